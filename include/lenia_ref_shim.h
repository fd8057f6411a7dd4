/*
 * lenia_ref_shim.h -- the reference's simulation-step entry points, re-provided over
 * the liblenia_hip C-ABI (liblenia_ref_shim.so).
 *
 * The reference declares its step as C++ functions on module globals (no extern "C",
 * all void): lenia.h:5-28, reintegration.h:74-107, transposition.h:5-12.  This header
 * declares the same functions with the same signatures and C++ linkage, so the
 * reference's caller (sim_thread_loop render.cpp:357-604, Setup render.cpp:813-903,
 * EnterMsgLoop main.cpp:301-308, the UI actions config.cpp:398-402) links against the
 * shim instead of lenia.cpp / reintegration.cpp / transposition.cpp.
 *
 * The GPU engine performs a whole step per launch sequence, so the phases map as:
 *   fft_step (thread id 0)        enqueue the full step on the GPU (after re-uploading
 *                                 parameters or host state the caller changed)
 *   mt_fft_step ... reintegration_step and the border / srt / rt / adv / mu functions
 *                                 no GPU work of their own; they mark the thread done
 *   transpose_to_row_major (id 0) wait for the step; copy A, H, U_SUM back into GRID,
 *                                 H, U_SUM (padded layout)
 *   transpose_mesh_borders        fill the torus face pads of RBB, H_RB, U_SUM_RB
 * Thread 1 of every phase returns at once (the reference splits each phase by z-slab
 * between SIM_THREADS=2 threads).  Host arrays other than GRID, H, U_SUM and the mesh
 * read-backs (RBB, H_RB, U_SUM_RB) are not mirrored.
 *
 * Linux stand-ins replace the Win32/DX9 types the reference's render.h brings in; with
 * LENIA_SHIM_REFERENCE_HEADERS defined the shim is compiled against the reference's own
 * render.h instead (it only touches TD::id, TD::is_done and the Kernel parameters).
 */
#ifndef LENIA_REF_SHIM_H
#define LENIA_REF_SHIM_H

#ifdef LENIA_SHIM_REFERENCE_HEADERS
#include "render.h"
#else

#ifndef GRID_WIDTH
#define GRID_WIDTH 32           /* render.h:27 */
#endif
#ifndef GRID_HEIGHT
#define GRID_HEIGHT 32          /* render.h:28 */
#endif
#define SIM_THREADS 2           /* render.h:14 */
#define GRID_SIZE (GRID_WIDTH * GRID_HEIGHT * GRID_WIDTH)
#define MAX_KERNELS 16          /* render.h:33 has 12; the engine takes 16 */
#define MAX_CHANNELS 4          /* render.h:34 */
#define MAX_BETA 12             /* render.h:35 */
#define PADDING 4               /* render.h:291 */
#define PW (GRID_WIDTH + PADDING * 2)
#define PH (GRID_HEIGHT + PADDING * 2)
#define PZOFFSET (PW * PH)
#define GRID_PADDED (PW * PH * (GRID_WIDTH + PADDING * 2))
#define IDX(x, y, z) ((x) + (y) * GRID_WIDTH + (z) * GRID_WIDTH * GRID_HEIGHT)
#define PIDX(x, y, z) (((x) + PADDING) + ((y) + PADDING) * PW + ((z) + PADDING) * PZOFFSET)
#define CPOFFSET(c) ((c) * GRID_PADDED)

typedef float fftwf_complex[2];

typedef enum { ALPHA_GRID, ALPHA_GRID_GROWTH, ALPHA_GA, ALPHA_GA_DIFF, ALPHA_MAX } AlphaMode;
typedef enum { GROWTH_NOGROWTH, GROWTH_RT, GROWTH_TRANSPOSITION, GROWTH_MAX } GrowthType;
typedef enum { MODE_SRT, MODE_RT, MODE_ADV, MODE_MAX } ModeType;
typedef enum { BORDER_TORUS, BORDER_WALL, BORDER_NONE, BORDER_MAX } Border;
typedef enum { GRADIENT_CENTRAL_DIFF, GRADIENT_SOBEL, GRADIENT_MAX } GradientType;

/* struct Kernel render.h:347-361 (the spectra stay on the GPU; grid/kg are not filled) */
typedef struct Kernel {
    fftwf_complex grid[GRID_SIZE];
    fftwf_complex kg[GRID_SIZE];
    float real_k[GRID_SIZE];
    int beta_size;
    float b[MAX_BETA];
    float r_a;
    float m;
    float s;
    float h;
    int radius;
    int c0;
    int c1;
} Kernel;

/* struct OffsetTable render.h:374-386 (offsets are not used by the engine) */
typedef struct OffsetTable {
    float sigma;
    float dd;
    int row_size;
    float ma;
    int *X, *Y, *Z;
} OffsetTable;

/* struct ChannelData render.h:486-492 without the D3DXCOLOR */
typedef struct ChannelData {
    bool asymptotic;
    bool soft_clip;
} ChannelData;

/* struct TD render.h:498-527: the step only reads id and writes is_done */
typedef struct TD {
    bool is_done;
    int id;
} TD;

struct Button; /* render.h:560-577; init_offset_table ignores it */

/* Parameters the reference's main.cpp:30-66 defines and the step reads. */
extern float timestep, GTA, STA, V_DAMP, N, THETA_A, S_NORM, ISO_VALUE;
extern float OSC_LA_AMT, OSC_FLOW_AMT, OSC_DIV_FLOW_AMT, neighborhood;
extern int channels, kernel_count;
extern bool ADV_VEL, MU_VEL, BLUR_CENTRAL_DIFF, OSC_LA, OSC_FLOW, OSC_DIV_FLOW, OSC_VEL;
extern Border SOBEL_BORDER, BORDER;
extern AlphaMode ALPHA_MODE;
extern GrowthType GROWTH_TYPE;
extern ModeType MODE_TYPE;
extern GradientType GRADIENT_MODE;
extern Kernel kernels[MAX_KERNELS];
extern Kernel blur_kernel;
extern OffsetTable off_table;
extern ChannelData ch[MAX_CHANNELS];

/* State arrays the shim defines (lenia.cpp:9-89): the mirrored ones. */
extern float GRID[GRID_PADDED * MAX_CHANNELS];
extern float H[GRID_PADDED * MAX_CHANNELS];
extern float U_SUM[GRID_PADDED * MAX_CHANNELS];
extern float RBB[GRID_PADDED * MAX_CHANNELS];
extern float H_RB[GRID_PADDED * MAX_CHANNELS];
extern float U_SUM_RB[GRID_PADDED * MAX_CHANNELS];
extern float GA[GRID_SIZE * MAX_CHANNELS];
extern float SUM[GRID_PADDED];
extern float V_x[GRID_PADDED * MAX_CHANNELS], V_y[GRID_PADDED * MAX_CHANNELS], V_z[GRID_PADDED * MAX_CHANNELS];
extern float POS_ARRAY_X[GRID_SIZE], POS_ARRAY_Y[GRID_SIZE], POS_ARRAY_Z[GRID_SIZE];
#endif /* LENIA_SHIM_REFERENCE_HEADERS */

/* ---- lenia.h:5-28 ---------------------------------------------------------------- */
void init_conway();                       /* lenia.h:5   uniform noise in GRID (lenia.cpp:285-302) */
void fft_step(TD *t);                     /* lenia.h:6   id 0: enqueue the whole GPU step */
void mt_fft_step(TD *t);                  /* lenia.h:7   fused into the step */
void init_kernel(Kernel *kernel);         /* lenia.h:8   marks kernels dirty; spectrum built on GPU */
void init_fftw_plans();                   /* lenia.h:9   creates the engine */
void generate_random();                   /* lenia.h:10  randomiser ranges lenia.cpp:370-429 */
void init_pos_array();                    /* lenia.h:11  fills POS_ARRAY_* */
void init_offset_table(Button *b);        /* lenia.h:12  ma = dd - sigma */
void sobel_convolve(TD *t);               /* lenia.h:13  fused */
void sobel_convolve_h(TD *t);             /* lenia.h:14  fused */
void transpose_to_row_major(TD *t);       /* lenia.h:16  id 0: wait, copy A/H/U_SUM back */
void compute_f(TD *t);                    /* lenia.h:17  fused */
void reintegration_step(TD *t);           /* lenia.h:21  fused */
void transpose_sobel_borders(TD *t);      /* lenia.h:22  halos are resolved on the GPU */
void transpose_borders(TD *t);            /* lenia.h:23  halos are resolved on the GPU */
void transpose_mesh_borders(TD *t);       /* lenia.h:24  fills RBB / H_RB / U_SUM_RB face pads */
void init_bell_lut();                     /* lenia.h:28  tables are built per kernel */
void init_gaussian_kernel();              /* lenia.h:25  unused 3^3 kernel (lenia.cpp:105-124) */

/* ---- reintegration.h:74-107 -------------------------------------------------------- */
void precompute_blur_kernel(Kernel *kernel);                                   /* :75 */
void h_sum_blur(TD *t);                                                        /* :82 */
void compute_velocity_mu_wall(TD *t);                                          /* :97 */
void compute_velocity_mu(TD *t);                                               /* :98 */
void compute_f_mu_wall(TD *t);                                                 /* :99 */
void compute_f_mu(TD *t);                                                      /* :100 */
void srt(float *ARR, int tid, int section, int end, float clip, float sn);            /* :103 */
void rt_border_torus(float *ARR, int tid, int section, int end, float clip, float sn); /* :104 */
void rt_border_other(float *ARR, int tid, int section, int end, float clip, float sn); /* :105 */
void adv_vel(float *ARR, int tid, int section, int end, float clip, float sn);         /* :106 */
void adv_f(float *ARR, int tid, int section, int end, float clip, float sn);           /* :107 */

/* ---- transposition.h:5-12 ------------------------------------------------------------ */
void transpose_central_diff_borders(TD *t);  /* :5 */
void wall_central_diff_borders(TD *t);       /* :6 */
void zero_central_diff_borders(TD *t);       /* :7 */
void wall_sobel_borders(TD *t);              /* :9 */
void zero_sobel_borders(TD *t);              /* :10 */

/* ---- shim control (no counterpart in the reference) ---------------------------------- */
/* Device the engine runs on (default 0); call before init_fftw_plans. */
void lenia_shim_set_device(int device);
/* Marks host GRID / GA / SUM / V as changed by the caller: uploaded at the next fft_step. */
void lenia_shim_state_changed();
/* Releases the engine (the next fft_step creates it again). */
void lenia_shim_shutdown();
/* Last engine error ("" if none). */
const char *lenia_shim_last_error();

#endif
