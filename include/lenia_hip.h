/*
 * lenia_hip.h -- C-ABI of the MI355X Flow Lenia step engine (liblenia_hip.so).
 *
 * One engine owns a batch of independent worlds (any mix of supported shapes, each
 * with its own channels, kernels and parameters) resident in HBM, and advances all of
 * them by whole simulation steps on one HIP stream.  Plain pointers, sizes and POD
 * structs only; every call returns an lh_status (0 = ok) and never throws.
 *
 * The reference has no C ABI: its step lives behind C++ entry points that operate on
 * module globals, all returning void (lenia.h:5-28, reintegration.h:74-107,
 * transposition.h:5-12).  Those entry points are re-provided over this ABI by the
 * shim in lenia_ref_shim.h; the mapping of each reference phase to the calls below is
 * given there.  Calls on one engine must be serialised by the caller.
 */
#ifndef LENIA_HIP_H
#define LENIA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LH_MAX_K 16    /* per world; the reference caps at MAX_KERNELS=12 (render.h:33) */
#define LH_MAX_C 4     /* MAX_CHANNELS render.h:34 */
#define LH_MAX_BETA 12 /* MAX_BETA render.h:35 */

typedef enum {
    LH_OK = 0,
    LH_EINVAL = -1,       /* bad argument, shape or parameter */
    LH_EUNSUPPORTED = -2, /* valid request the engine does not implement */
    LH_EHIP = -3,         /* HIP runtime error (message in lh_last_error) */
    LH_ESTATE = -4,       /* call out of order (e.g. step before commit) */
    LH_ENOMEM = -5
} lh_status;

/* ModeType / Border / GradientType / GrowthType, render.h:125-233 */
enum { LH_MODE_SRT = 0, LH_MODE_RT = 1, LH_MODE_ADV = 2 };
enum { LH_BORDER_TORUS = 0, LH_BORDER_WALL = 1, LH_BORDER_NONE = 2 };
enum { LH_GRAD_CENTRAL = 0, LH_GRAD_SOBEL = 1 };
enum { LH_GROWTH_NONE = 0, LH_GROWTH_RT = 1, LH_GROWTH_TRANSPOSITION = 2 };

/* Reference-behaviour switches; bit set = reproduce the reference, clear = physics
 * semantics (mass-conserving transport, true gradient axes, analytic growth ...). */
#define LH_Q_SOBEL_PERMUTE   (1u << 0)  /* Sobel axes (z,x,y): lenia.h:30-64 */
#define LH_Q_PREALPHA_INDEX  (1u << 1)  /* alpha written at c*P+i, read at c*N+i: lenia.cpp:660,708 */
#define LH_Q_SUM_RESET       (1u << 2)  /* S zeroed each step: lenia.cpp:912-915 */
#define LH_Q_FACE_HALO       (1u << 3)  /* halos fill faces only: transposition.cpp */
#define LH_Q_TRANSPORT_TORUS (1u << 4)  /* transport halos always periodic: render.cpp:538-548 */
#define LH_Q_RT_3CH          (1u << 5)  /* SRT/RT move channels 0..2 only: reintegration.cpp:642 */
#define LH_Q_GROWTH_LUT      (1u << 6)  /* 8-bit bell LUT: lenia.cpp:93-101, lenia.h:79-84 */
#define LH_Q_RT_SNORM        (1u << 7)  /* sn=(1-S)/(4s^2)+S/(2s^3): lenia.cpp:730-733 */
#define LH_Q_CLAMP01         (1u << 8)  /* clamp to [0,1] at the end of the step: lenia.cpp:978-992 */
#define LH_Q_CD_COPY         (1u << 9)  /* non-blur central-diff copy extent: reintegration.cpp:179-186 */
#define LH_Q_SOFTCLIP_TR     (1u << 10) /* transposition soft-clip writes BACK: lenia.cpp:935-957 */
#define LH_Q_ALL             0x7FFu

/* Simulation parameters of one world; the reference keeps these as globals
 * (main.cpp:30-56, off_table render.h:374-386, ch[] render.h:486-492). */
typedef struct {
    int   mode, border, sobel_border, gradient, growth_type;
    int   mu_vel, adv_vel, blur;
    float dt, sigma, dd, s_norm, theta_a, gta, sta, v_damp;
    int   asymptotic[LH_MAX_C], soft_clip[LH_MAX_C];
    int   osc_la, osc_flow, osc_div_flow, osc_vel;
    float osc_la_amt, osc_flow_amt, osc_div_flow_amt;
    uint32_t quirks;
} lh_params;

/* One growth kernel: the parameter part of struct Kernel (render.h:347-361).  The
 * engine derives the kernel field and spectrum (init_kernel lenia.cpp:199-261). */
typedef struct {
    int   c0, c1, radius, beta_size;
    float b[LH_MAX_BETA];
    float r_a, m, s, h;
} lh_kernel;

/* World extents: x = W, y = H, z = D (the reference grid is W x H x W). */
typedef struct { int W, H, D, C, K; } lh_shape;

/* Per-world fields, unpadded, channel-major, x fastest. */
enum {
    LH_F_A = 0,    /* GRID interior          [C][N]    (state) */
    LH_F_GA = 1,   /* GA moving average      [C][N]    (state) */
    LH_F_S = 2,    /* SUM interior           [N]       (state) */
    LH_F_V = 3,    /* V_x,V_y,V_z            [C][3][N] (state) */
    LH_F_H = 4,    /* growth H of last step  [C][N] */
    LH_F_B = 5,    /* BACK of last step      [C][N] */
    LH_F_MU = 6,   /* MU (RT/SRT) or clamped displacement (ADV) [C][3][N]
                      (for SRT worlds without oscillators needs LH_OPT_KEEP_MU) */
    LH_F_USUM = 7  /* U_SUM of last step     [C][N] (needs LH_OPT_KEEP_USUM) */
};

/* Options (lh_set_option). */
enum {
    LH_OPT_KEEP_VELOCITY = 0, /* maintain V even where the mode never reads it (default 0) */
    LH_OPT_KEEP_USUM = 1,     /* write U_SUM for mesh views (default 0) */
    LH_OPT_USE_GRAPH = 2,     /* replay steps from a captured hipGraph (default 1) */
    LH_OPT_PROFILE = 3,       /* time every kernel launch with HIP events (default 0) */
    LH_OPT_AUTOTUNE = 4,      /* lh_commit times the candidate schedules of each shape class
                                 (sweep segment lengths, sweep vs tiled transport) on the
                                 class's worlds and keeps the fastest (default 1) */
    LH_OPT_SCHEDULE = 5,      /* forces one schedule on every shape class instead (default -1:
                                 none): zseg_local | zseg_transport << 8 | tiled << 16 |
                                 growth << 17, the z-planes per sweep workgroup of the gradient
                                 and transport sweeps (0: the launchers' default, above the
                                 depth: the whole depth), 1 to move the sweep's worlds by the
                                 tiled kernel, and the spectral growth as two launches (1) or
                                 one fused launch where the shape's half spectrum fits in LDS
                                 (2; 0: the default choice).  Every schedule gives
                                 bit-identical results; for tests. */
    LH_OPT_KEEP_MU = 6        /* write MU for every world (default 0: SRT worlds without
                                 oscillators are moved inside the gradient sweep from each
                                 cell's own mu, which then never reaches HBM) */
};

/* Growth convolution of a world (lh_set_world_conv).  LH_CONV_FFT is the fp32 FFT
 * path that matches the reference's FFTW convolution (lenia.cpp:435-538); the others
 * evaluate the same convolution as dense DFT matrix products on the matrix cores with
 * 16-bit operands and fp32 accumulation, split into hi + lo parts (three products) or
 * unsplit (one product).  No reference counterpart: a precision/throughput option. */
enum {
    LH_CONV_FFT = 0,
    LH_CONV_GEMM_BF16X3 = 1,
    LH_CONV_GEMM_FP16X3 = 2,
    LH_CONV_GEMM_BF16 = 3,
    LH_CONV_GEMM_FP16 = 4
};

/* ---- engine lifetime --------------------------------------------------------- */
lh_status lh_create(int device, void **engine);
lh_status lh_destroy(void *engine);
lh_status lh_set_option(void *engine, int option, int value);

/* Adds a world; returns its id.  Must precede lh_commit. */
lh_status lh_add_world(void *engine, const lh_shape *shape, const lh_params *params,
                       const lh_kernel *kernels, int *world_id);
/* Selects the growth convolution of a world added before lh_commit (default
 * LH_CONV_FFT).  The matrix-core methods need a cube of side 32, 64 or 128 and no
 * blur pass (LH_EUNSUPPORTED otherwise).  No reference counterpart. */
lh_status lh_set_world_conv(void *engine, int world, int method);
/* Allocates HBM for all worlds, builds kernel spectra and growth tables on the GPU. */
lh_status lh_commit(void *engine);

/* ---- state --------------------------------------------------------------------- */
/* count is in floats and must equal the field's size. Host pointers; synchronous. */
lh_status lh_set_field(void *engine, int world, int field, const float *src, size_t count);
lh_status lh_get_field(void *engine, int world, int field, float *dst, size_t count);
/* Device address of a field (valid until lh_destroy); for zero-copy consumers. */
lh_status lh_field_device_ptr(void *engine, int world, int field, float **dptr, size_t *count);
/* Re-derives the spectra after fields were written through lh_field_device_ptr. */
lh_status lh_state_updated(void *engine);

/* ---- stepping ------------------------------------------------------------------ */
/* Enqueues nsteps full steps of every world on the engine's stream (asynchronous).
 * With graphs on (the default), up to 256 steps of every shape class are one launch of
 * a captured hipGraph (captured on the first call with that step count and kept; calls
 * of more steps launch 256-step graphs and one for the remainder).  Replaces the
 * reference's per-step phase hand-off between its simulation threads
 * (render.cpp:357-604, main.cpp:173-247). */
lh_status lh_step(void *engine, int nsteps);
/* Captures the graphs lh_step(engine, nsteps) will launch (1 <= nsteps <= 256, both A
 * buffer parities) without running them, so a later lh_step of that count starts at
 * once.  No reference counterpart (graph capture is set-up work). */
lh_status lh_prepare_steps(void *engine, int nsteps);
lh_status lh_synchronize(void *engine);
/* The engine's hipStream_t. */
void *lh_stream(void *engine);

/* ---- reductions and read-backs ------------------------------------------------- */
/* Total mass sum_c sum_i A, accumulated in FP64 per world, into out[n_worlds]. */
lh_status lh_world_mass(void *engine, double *out, int n_worlds);
/* Mass and circular first moments per world (FP64), out[7 * n_worlds]: sum A, then per
 * axis x, y, z the sums of A cos(t) and A sin(t), t = 2 pi (i + 1/2) / extent, A summed
 * over channels -- the centre of mass on the torus is atan2 of each pair.  No reference
 * counterpart: labels for the learned parameter proposals. */
lh_status lh_world_moments(void *engine, double *out, int n_worlds);
/* Marching-cubes case index (mesh.cpp:173-185) of one channel of field LH_F_A,
 * LH_F_H or LH_F_USUM with the torus face pads of transpose_mesh_borders
 * (transposition.cpp:881-995): out[N] bytes. */
lh_status lh_mc_case(void *engine, int world, int field, int channel, float iso,
                     int face_only, uint8_t *out);

/* lh_mc_case with no allocation: the same result, staged in a static device scratch array
 * of 2 MiB per GPU (held until the process ends; worlds of at most 2,097,152 cells, i.e.
 * up to 128^3; larger ones give LH_EUNSUPPORTED) instead of memory the engine allocates
 * on first use.  It calls no hipMalloc / hipFree.  Every engine on one GPU shares that
 * array, so calls that overlap on one GPU (from several host threads) race.  No
 * counterpart in the reference: for callers that cannot allocate device memory (tight
 * serving loops, code that must not call hipMalloc). */
lh_status lh_mc_case_noalloc(void *engine, int world, int field, int channel, float iso,
                             int face_only, uint8_t *out);

/* ---- meshes ------------------------------------------------------------------------ */
/* One mesh vertex: the reference's 28-byte Vertex (render.h:443-501) as the mesh
 * threads fill it -- position, unnormalised accumulated normal, packed colour. */
typedef struct lh_vertex {
    float x, y, z, nx, ny, nz;
    uint32_t color;
} lh_vertex;

/* This library's default triangulation table, tri[256 * 16], -1 terminated rows of
 * edge triples (the layout of TRIANGLE_TABLE, marching_cubes.h:98-355), derived from
 * the cube's topology.  Host only. */
lh_status lh_mc_table(int *tri_table);

/* Marching-cubes mesh of one world, replacing the mesh threads' generate_mesh_v1 /
 * v2 / v3 (mesh.cpp:34, 434, 1115) for view 1 (A), 2 (A, U_SUM) or 3 (A, U_SUM, H),
 * with one section over the whole grid.  Corner values are field > iso with the torus
 * face pads of transpose_mesh_borders (zero edge/corner pads when face_only); corner
 * positions are (g + delta) * cell_size.  tri_table: the caller's 256x16 table (the
 * reference passes TRIANGLE_TABLE) or NULL for lh_mc_table's; colors[C]: the packed
 * colour of each channel (NULL: white).  Vertices are numbered in the order the
 * reference's single mesh thread creates them, indices follow its sweep, and normals
 * are summed in its order, so the buffers match it bit for bit.  Writes the counts to
 * *n_verts / *n_indices; fills verts / indices only when both fit max_verts /
 * max_indices (pass NULL and 0 to query sizes).  View 2 and 3 need LH_OPT_KEEP_USUM. */
lh_status lh_mesh(void *engine, int world, int view, float iso, float cell_size, int face_only,
                  const int *tri_table, const uint32_t *colors, lh_vertex *verts, int64_t max_verts,
                  uint32_t *indices, int64_t max_indices, int64_t *n_verts, int64_t *n_indices);

/* The same mesh built from the fields held in snapshot slot `slot` (lh_snapshot) --
 * the mesh threads' RBB / U_SUM_RB / H_RB copies (main.cpp:212-244) -- on the
 * read-back stream, so the steps enqueued after the snapshot keep running meanwhile.
 * The snapshot must hold the view's fields (A; U_SUM; H).  Returns when the mesh is in
 * the host buffers.  Calls of lh_mesh and lh_mesh_snapshot on one engine share its mesh
 * buffers: make them from one host thread at a time. */
lh_status lh_mesh_snapshot(void *engine, int slot, int world, int view, float iso, float cell_size,
                           int face_only, const int *tri_table, const uint32_t *colors, lh_vertex *verts,
                           int64_t max_verts, uint32_t *indices, int64_t max_indices, int64_t *n_verts,
                           int64_t *n_indices);

/* ---- checkpoints ----------------------------------------------------------------- */
/* One world's complete state as a flat, versioned byte image: the intent of the
 * reference's disabled save_config / load_config (config.cpp:75-261: the simulation
 * globals, then per cell grid, GA, SUM and V, then channel flags and kernels), for one
 * world of a batch.  Little-endian, in this order:
 *   lh_checkpoint_header (64 bytes), lh_params, lh_kernel[K],
 *   float32 A [C][N], GA [C][N], S [N], V [C][3][N]   (N = W*H*D, x fastest).
 * The spectra, growth tables and per-step outputs (H, B, MU, U_SUM) are derived from
 * these and are not stored. */
#define LH_CHECKPOINT_VERSION 1u
typedef struct {
    char     magic[4];      /* "LHCK" */
    uint32_t version;       /* LH_CHECKPOINT_VERSION */
    uint32_t header_bytes;  /* sizeof(lh_checkpoint_header) = 64 */
    uint32_t params_bytes;  /* sizeof(lh_params) */
    uint32_t kernel_bytes;  /* sizeof(lh_kernel) */
    int32_t  conv;          /* growth convolution (LH_CONV_*) */
    lh_shape shape;
    uint32_t reserved;      /* 0 */
    uint64_t steps;         /* steps the world has taken (lh_world_steps) */
    uint64_t payload_bytes; /* bytes of the float32 fields that follow the kernels */
} lh_checkpoint_header;

/* Size in bytes of world's checkpoint image. */
lh_status lh_checkpoint_size(void *engine, int world, size_t *bytes);
/* Writes world's image into buf (bytes >= lh_checkpoint_size); synchronous, after the
 * steps enqueued so far. */
lh_status lh_checkpoint_save(void *engine, int world, void *buf, size_t bytes);
/* Restores world's state (A, GA, S, V, step counter) from an image.  The world must have
 * the image's shape, parameters, kernels and convolution, bit for bit (LH_EINVAL
 * otherwise, as for a wrong magic, version, size or truncated image); the spectra are
 * re-derived before the next step, so stepping continues exactly as it would have
 * from the saved state. */
lh_status lh_checkpoint_load(void *engine, int world, const void *buf, size_t bytes);
/* Steps world has taken since it was added (or since the step count of the image it
 * was last loaded from). */
lh_status lh_world_steps(void *engine, int world, uint64_t *steps);

/* ---- random worlds (generate_random lenia.cpp:370-429, seed_grid_gaussian
 * lenia.cpp:305-334) -------------------------------------------------------------- */
/* Samples parameter sets first .. first+n-1 on the GPU with generate_random's ranges:
 * neighbourhood 1..8 (every kernel's radius), 1..10 kernels, 1..2 rings with peaks
 * k/100, m = (100 + k)/1000 (k < 500), s = m/10 + k/1000 (k < 200), h = 1, r_a = 1,
 * c0 = c1 = 0 for kernel 0 or one channel, else uniform over the channels.  Draws are
 * splitmix64 of (seed, set, slot) in place of rand(), so a set is reproducible from its
 * index.  kernels[n][LH_MAX_K] (entries past the count zeroed), kernel_count[n]; host
 * buffers, synchronous. */
lh_status lh_sample_random(int device, uint64_t seed, int64_t first, int n, int channels,
                           lh_kernel *kernels, int *kernel_count);
/* Seeds worlds first_world .. first_world+n_worlds-1 on the GPU: world j gets the five
 * Gaussian blobs of set first_set + j (centres drawn like the kernels, width its
 * neighbourhood, every channel, plain distance) in A, and zero GA, S, V and step count,
 * as generate_random does after sampling.  Synchronous. */
lh_status lh_seed_random(void *engine, int first_world, int n_worlds, uint64_t seed, int64_t first_set);

/* ---- asynchronous read-backs ------------------------------------------------------ */
/* The render side of the reference copies GRID -> RBB (and reads U_SUM_RB / H_RB) on its
 * main thread while the simulation threads may be transporting (check_and_dispatch_threads
 * main.cpp:212-244; the phase protocol render.cpp:357-604).  Here a snapshot is a device
 * copy of selected fields, enqueued on the engine stream after every step enqueued so far
 * and before any later one, so it is always one whole step's state; two slots let a
 * consumer read one while the next is taken.  No reference counterpart as an API. */
enum { LH_SNAP_A = 1, LH_SNAP_H = 2, LH_SNAP_USUM = 4 };
/* Enqueues a snapshot of fields (mask of LH_SNAP_*) of worlds first_world ..
 * first_world+n_worlds-1 into slot 0 or 1 and returns (U_SUM needs LH_OPT_KEEP_USUM). */
lh_status lh_snapshot(void *engine, int slot, int first_world, int n_worlds, int fields);
/* ready = 1 once the slot's copy has completed, 0 while it waits behind earlier steps. */
lh_status lh_snapshot_ready(void *engine, int slot, int *ready);
/* Copies one field (LH_F_A, LH_F_H or LH_F_USUM) of one world of the slot to host memory;
 * waits for that snapshot only, not for steps enqueued after it. */
lh_status lh_snapshot_read(void *engine, int slot, int world, int field, float *dst, size_t count);

/* ---- introspection -------------------------------------------------------------- */
int lh_world_count(void *engine);
/* Number of distinct kernel names with recorded launches (LH_OPT_PROFILE). */
int lh_profile_count(void *engine);
/* i-th kernel: name, launches and total milliseconds since the last reset. */
lh_status lh_profile_entry(void *engine, int i, const char **name, long *launches, double *ms);
lh_status lh_profile_reset(void *engine);
/* Measurement aid, no counterpart in the reference: one plain device-to-device copy
 * kernel moving `bytes` bytes (rounded down to 16) on the engine stream, synchronous;
 * its FETCH_SIZE / WRITE_SIZE counters calibrate HBM-traffic readings. */
lh_status lh_calibration_copy(void *engine, size_t bytes);
/* Whether a shape is supported: 1 built into this library, 2 as a shape module
 * (liblenia_shape_<W>x<H>x<D>.so in $LENIA_SHAPE_DIR or in shapes/ beside the library,
 * built by `python -m lenia_amd.build --shape WxHxD` for any sides divisible by 4 from 4
 * to 128, as the reference allows any such GRID_WIDTH x GRID_HEIGHT, render.h:27-29,
 * README.md:30), 0 neither.  Host only: no HIP call, no module loaded. */
int lh_shape_supported(int W, int H, int D);
const char *lh_last_error(void);
const char *lh_version(void);

#ifdef __cplusplus
}
#endif
#endif
