/*
 * lenia_oracle.h -- CPU oracle for the Flow Lenia step (TEST INFRASTRUCTURE ONLY).
 *
 * This is the checker, not the product.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product path (lenia_amd/, the HIP
 * library) never links, imports or calls anything under oracle/.
 *
 * It restates the reference step (lenia.cpp:435-1046, reintegration.cpp:13-1410,
 * transposition.cpp:6-995, lenia.h:30-84, render.h:291-341) over the reference's own
 * padded array layout (PADDING=4, face-only halo fills), single-threaded.  Every
 * floating-point operation is written out in a fixed order so that a second
 * implementation which follows the same order (the HIP kernels) produces identical
 * fp32 bits.  Two precisions are built from one source: lo_f32_* (float) and
 * lo_f64_* (double, used as the FP64 oracle for error statements).
 *
 * Parity status per function (pinning checks live in tests/test_oracle_*.py; each one
 * recomputes the function from the reference's definition in numpy, float64):
 *   fft1 / rfft3 / irfft3  pinned: numpy.fft at 1e-6 (f32) / 1e-12 (f64), brute-force DFT,
 *                          delta -> ones, Parseval, linearity (test_oracle_fft.py).
 *   kernel_field/spectrum  pinned: sum = 1, torus-even symmetry, real spectrum, K(0)=1,
 *                          closed-form ring profile at sampled radii.
 *   bell LUT               pinned: closed form exp(-((u-m)/s)^2/2) per entry.
 *   phase_growth           pinned (test_oracle_pins.py): U_k against a numpy circular
 *                          convolution of A_c0 with the closed-form kernel field at three
 *                          shapes (octant-mirror indexing, c0 routing, /N) and one cell by
 *                          direct summation; U_SUM per c1; H = sum h (LUT[(int)(255u)] - A
 *                          if asymptotic); B = clamp(A + dt H), soft clip through the
 *                          degree-7 Taylor fastexp; the analytic bell in physics mode.
 *   phase_blur             pinned: central differences of a numpy circular 5^3 Gaussian
 *                          blur of H and S at zero, torus and wall pads; the non-blur copy's
 *                          extent U_S[0, 5P/2) <- H over the padded layout.
 *   Sobel / central diff   pinned: linear fields (axis permutation, x32 scale), torus wrap,
 *                          numpy Sobel of S and H (face-only and full pads).
 *   alpha / F              pinned: F = NU (1 - a) - NA a with PREALPHA written at c*P+i and
 *                          read at c*N+i, checked against the closed form of a~_1, a~_2 at
 *                          32^3, C = 3, and without the quirk.
 *   phase_mu (mu, V)       pinned: both MU_VEL orders, damping, the +-ma clamp and the wall
 *                          clamp against numpy from F and V; mu from a linear S end to end.
 *   SRT                    pinned: per-cell closed form, reference and physics sn / clip,
 *                          channels < 3 under the quirk; the clamp equals the box overlap.
 *   RT torus / non-torus   pinned: brute-force box-overlap integral on the torus and at
 *                          zero / wall / periodic-image pads without the distance wrap;
 *                          identity at sigma = 1/2; mass conservation in physics mode; the
 *                          reference's x1.3 mass factor at sigma = 0.65.
 *   ADV                    pinned: integer translations, zero velocity.
 *   oscillators, TRM       pinned: Laplacian / flow / divergence closed forms; S = STA sum A,
 *                          GA = (1-GTA) GA + GTA A.
 *   MC case index          pinned: brute force + generated edge table vs the reference's
 *                          EDGE_TABLE fixture (tests/golden/edge_table.json).
 *   mesh (lo_mesh)         pinned: sphere / torus Euler characteristic 2 / 0, closed and
 *                          consistently oriented, outward normals, vertices on a linear
 *                          field's iso-plane, single-cell octahedron, per-field and
 *                          channel-shared edge maps (tests/test_mesh_oracle.py).
 *   full step              the composition of the pinned phases above, in the reference's
 *                          phase order (render.cpp:376-592), plus the f32-vs-f64 check.
 *   Not reproduced: FFTW's own float rounding (FFTW is absent; the FFT is this checker's
 *   canonical Stockham order, pinned to numpy above).
 */
#ifndef LENIA_ORACLE_H
#define LENIA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LO_MAX_K 16
#define LO_MAX_C 4      /* MAX_CHANNELS render.h:34 */
#define LO_MAX_BETA 12  /* MAX_BETA render.h:35 */

/* modes, render.h:125-233 */
enum { LO_MODE_SRT = 0, LO_MODE_RT = 1, LO_MODE_ADV = 2 };
enum { LO_BORDER_TORUS = 0, LO_BORDER_WALL = 1, LO_BORDER_NONE = 2 };
enum { LO_GRAD_CENTRAL = 0, LO_GRAD_SOBEL = 1 };
enum { LO_GROWTH_NONE = 0, LO_GROWTH_RT = 1, LO_GROWTH_TRANSPOSITION = 2 };

/* Reference-behaviour switches.  Bit set = reference-compatible behaviour. */
#define LO_Q_SOBEL_PERMUTE   (1u << 0)  /* Q1  lenia.h:30-64, reintegration.cpp:201-203 */
#define LO_Q_PREALPHA_INDEX  (1u << 1)  /* Q2  lenia.cpp:652,660,670 vs 696,708 */
#define LO_Q_SUM_RESET       (1u << 2)  /* Q4  lenia.cpp:912-915 */
#define LO_Q_FACE_HALO       (1u << 3)  /* Q5  transposition.cpp loop bounds */
#define LO_Q_TRANSPORT_TORUS (1u << 4)  /* Q6  render.cpp:538-548 */
#define LO_Q_RT_3CH          (1u << 5)  /* Q7  reintegration.cpp:642,664,791,866 */
#define LO_Q_GROWTH_LUT      (1u << 6)  /* Q8  lenia.cpp:93-101, lenia.h:79-84 */
#define LO_Q_RT_SNORM        (1u << 7)  /* Q9  lenia.cpp:730-733 */
#define LO_Q_CLAMP01         (1u << 8)  /* Q10 lenia.cpp:978-992 */
#define LO_Q_CD_COPY         (1u << 9)  /* Q12 reintegration.cpp:179-186 */
#define LO_Q_SOFTCLIP_TR     (1u << 10) /* Q13 lenia.cpp:935-957 */
#define LO_Q_ALL             0x7FFu

typedef struct {
    int   mode, border, sobel_border, gradient, growth_type;
    int   mu_vel, adv_vel, blur;
    float dt, sigma, dd, s_norm, theta_a, gta, sta, v_damp;
    int   asymptotic[LO_MAX_C], soft_clip[LO_MAX_C];
    int   osc_la, osc_flow, osc_div_flow, osc_vel;
    float osc_la_amt, osc_flow_amt, osc_div_flow_amt;
    uint32_t quirks;
} lo_params;

/* Kernel parameters, render.h:347-361 (spectra are derived, not stored here). */
typedef struct {
    int   c0, c1, radius, beta_size;
    float b[LO_MAX_BETA];
    float r_a, m, s, h;
} lo_kernel;

/* Fields that can be read or written (unpadded, [c][z][y][x]). */
enum {
    LO_F_A = 0,     /* GRID interior  [C][N] */
    LO_F_GA = 1,    /* GA             [C][N] */
    LO_F_S = 2,     /* SUM interior   [N]    */
    LO_F_V = 3,     /* V_x,V_y,V_z    [C][3][N] */
    LO_F_H = 4,     /* growth H       [C][N] */
    LO_F_B = 5,     /* BACK           [C][N] */
    LO_F_MU = 6,    /* MU_x,y,z       [C][3][N] */
    LO_F_USUM = 7,  /* U_SUM          [C][N] */
    LO_F_F = 8,     /* F_x,y,z        [C][3][N] */
    LO_F_NA = 9,    /* NA_x,y,z       [3][N] */
    LO_F_NU = 10,   /* NU_x,y,z       [C][3][N] */
    LO_F_TT = 11,   /* TT after reintegration (before TRM) [C][N] */
    LO_F_U = 12,    /* U_k = Re IDFT / N per kernel [K][N] */
};

#define LO_DECLARE(P, T)                                                                  \
    typedef struct P##world P##world;                                                     \
    P##world *P##create(int W, int H, int D, int C, int K, const lo_params *p,           \
                        const lo_kernel *k);                                              \
    void P##destroy(P##world *w);                                                         \
    int  P##set_field(P##world *w, int field, const T *src);                              \
    int  P##get_field(P##world *w, int field, T *dst);                                    \
    void P##step(P##world *w, int nsteps);                                                \
    void P##step_phases(P##world *w, int nphases);                                        \
    double P##mass(P##world *w);                                                          \
    void P##fft1(int n, T *re_im, int inverse);                                           \
    void P##rfft3(int W, int H, int D, const T *in, T *out);                              \
    void P##irfft3(int W, int H, int D, const T *in, T *out);                             \
    void P##kernel_field(int W, int H, int D, const lo_kernel *k, T *out);               \
    void P##kernel_spectrum(int W, int H, int D, const T *field, T *octant);              \
    void P##blur_field(int W, int H, int D, T *out);                                      \
    void P##bell_table(float m, float s, T *out256);                                      \
    void P##mc_case(int W, int H, int D, const T *field, T iso, int face_only,            \
                    uint8_t *out);

LO_DECLARE(lo_f32_, float)
LO_DECLARE(lo_f64_, double)
double lo_canon_exp(double x);
/* Host threads for the step's parallel loops (default 1); results do not depend on it. */
int lo_set_threads(int n);

/* Edge mask of a marching-cubes case, generated from the cube's edge list. */
int lo_edge_mask(int case_index);

/* One mesh vertex, the 28-byte Vertex of render.h:443-501: position, accumulated
 * (unnormalised) normal, packed colour. */
typedef struct lo_vertex {
    float x, y, z, nx, ny, nz;
    uint32_t color;
} lo_vertex;

/* Marching-cubes mesh of F fields (f0 = A, f1 = U_SUM, f2 = H; each C channels of
 * W*H*D) after generate_mesh_v1/v2/v3 (lenia_mesh_oracle.c).  tri is a 256x16
 * triangulation table (-1 terminated rows), colors[C].  Returns 0, or -1 when the
 * counts exceed max_verts / max_indices (counts still written), -2 out of memory. */
int lo_mesh(int W, int H, int D, int C, int F, const float *f0, const float *f1, const float *f2, float iso,
            float cell, int face_only, const int *tri, const uint32_t *colors, lo_vertex *verts, int64_t max_verts,
            uint32_t *indices, int64_t max_indices, int64_t *n_verts, int64_t *n_indices);

#ifdef __cplusplus
}
#endif
#endif
