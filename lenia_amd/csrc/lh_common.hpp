// lh_common.hpp -- data shared by the host engine and the HIP kernels.
//
// Every world lives in one HBM pool; a DevWorld holds its parameters and the float
// offsets of its fields.  Fields are unpadded, channel-major, x fastest.  Spectra are
// half spectra [z][y][kx] of float2 with kx in [0, W/2] (W/2+1 bins).
#pragma once
#include <stdint.h>

#define LHD_MAX_K 16
#define LHD_MAX_C 4

struct DevKernel {
    int c0, c1;
    float h, m, s;
};

struct DevWorld {
    int W, H, D, C, K;
    int mode, border, sobel_border, gradient, growth_type;
    int mu_vel, adv_vel, blur;
    int keep_v, keep_usum, keep_mu, v_live, need_f, need_alpha;
    int asym_mask, soft_mask;
    int osc_la, osc_flow, osc_div_flow, osc_vel;
    uint32_t quirks;
    int dd;            // transport halo depth (int)off_table.dd
    float dt, sigma, ma, sn, clip, si, theta_a, gtam, gtas, stam, stas, damp;
    float inv_theta;   // 1 / theta_a when theta_a is a power of two (x / theta == x * inv_theta exactly), else 0
    float osc_la_amt, osc_flow_amt, osc_div_flow_amt;
    // float offsets into the pool
    // A / A2 and S / S2 alternate with the step parity (read one, write the other)
    int64_t A, A2, GA, S, S2, V, Bk, Hg, USUM, MU, Fv, Fsp, Gsp, KH, LUT, US, SUMS, BH;
    int64_t AL;        // ALPHA = |A - GA| for the divergence oscillator [C][N]
    int64_t KST, KSP;  // staging (second pool, commit only): kernel fields and spectra
    int nkf;           // staged fields: K kernels (+1 blur kernel)
    DevKernel kern[LHD_MAX_K];
};

// Worlds whose transport runs as the z-sweep (RT, dd = 1, no oscillators); the tiled
// transport kernel takes the rest.
__host__ __device__ inline bool sweep_world(const DevWorld &w) {
    return w.mode == 1 && w.dd == 1 && !(w.osc_la || w.osc_flow || w.osc_div_flow);
}

// Worlds moved by SRT without oscillators: each cell's transport reads only its own mu,
// so the gradient sweep moves them and applies the end-of-step transform itself (no
// transport launch, mu never written to HBM unless kept).
__host__ __device__ inline bool srt_local_world(const DevWorld &w) {
    return w.mode == 0 && !(w.osc_la || w.osc_flow || w.osc_div_flow);
}

// A world whose gradient-sweep cell needs none of the optional paths: Sobel gradients of
// S and H (no blur), SRT or RT, no MU_VEL, V not live, F and alpha not kept, no wall clamp
// of mu, and theta_A a power of two (alpha's division an exact multiply).  A shape class of
// only such worlds runs the sweep compiled without those paths.
__host__ __device__ inline bool plain_local_world(const DevWorld &w) {
    return w.gradient == 1 && !w.blur && w.mode != 2 && !w.mu_vel && !w.v_live && !w.need_f && !w.need_alpha &&
           w.border != 1 && w.inv_theta != 0.0f;
}

// quirk bits (same values as lenia_hip.h)
#define LHQ_SOBEL_PERMUTE   (1u << 0)
#define LHQ_PREALPHA_INDEX  (1u << 1)
#define LHQ_SUM_RESET       (1u << 2)
#define LHQ_FACE_HALO       (1u << 3)
#define LHQ_TRANSPORT_TORUS (1u << 4)
#define LHQ_RT_3CH          (1u << 5)
#define LHQ_GROWTH_LUT      (1u << 6)
#define LHQ_RT_SNORM        (1u << 7)
#define LHQ_CLAMP01         (1u << 8)
#define LHQ_CD_COPY         (1u << 9)
#define LHQ_SOFTCLIP_TR     (1u << 10)

// Arguments common to every phase launch of one shape class.
struct PhaseArgs {
    float *pool;
    float *pool2;             // staging pool (kernel spectra at commit), may be null
    int parity;               // current A buffer: 0 -> A, 1 -> A2
    const DevWorld *worlds;   // all worlds
    const int *ids;           // worlds of this shape class
    int nworlds;              // entries in ids
    const float2 *twW, *twH, *twD;  // twiddle tables (cos, sin) for the three lengths
    const unsigned short *dft;      // DFT-as-GEMM matrices of the class (fragment order), or null
};

// Which real field fwd_xy transforms.
enum { SRC_A = 0, SRC_HBLUR = 1, SRC_KERNEL = 2 };

typedef void (*launch_fn)(const PhaseArgs &, int arg, void *stream);

struct ShapeOps {
    int W, H, D;
    launch_fn fwd_xy;       // arg: SRC_*
    launch_fn spec_z;       // arg: 0 = growth kernels, 1 = blur
    launch_fn inv_growth;   // growth: y-IFFT, x-c2r, bell, H, B, U_SUM
    launch_fn inv_blur;     // blur: y-IFFT, x-c2r -> U_S, SUM_S
    launch_fn local;        // gradients, alpha, F, mu, V
    launch_fn transport;    // SRT / RT / ADV, oscillators, TRM
    launch_fn kspec_z;      // forward z on kernel fields -> real octant
    launch_fn growth_fused; // spec_z + inv_growth in one launch, or nullptr (spectrum too big for LDS)
};

// The launchers of a shape: built into the library, or from a shape module
// (liblenia_shape_<W>x<H>x<D>.so) loaded on first use; nullptr when neither exists.
const ShapeOps *lh_shape_ops(int W, int H, int D);
bool lh_shape_builtin(int W, int H, int D);
bool lh_shape_module_exists(int W, int H, int D);

// One world to seed (lh_seed_random): float offsets of its current A, GA, S, V in the
// pool, and its extents.
struct SeedJob {
    int64_t A, GA, S, V;
    int W, H, D, C;
};
void lh_launch_seed_random(float *pool, const SeedJob *d_jobs, int n, int max_cells, uint64_t seed,
                           long long first_set, hipStream_t st);

// Marching-cubes mesh of one world (lh_mesh): the view's fields (A; A, U_SUM; A, U_SUM,
// H), each C channel planes of W*H*D floats.
struct MeshJob {
    const float *field[3];
    int W, H, D, C, F, face_only;
    float iso, cell;
};
// Device memory a mesh needs, grown on demand and kept by the engine.
struct MeshBufs {
    void *work = nullptr;
    size_t work_cap = 0;
    void *out = nullptr;
    size_t out_cap = 0;
};
// The built-in triangulation table, tri[256 * 16] with -1 terminated rows.
void lh_mc_derive(int *tri);
// nullptr when tri is a usable triangulation table, else what is wrong with it.
const char *lh_mc_invalid(const int *tri);
// Counts the mesh of job (writes *nv, *ni) and, when verts and idx are given and the
// counts fit max_v and max_i, emits it into those host buffers.  Synchronises st.
hipError_t lh_mesh_run(const MeshJob &job, const int *tri, const uint32_t *colors, MeshBufs &b, hipStream_t st,
                       void *verts, int64_t max_v, uint32_t *idx, int64_t max_i, int64_t *nv, int64_t *ni);

// Growth convolution as DFT-as-GEMM on matrix cores (cubic worlds of side N); arg of
// fx/fy/fz: channels per world, of iz/iy: kernels per world.
struct GemmOps {
    int N, conv;
    launch_fn fx, fy, fz, iz, iy, ix;
};
const GemmOps *lh_gemm_ops(int n, int conv);
