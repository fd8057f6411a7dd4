// lh_engine.hip -- host side of liblenia_hip.so: world table, HBM pool, kernel spectra,
// step scheduling (eager or replayed from a captured hipGraph), C-ABI entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lenia_hip.h"
#include "lh_common.hpp"

void lh_launch_mass(const PhaseArgs &a, int nworlds, double *out, void *st);
void lh_launch_moments(const PhaseArgs &a, int nworlds, double *out, void *st);
void lh_launch_mc(const float *f, int W, int H, int D, float iso, int face_only, unsigned char *out, void *st);
void lh_launch_copy(const void *src, void *dst, size_t bytes, void *st);
void lh_build_dft_mats(int n, int conv, std::vector<uint16_t> &out);

namespace {

thread_local std::string g_err;

lh_status fail(lh_status s, const std::string &msg) {
    g_err = msg;
    return s;
}

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return fail(LH_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---- scratch of the allocation-free entry points --------------------------------------
// lh_mc_case_noalloc stages its case bytes here instead of in memory it allocates: one
// static device array, which every GPU the library runs on holds until the process ends.
// In host code `scratch` only names a host-side stand-in; its device address on a GPU
// comes from hipGetSymbolAddress with that GPU current (scratch_init, called by lh_create)
// and is kept per device in g_scratch_ptr.  Host code uses only those addresses.
constexpr size_t LH_SCRATCH_BYTES = (size_t)2 << 20;  // one byte per cell: up to 128^3 cells
static __device__ __align__(256) uint8_t scratch[LH_SCRATCH_BYTES];
constexpr int LH_SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[LH_SCRATCH_DEVICES] = {};

// Looks up the scratch array's address on `device`, which must be current.
lh_status scratch_init(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= LH_SCRATCH_DEVICES) return fail(LH_EUNSUPPORTED, "device index too large");
    if (g_scratch_ptr[device]) return LH_OK;
    void *ptr = nullptr;
    hipError_t r = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !ptr)
        return fail(LH_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(r));
    g_scratch_ptr[device] = ptr;
    return LH_OK;
}

void *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < LH_SCRATCH_DEVICES ? g_scratch_ptr[device] : nullptr;
}

int64_t align64(int64_t v) { return (v + 63) & ~int64_t(63); }

// ---- kernel field (init_kernel lenia.cpp:199-261) ------------------------------------
// Each cell's value depends on |r_x|, |r_y|, |r_z| only; values are evaluated on the
// octant, then placed and summed in the reference's loop order (r_x, r_y, r_z with r_z
// innermost) so the fp32 normaliser matches the reference's accumulation order.
float bell_f(float x, float m, float s) {
    const float q = (x - m) / s;
    const double d = (double)q;
    return (float)std::exp(-(d * d) / 2.0);
}

void kernel_field(int W, int H, int D, const lh_kernel &k, float *out) {
    const size_t N = (size_t)W * H * D;
    std::memset(out, 0, sizeof(float) * N);
    const int ow = W / 2, oh = H / 2, od = D / 2;
    const int mx = std::max(ow, std::max(oh, od)) + 1;
    std::vector<float> oct((size_t)mx * mx * mx, 0.0f);
    std::vector<char> in((size_t)mx * mx * mx, 0);
    for (int ax = 0; ax <= ow; ax++)
        for (int ay = 0; ay <= oh; ay++)
            for (int az = 0; az <= od; az++) {
                const float x = (float)ax, y = (float)ay, z = (float)az;
                float dist = (float)std::sqrt((double)(x * x + y * y + z * z));
                dist = dist / (float)k.radius * (float)k.beta_size / k.r_a;
                if (dist < (float)k.beta_size) {
                    const int fl = (int)std::floor((double)dist);
                    const int index = fl < k.beta_size - 1 ? fl : k.beta_size - 1;
                    const float fr = (float)std::fmod((double)dist, 1.0);
                    const size_t o = ((size_t)ax * mx + ay) * mx + az;
                    oct[o] = k.b[index] * bell_f(fr, 0.5f, 0.15f);
                    in[o] = 1;
                }
            }
    float k_cells = 0.0f;
    for (int rx = -ow; rx < ow; rx++)
        for (int ry = -oh; ry < oh; ry++)
            for (int rz = -od; rz < od; rz++) {
                const size_t o = ((size_t)std::abs(rx) * mx + std::abs(ry)) * mx + std::abs(rz);
                if (!in[o]) continue;
                const int px = ((rx % W) + W) % W, py = ((ry % H) + H) % H, pz = ((rz % D) + D) % D;
                out[px + (size_t)py * W + (size_t)pz * W * H] = oct[o];
                k_cells += oct[o];
            }
    for (size_t i = 0; i < N; i++) out[i] /= k_cells;
}

// 5^3 Gaussian blur kernel (precompute_blur_kernel reintegration.cpp:13-78).
void blur_field(int W, int H, int D, float *out) {
    const size_t N = (size_t)W * H * D;
    std::memset(out, 0, sizeof(float) * N);
    float k_cells = 0.0f;
    const float s = 1.5f;
    for (int rx = -2; rx <= 2; rx++)
        for (int ry = -2; ry <= 2; ry++)
            for (int rz = -2; rz <= 2; rz++) {
                const float x = std::fabs((float)rx), y = std::fabs((float)ry), z = std::fabs((float)rz);
                const float dist = (x * x + y * y + z * z);
                const float g = std::exp(-dist / (s * s * s));
                const int px = ((rx % W) + W) % W, py = ((ry % H) + H) % H, pz = ((rz % D) + D) % D;
                out[px + (size_t)py * W + (size_t)pz * W * H] = g;
                k_cells += g;
            }
    for (size_t i = 0; i < N; i++) out[i] /= k_cells;
}

// The 256 BELL_LUT entries one kernel reads (init_bell_lut lenia.cpp:93-101,
// bell_lookup lenia.h:79-84).
void bell_table(float m, float s, float *out) {
    const int im = (int)(m * 255.0f), is = (int)(s * 255.0f);
    for (int iu = 0; iu < 256; iu++)
        out[iu] = bell_f((float)iu / 255.0f, (float)im / 255.0f, (float)is / 255.0f);
}

void twiddles(int n, float2 *out) {
    for (int e = 0; e < n; e++) {
        const double a = (2.0 * 3.14159265358979323846 * (double)e) / (double)n;
        out[e] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
}

struct HostWorld {
    lh_shape shape;
    lh_params p;
    std::vector<lh_kernel> k;
    DevWorld dev;
    int cls;
    int conv = LH_CONV_FFT;
    uint64_t steps = 0;  // steps taken (restored from a checkpoint image)
};

struct ShapeClass {
    int W, H, D;
    const ShapeOps *ops;
    std::vector<int> ids;
    int *d_ids = nullptr;
    int maxC = 0, maxNkf = 0;
    bool any_blur = false;
    bool any_osc = false;
    bool any_sweep = false, any_other = false;  // transport by the z-sweep / by tiles
    bool any_srt_local = false;                 // SRT worlds moved by the gradient sweep
    bool all_plain = true;                      // every world is plain_local_world
    // schedule (lh_commit's autotuner; 0 / false = the launchers' defaults)
    int zseg_local = 0, zseg_transport = 0;
    bool tiled_transport = false;  // the tiled kernel also moves the sweep's worlds
    bool fused_growth = false;     // growth_fused instead of spec_z + inv_growth
    int conv = LH_CONV_FFT;          // growth convolution of every world in the class
    const GemmOps *gops = nullptr;   // matrix-core passes when conv != LH_CONV_FFT
    unsigned short *d_dft = nullptr; // their DFT matrices
    int maxK = 0;
};

struct Prof {
    long launches = 0;
    double ms = 0.0;
};

// One read-back slot (lh_snapshot): device copies of selected fields of a range of
// worlds, taken in stream order between steps.
struct Snapshot {
    float *d = nullptr;
    size_t cap = 0;  // floats
    hipEvent_t done = nullptr;
    bool taken = false;
    int first = 0, n = 0, fields = 0;
    std::vector<size_t> off;  // [world - first][field bit] float offset in d
};

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t rb_stream = nullptr;  // read-backs of snapshots (created on first use)
    Snapshot snap[2];
    std::vector<HostWorld> worlds;
    std::vector<ShapeClass> classes;
    bool committed = false;
    float *pool = nullptr;
    size_t pool_floats = 0;
    DevWorld *d_worlds = nullptr;
    float2 *d_tw = nullptr;
    std::map<int, size_t> tw_off;
    double *d_mass = nullptr;
    double *d_mom = nullptr;
    unsigned char *d_mc = nullptr;
    size_t mc_cap = 0;
    MeshBufs mesh;
    int parity = 0;
    bool spectra_dirty = true;
    int opt_keep_v = 0, opt_keep_usum = 0, opt_keep_mu = 0, opt_graph = 1, opt_profile = 0, opt_autotune = 1;
    int opt_schedule = -1;  // LH_OPT_SCHEDULE: -1 = autotuner / defaults, else the forced schedule
    // captured step sequences, keyed by (starting parity, steps): one graph launch
    // advances every shape class by that many steps; least recently used evicted
    struct StepGraph {
        int parity, steps;
        hipGraphExec_t exec;
        uint64_t used;
    };
    std::vector<StepGraph> graphs;
    uint64_t graph_clock = 0;
    // with several shape classes, a captured step sequence holds one branch per class
    // (classes share no data a step writes): its streams and the fork / join events,
    // created on the first such capture
    std::vector<hipStream_t> branch;
    hipEvent_t fork_ev = nullptr;
    std::vector<hipEvent_t> join_ev;
    std::vector<std::string> prof_names;
    std::map<std::string, Prof> prof;
};

PhaseArgs args_for(Engine &e, ShapeClass &c, int parity, float *pool2 = nullptr) {
    PhaseArgs a;
    a.pool = e.pool;
    a.pool2 = pool2;
    a.parity = parity;
    a.worlds = e.d_worlds;
    a.ids = c.d_ids;
    a.nworlds = (int)c.ids.size();
    a.twW = e.d_tw + e.tw_off[c.W];
    a.twH = e.d_tw + e.tw_off[c.H];
    a.twD = e.d_tw + e.tw_off[c.D];
    a.dft = c.d_dft;
    return a;
}

// Launches one phase on stream st; in profile mode brackets it with events.
void run(Engine &e, const char *name, launch_fn fn, const PhaseArgs &a, int arg, hipStream_t st) {
    if (!e.opt_profile) {
        fn(a, arg, st);
        return;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float ms = -1.0f;
    if (hipEventCreate(&t0) == hipSuccess && hipEventCreate(&t1) == hipSuccess &&
        hipEventRecord(t0, st) == hipSuccess) {
        fn(a, arg, st);
        if (hipEventRecord(t1, st) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
            hipEventElapsedTime(&ms, t0, t1) != hipSuccess)
            ms = -1.0f;
    } else {
        fn(a, arg, st);  // launched untimed
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (ms < 0.0f) return;  // no timing recorded for this launch
    const std::string key = name;
    auto it = e.prof.find(key);
    if (it == e.prof.end()) {
        e.prof_names.push_back(key);
        it = e.prof.emplace(key, Prof{}).first;
    }
    it->second.launches++;
    it->second.ms += ms;
}

int transport_arg(const ShapeClass &c) {
    const bool sweep = c.any_sweep && !c.tiled_transport, tiles = c.any_other || (c.any_sweep && c.tiled_transport);
    return (c.any_osc ? 1 : 0) | (sweep ? 2 : 0) | (tiles ? 4 : 0) | (c.maxC << 8) | (c.zseg_transport << 16);
}

int local_arg(const ShapeClass &c) {
    return (c.any_srt_local ? 1 : 0) | (c.all_plain ? 2 : 0) | (c.maxC << 8) | (c.zseg_local << 16);
}

void enqueue_spectra(Engine &e) {
    for (ShapeClass &c : e.classes) {
        if (c.gops) continue;  // the matrix-core path transforms A at the start of each step
        PhaseArgs a = args_for(e, c, e.parity);
        run(e, "fwd_xy", c.ops->fwd_xy, a, SRC_A | (c.maxC << 8), e.stream);
    }
}

// One step of one shape class, on stream st.
void enqueue_class_step(Engine &e, ShapeClass &c, int parity, hipStream_t st) {
    {
        PhaseArgs a = args_for(e, c, parity);
        if (c.gops) {
            run(e, "dft_fx", c.gops->fx, a, c.maxC, st);
            run(e, "dft_fy", c.gops->fy, a, c.maxC, st);
            run(e, "dft_fz", c.gops->fz, a, c.maxC, st);
            run(e, "dft_iz", c.gops->iz, a, c.maxK, st);
            run(e, "dft_iy", c.gops->iy, a, c.maxK, st);
            run(e, "dft_ix_growth", c.gops->ix, a, 0, st);
            run(e, "local", c.ops->local, a, local_arg(c), st);
            if (c.any_sweep || c.any_other) run(e, "transport", c.ops->transport, a, transport_arg(c), st);
            return;
        }
        if (c.fused_growth) {
            run(e, "growth_fused", c.ops->growth_fused, a, (e.opt_keep_usum ? 1 : 0) | (c.maxC << 8), st);
        } else {
            run(e, "spec_z", c.ops->spec_z, a, 0, st);
            run(e, "inv_growth", c.ops->inv_growth, a, (e.opt_keep_usum ? 1 : 0) | (c.maxC << 8), st);
        }
        if (c.any_blur) {
            run(e, "fwd_xy_blur", c.ops->fwd_xy, a, SRC_HBLUR | ((c.maxC + 1) << 8), st);
            run(e, "spec_z_blur", c.ops->spec_z, a, 1, st);
            run(e, "inv_blur", c.ops->inv_blur, a, 0, st);
        }
        run(e, "local", c.ops->local, a, local_arg(c), st);
        // (a class of SRT worlds only was moved by the gradient sweep)
        if (c.any_sweep || c.any_other) run(e, "transport", c.ops->transport, a, transport_arg(c), st);
        PhaseArgs an = args_for(e, c, parity ^ 1);
        run(e, "fwd_xy", c.ops->fwd_xy, an, SRC_A | (c.maxC << 8), st);
    }
}

void enqueue_step(Engine &e, int parity) {
    for (ShapeClass &c : e.classes) enqueue_class_step(e, c, parity, e.stream);
}

lh_status validate(const lh_shape &s, const lh_params &p, const lh_kernel *k) {
    const std::string dims = std::to_string(s.W) + "x" + std::to_string(s.H) + "x" + std::to_string(s.D);
    if (s.W % 4 || s.H % 4 || s.D % 4 || s.W < 4 || s.H < 4 || s.D < 4 || s.W > 128 || s.H > 128 || s.D > 128)
        return fail(LH_EUNSUPPORTED, "shape " + dims + " is not built: sides must be divisible by 4, 4..128");
    if (!lh_shape_ops(s.W, s.H, s.D))
        return fail(LH_EUNSUPPORTED, "shape " + dims + " is not built into this library and has no shape module "
                                     "(python -m lenia_amd.build --shape " + dims + ")");
    if (s.C < 1 || s.C > LH_MAX_C) return fail(LH_EINVAL, "channel count must be 1..4");
    if (s.K < 0 || s.K > LH_MAX_K) return fail(LH_EINVAL, "kernel count must be 0..16");
    if (s.K > 0 && !k) return fail(LH_EINVAL, "kernels pointer is null");
    for (int i = 0; i < s.K; i++) {
        if (k[i].c0 < 0 || k[i].c0 >= s.C || k[i].c1 < 0 || k[i].c1 >= s.C)
            return fail(LH_EINVAL, "kernel " + std::to_string(i) + ": channel out of range");
        if (k[i].radius < 1 || k[i].beta_size < 1 || k[i].beta_size > LH_MAX_BETA || !(k[i].r_a > 0.0f))
            return fail(LH_EINVAL, "kernel " + std::to_string(i) + ": radius/beta_size/r_a out of range");
        if (!(k[i].s > 0.0f)) return fail(LH_EINVAL, "kernel " + std::to_string(i) + ": s must be > 0");
    }
    if (p.mode < 0 || p.mode > 2 || p.border < 0 || p.border > 2 || p.sobel_border < 0 || p.sobel_border > 2 ||
        p.gradient < 0 || p.gradient > 1 || p.growth_type < 0 || p.growth_type > 2)
        return fail(LH_EINVAL, "mode/border/gradient/growth enum out of range");
    if (!(p.sigma > 0.0f) || !(p.dd >= 1.0f) || (int)p.dd > 4 || !(p.dd > p.sigma))
        return fail(LH_EINVAL, "need sigma > 0, 1 <= dd <= 4, dd > sigma");
    if (!(p.theta_a != 0.0f)) return fail(LH_EINVAL, "theta_a must be nonzero");
    return LH_OK;
}

DevWorld make_dev(const HostWorld &hw, const Engine &e) {
    const lh_params &p = hw.p;
    DevWorld d;
    std::memset(&d, 0, sizeof(d));
    d.W = hw.shape.W; d.H = hw.shape.H; d.D = hw.shape.D; d.C = hw.shape.C; d.K = hw.shape.K;
    d.mode = p.mode; d.border = p.border; d.sobel_border = p.sobel_border; d.gradient = p.gradient;
    d.growth_type = p.growth_type; d.mu_vel = p.mu_vel; d.adv_vel = p.adv_vel; d.blur = p.blur && p.gradient == 0;
    d.keep_v = e.opt_keep_v;
    d.keep_usum = e.opt_keep_usum;
    d.keep_mu = e.opt_keep_mu;
    const bool osc_vf = p.osc_flow || p.osc_div_flow;
    d.v_live = p.mu_vel || (p.mode == LH_MODE_ADV && p.adv_vel) || (osc_vf && p.osc_vel) || e.opt_keep_v;
    d.need_f = osc_vf && !p.osc_vel;
    d.need_alpha = p.osc_div_flow != 0;
    d.osc_la = p.osc_la;
    d.osc_flow = p.osc_flow;
    d.osc_div_flow = p.osc_div_flow;
    d.osc_vel = p.osc_vel;
    d.osc_la_amt = p.osc_la_amt;
    d.osc_flow_amt = p.osc_flow_amt;
    d.osc_div_flow_amt = p.osc_div_flow_amt;
    for (int c = 0; c < LH_MAX_C; c++) {
        d.asym_mask |= (p.asymptotic[c] ? 1 : 0) << c;
        d.soft_mask |= (p.soft_clip[c] ? 1 : 0) << c;
    }
    d.quirks = p.quirks;
    d.dd = (int)p.dd;
    d.dt = p.dt;
    d.sigma = p.sigma;
    d.ma = p.dd - p.sigma;  // init_offset_table lenia.cpp:143
    // RT normalisation (lenia.cpp:730-733) or the exact per-source total 1/(2 sigma)^3
    const float sigma = p.sigma;
    if (p.quirks & LH_Q_RT_SNORM) {
        const float s2 = 1.0f / (4.0f * sigma * sigma);
        const float s23 = 1.0f / (2.0f * sigma * sigma * sigma);
        d.sn = s2 * (1.0f - p.s_norm) + (p.s_norm * s23);
    } else {
        d.sn = 1.0f / (8.0f * sigma * sigma * sigma);
    }
    const float two_s = 2.0f * sigma;
    d.clip = two_s < 1.0f ? two_s : 1.0f;
    d.si = 0.5f + sigma;
    d.theta_a = p.theta_a;
    {
        // dividing by a power of two and multiplying by its exact reciprocal round the
        // same exact quotient, so the sweep may multiply (one instruction, not ten)
        int ex = 0;
        const float inv = 1.0f / p.theta_a;
        const bool pow2 = std::frexp(std::fabs(p.theta_a), &ex) == 0.5f && std::isfinite(inv) &&
                          std::frexp(std::fabs(inv), &ex) == 0.5f && (double)inv * (double)p.theta_a == 1.0;
        d.inv_theta = pow2 ? inv : 0.0f;
    }
    d.gtam = 1.0f - p.gta; d.gtas = p.gta;
    d.stam = 1.0f - p.sta; d.stas = p.sta;
    d.damp = p.v_damp;
    for (int i = 0; i < hw.shape.K; i++) {
        d.kern[i].c0 = hw.k[i].c0;
        d.kern[i].c1 = hw.k[i].c1;
        d.kern[i].h = hw.k[i].h;
        d.kern[i].m = hw.k[i].m;
        d.kern[i].s = hw.k[i].s;
    }
    d.nkf = hw.shape.K + (d.blur ? 1 : 0);
    return d;
}

// Field layout helpers.
size_t field_count(const HostWorld &w, int field) {
    const size_t N = (size_t)w.shape.W * w.shape.H * w.shape.D, C = w.shape.C;
    switch (field) {
    case LH_F_A: case LH_F_GA: case LH_F_H: case LH_F_B: case LH_F_USUM: return C * N;
    case LH_F_S: return N;
    case LH_F_V: case LH_F_MU: return 3 * C * N;
    }
    return 0;
}

int64_t field_off(const Engine &e, const HostWorld &w, int field) {
    const DevWorld &d = w.dev;
    switch (field) {
    case LH_F_A: return e.parity ? d.A2 : d.A;
    case LH_F_GA: return d.GA;
    case LH_F_S: return e.parity ? d.S2 : d.S;
    case LH_F_V: return d.V;
    case LH_F_H: return d.Hg;
    case LH_F_B: return d.Bk;
    case LH_F_MU: return d.MU;
    case LH_F_USUM: return d.USUM;
    }
    return -1;
}

// Average time of reps launches of one phase (after one untimed launch).
float time_phase(Engine &e, launch_fn fn, const PhaseArgs &a, int arg, int reps) {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float ms = -1.0f;
    if (hipEventCreate(&t0) == hipSuccess && hipEventCreate(&t1) == hipSuccess) {
        fn(a, arg, e.stream);
        if (hipEventRecord(t0, e.stream) == hipSuccess) {
            for (int r = 0; r < reps; r++) fn(a, arg, e.stream);
            if (hipEventRecord(t1, e.stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                hipEventElapsedTime(&ms, t0, t1) != hipSuccess)
                ms = -1.0f;
        }
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    return ms < 0.0f ? 1e30f : ms / reps;  // a candidate that could not be timed is never picked
}

// Schedule autotuner: per shape class, times the gradient sweep at several segment
// lengths and the transport as z-sweep (same lengths) or as tiles, on the class's own
// worlds, and keeps the fastest.  Candidates are tried from the whole depth down; a
// shorter segment (more workgroups, more halo planes staged) or the tiled transport
// replaces the current pick only when faster by more than the timing spread
// (kTuneMargin), so that two near-equal candidates do not flip between runs.  Runs on
// the freshly zeroed pool (every kernel is defined on zeros); the caller clears the pool
// afterwards.
constexpr float kTuneMargin = 0.97f;
// Worlds of one shape class from which it runs as two halves (graph branches that
// overlap): 4096 x 32^3 7.72 -> 7.58 ms per step (four quarters: 7.62).
constexpr size_t kSplitWorlds = 2048;
lh_status autotune(Engine &e) {
    for (ShapeClass &c : e.classes) {
        PhaseArgs a = args_for(e, c, 0);
        const int D = c.D;
        std::vector<int> cands = {D};
        for (int zs : {D / 2, D / 4, 8, 4, 2, 1})
            if (zs >= 1 && std::find(cands.begin(), cands.end(), zs) == cands.end()) cands.push_back(zs);
        const int reps = c.ids.size() >= 256 ? 3 : 5;
        float best = 1e30f;
        int pick_local = D;
        for (int zs : cands) {
            c.zseg_local = zs;
            const float t = time_phase(e, c.ops->local, a, local_arg(c), reps);
            if (t < best * kTuneMargin) { best = t; pick_local = zs; }
        }
        c.zseg_local = pick_local;
        if (c.ops->growth_fused && !c.gops) {
            // the spectral growth as spec_z + inv_growth, or as one fused launch
            const int garg = (e.opt_keep_usum ? 1 : 0) | (c.maxC << 8);
            const float split = time_phase(e, c.ops->spec_z, a, 0, reps) + time_phase(e, c.ops->inv_growth, a, garg, reps);
            const float fused = time_phase(e, c.ops->growth_fused, a, garg, reps);
            c.fused_growth = fused < split;
        }
        if (c.any_sweep) {
            best = 1e30f;
            int pick = 0;
            bool tiled = false;
            for (int zs : cands) {
                c.zseg_transport = zs;
                c.tiled_transport = false;
                const float t = time_phase(e, c.ops->transport, a, transport_arg(c), reps);
                if (t < best * kTuneMargin) { best = t; pick = zs; tiled = false; }
            }
            c.tiled_transport = true;
            c.zseg_transport = 0;
            const float t = time_phase(e, c.ops->transport, a, transport_arg(c), reps);
            if (t < best * kTuneMargin) { best = t; tiled = true; pick = 0; }
            c.tiled_transport = tiled;
            c.zseg_transport = pick;
        }
        HIPCHK(hipGetLastError());
    }
    return LH_OK;
}

lh_status build_kernel_spectra(Engine &e) {
    // Kernel fields are built on the host in chunks of worlds, staged in HBM, and
    // transformed by the same fwd_xy / z kernels as the state (octant of the real part).
    for (ShapeClass &c : e.classes) {
        const size_t N = (size_t)c.W * c.H * c.D, Nh = (size_t)(c.W / 2 + 1) * c.H * c.D;
        const int nkf = c.maxNkf;
        if (nkf == 0) continue;
        const size_t per_world = align64((int64_t)(nkf * N)) + align64((int64_t)(nkf * Nh * 2));
        const size_t chunk = std::max<size_t>(1, std::min<size_t>(c.ids.size(), (size_t)(2ull << 30) / (per_world * 4)));
        float *stage = nullptr;
        HIPCHK(hipMalloc(&stage, chunk * per_world * sizeof(float)));
        std::vector<float> host(chunk * nkf * N);
        int *d_chunk_ids = nullptr;
        HIPCHK(hipMalloc(&d_chunk_ids, chunk * sizeof(int)));
        for (size_t b = 0; b < c.ids.size(); b += chunk) {
            const size_t n = std::min(chunk, c.ids.size() - b);
            // host fields, one thread per world slice
            const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> th;
            for (unsigned t = 0; t < nth; t++)
                th.emplace_back([&, t]() {
                    for (size_t j = t; j < n; j += nth) {
                        HostWorld &hw = e.worlds[c.ids[b + j]];
                        float *dst = host.data() + j * nkf * N;
                        for (int k = 0; k < hw.shape.K; k++) kernel_field(c.W, c.H, c.D, hw.k[k], dst + k * N);
                        if (hw.dev.blur) blur_field(c.W, c.H, c.D, dst + hw.shape.K * N);
                    }
                });
            for (auto &t : th) t.join();
            for (size_t j = 0; j < n; j++) {
                HostWorld &hw = e.worlds[c.ids[b + j]];
                hw.dev.KST = (int64_t)(j * per_world);
                hw.dev.KSP = hw.dev.KST + align64((int64_t)(nkf * N));
                HIPCHK(hipMemcpyAsync(stage + hw.dev.KST, host.data() + j * nkf * N, sizeof(float) * hw.dev.nkf * N,
                                      hipMemcpyHostToDevice, e.stream));
                HIPCHK(hipMemcpyAsync(e.d_worlds + c.ids[b + j], &hw.dev, sizeof(DevWorld), hipMemcpyHostToDevice,
                                      e.stream));
            }
            HIPCHK(hipMemcpyAsync(d_chunk_ids, c.ids.data() + b, n * sizeof(int), hipMemcpyHostToDevice, e.stream));
            PhaseArgs a = args_for(e, c, e.parity, stage);
            a.ids = d_chunk_ids;
            a.nworlds = (int)n;
            c.ops->fwd_xy(a, SRC_KERNEL | (nkf << 8), e.stream);
            c.ops->kspec_z(a, 0, e.stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(e.stream));
        }
        HIPCHK(hipFree(d_chunk_ids));
        HIPCHK(hipFree(stage));
    }
    // growth tables
    std::vector<float> lut;
    for (HostWorld &hw : e.worlds) {
        if (hw.shape.K == 0) continue;
        lut.assign((size_t)hw.shape.K * 256, 0.0f);
        for (int k = 0; k < hw.shape.K; k++) bell_table(hw.k[k].m, hw.k[k].s, lut.data() + k * 256);
        HIPCHK(hipMemcpyAsync(e.pool + hw.dev.LUT, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice,
                              e.stream));
        HIPCHK(hipStreamSynchronize(e.stream));
    }
    return LH_OK;
}

void destroy_graphs(Engine &e) {
    for (Engine::StepGraph &g : e.graphs) (void)hipGraphExecDestroy(g.exec);
    e.graphs.clear();
}

// Longest step sequence captured as one graph, and how many graphs are kept.
constexpr int kMaxGraphSteps = 256;
constexpr size_t kMaxGraphs = 6;

// The graph advancing every world by `steps` steps from A-buffer parity `parity`,
// captured on first use: the kernels of a shape class step after step, each class's
// chain a branch of its own (forked from and joined into the engine stream) when there
// are several, so the classes' kernels may run side by side.
lh_status step_graph(Engine &e, int parity, int steps, hipGraphExec_t *out) {
    for (Engine::StepGraph &g : e.graphs)
        if (g.parity == parity && g.steps == steps) {
            g.used = ++e.graph_clock;
            *out = g.exec;
            return LH_OK;
        }
    if (e.graphs.size() >= kMaxGraphs) {
        auto lru = std::min_element(e.graphs.begin(), e.graphs.end(),
                                    [](const Engine::StepGraph &x, const Engine::StepGraph &y) { return x.used < y.used; });
        HIPCHK(hipStreamSynchronize(e.stream));  // a launched graph may still be running
        (void)hipGraphExecDestroy(lru->exec);
        e.graphs.erase(lru);
    }
    const size_t nc = e.classes.size();
    if (nc > 1 && e.branch.size() < nc) {
        if (!e.fork_ev) HIPCHK(hipEventCreateWithFlags(&e.fork_ev, hipEventDisableTiming));
        while (e.branch.size() < nc) {
            hipStream_t st = nullptr;
            hipEvent_t ev = nullptr;
            HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            e.branch.push_back(st);
            HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            e.join_ev.push_back(ev);
        }
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
    hipError_t ce = hipSuccess;
    if (nc > 1) {
        // one branch per shape class, forked from and joined back into the engine stream
        ce = hipEventRecord(e.fork_ev, e.stream);
        for (size_t i = 0; i < nc && ce == hipSuccess; i++) {
            ce = hipStreamWaitEvent(e.branch[i], e.fork_ev, 0);
            for (int s = 0; s < steps && ce == hipSuccess; s++) enqueue_class_step(e, e.classes[i], parity ^ (s & 1), e.branch[i]);
            if (ce == hipSuccess) ce = hipEventRecord(e.join_ev[i], e.branch[i]);
            if (ce == hipSuccess) ce = hipStreamWaitEvent(e.stream, e.join_ev[i], 0);
        }
    } else {
        for (int s = 0; s < steps; s++) enqueue_step(e, parity ^ (s & 1));
    }
    const hipError_t ee = hipStreamEndCapture(e.stream, &g);
    HIPCHK(ce);
    HIPCHK(ee);
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(ie);
    e.graphs.push_back(Engine::StepGraph{parity, steps, ex, ++e.graph_clock});
    *out = ex;
    return LH_OK;
}

Engine *as_engine(void *p) { return static_cast<Engine *>(p); }

}  // namespace

extern "C" {

lh_status lh_create(int device, void **out) {
    if (!out) return fail(LH_EINVAL, "null output pointer");
    *out = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(LH_EINVAL, "no such device");
    HIPCHK(hipSetDevice(device));
    if (lh_status st = scratch_init(device)) return st;
    Engine *e = new Engine();
    e->device = device;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return fail(LH_EHIP, "hipStreamCreate failed");
    }
    *out = e;
    return LH_OK;
}

lh_status lh_destroy(void *p) {
    Engine *e = as_engine(p);
    if (!e) return LH_OK;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    destroy_graphs(*e);
    for (ShapeClass &c : e->classes) {
        (void)hipFree(c.d_ids);
        (void)hipFree(c.d_dft);
    }
    (void)hipFree(e->pool);
    (void)hipFree(e->d_worlds);
    (void)hipFree(e->d_tw);
    (void)hipFree(e->d_mass);
    (void)hipFree(e->d_mom);
    (void)hipFree(e->d_mc);
    (void)hipFree(e->mesh.work);
    (void)hipFree(e->mesh.out);
    for (Snapshot &sn : e->snap) {
        (void)hipFree(sn.d);
        if (sn.done) (void)hipEventDestroy(sn.done);
    }
    if (e->rb_stream) (void)hipStreamDestroy(e->rb_stream);
    for (hipStream_t st : e->branch) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : e->join_ev) (void)hipEventDestroy(ev);
    if (e->fork_ev) (void)hipEventDestroy(e->fork_ev);
    (void)hipStreamDestroy(e->stream);
    delete e;
    return LH_OK;
}

lh_status lh_set_option(void *p, int option, int value) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    switch (option) {
    case LH_OPT_KEEP_VELOCITY:
    case LH_OPT_KEEP_USUM:
        if (e->committed) return fail(LH_ESTATE, "option must be set before lh_commit");
        (option == LH_OPT_KEEP_VELOCITY ? e->opt_keep_v : e->opt_keep_usum) = value != 0;
        return LH_OK;
    case LH_OPT_AUTOTUNE:
        if (e->committed) return fail(LH_ESTATE, "option must be set before lh_commit");
        e->opt_autotune = value != 0;
        return LH_OK;
    case LH_OPT_SCHEDULE:
        if (e->committed) return fail(LH_ESTATE, "option must be set before lh_commit");
        if (value < -1 || (value >> 19) != 0 || ((value >> 17) & 3) == 3)
            return fail(LH_EINVAL, "LH_OPT_SCHEDULE: -1, or zseg_local | zseg_transport << 8 | tiled << 16 | growth << 17");
        e->opt_schedule = value;
        return LH_OK;
    case LH_OPT_KEEP_MU:
        if (e->committed) return fail(LH_ESTATE, "option must be set before lh_commit");
        e->opt_keep_mu = value != 0;
        return LH_OK;
    case LH_OPT_USE_GRAPH: e->opt_graph = value != 0; return LH_OK;
    case LH_OPT_PROFILE: e->opt_profile = value != 0; return LH_OK;
    }
    return fail(LH_EINVAL, "unknown option");
}

lh_status lh_add_world(void *p, const lh_shape *shape, const lh_params *params, const lh_kernel *kernels,
                       int *world_id) {
    Engine *e = as_engine(p);
    if (!e || !shape || !params) return fail(LH_EINVAL, "null argument");
    if (e->committed) return fail(LH_ESTATE, "worlds must be added before lh_commit");
    lh_status s = validate(*shape, *params, kernels);
    if (s != LH_OK) return s;
    HostWorld hw;
    hw.shape = *shape;
    hw.p = *params;
    hw.k.assign(kernels, kernels + shape->K);
    hw.cls = -1;
    e->worlds.push_back(hw);
    if (world_id) *world_id = (int)e->worlds.size() - 1;
    return LH_OK;
}

lh_status lh_set_world_conv(void *p, int world, int method) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    if (e->committed) return fail(LH_ESTATE, "the convolution must be chosen before lh_commit");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    if (method < LH_CONV_FFT || method > LH_CONV_GEMM_FP16) return fail(LH_EINVAL, "unknown convolution method");
    HostWorld &hw = e->worlds[world];
    if (method != LH_CONV_FFT) {
        const lh_shape &s = hw.shape;
        if (s.W != s.H || s.W != s.D || !lh_gemm_ops(s.W, method))
            return fail(LH_EUNSUPPORTED, "matrix-core convolution needs a cube of side 32, 64 or 128");
        if (hw.p.blur) return fail(LH_EUNSUPPORTED, "matrix-core convolution has no blur pass");
    }
    hw.conv = method;
    return LH_OK;
}

lh_status lh_commit(void *p) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    if (e->committed) return fail(LH_ESTATE, "already committed");
    if (e->worlds.empty()) return fail(LH_ESTATE, "no worlds");
    HIPCHK(hipSetDevice(e->device));
    // shape classes
    for (int i = 0; i < (int)e->worlds.size(); i++) {
        HostWorld &hw = e->worlds[i];
        int ci = -1;
        for (int j = 0; j < (int)e->classes.size(); j++)
            if (e->classes[j].W == hw.shape.W && e->classes[j].H == hw.shape.H && e->classes[j].D == hw.shape.D &&
                e->classes[j].conv == hw.conv)
                ci = j;
        if (ci < 0) {
            ShapeClass c;
            c.W = hw.shape.W; c.H = hw.shape.H; c.D = hw.shape.D;
            c.ops = lh_shape_ops(c.W, c.H, c.D);
            c.conv = hw.conv;
            if (hw.conv != LH_CONV_FFT) c.gops = lh_gemm_ops(c.W, hw.conv);
            e->classes.push_back(c);
            ci = (int)e->classes.size() - 1;
        }
        hw.cls = ci;
        e->classes[ci].ids.push_back(i);
    }
    // a class of at least kSplitWorlds worlds runs as two halves, two branches of the
    // step graph whose kernels overlap (each half still fills the GPU)
    const size_t n0 = e->classes.size();
    for (size_t j = 0; j < n0; j++) {
        if (e->classes[j].ids.size() < kSplitWorlds) continue;
        ShapeClass half = e->classes[j];
        const size_t h = e->classes[j].ids.size() / 2;
        half.ids.assign(e->classes[j].ids.begin() + h, e->classes[j].ids.end());
        e->classes[j].ids.resize(h);
        e->classes.push_back(half);
        for (int i : e->classes.back().ids) e->worlds[i].cls = (int)e->classes.size() - 1;
    }
    // pool layout
    int64_t off = 0;
    auto take = [&](int64_t n) { const int64_t o = off; off += align64(std::max<int64_t>(n, 1)); return o; };
    for (HostWorld &hw : e->worlds) {
        DevWorld d = make_dev(hw, *e);
        const int64_t N = (int64_t)d.W * d.H * d.D, Nh = (int64_t)(d.W / 2 + 1) * d.H * d.D;
        const int64_t Noct = (int64_t)(d.W / 2 + 1) * (d.H / 2 + 1) * (d.D / 2 + 1);
        const int C = d.C, K = d.K;
        d.A = take(C * N);
        d.A2 = take(C * N);
        d.GA = take(C * N);
        d.S = take(N);
        d.S2 = take(N);
        d.V = take(3 * C * N);
        d.Bk = take(C * N);
        d.Hg = take(C * N);
        d.USUM = d.keep_usum ? take(C * N) : d.Hg;
        d.MU = take(3 * C * N);
        d.Fv = d.need_f ? take(3 * C * N) : d.MU;
        d.Fsp = take((C + (d.blur ? 1 : 0)) * Nh * 2);
        d.Gsp = take(std::max<int64_t>(K, d.blur ? C + 1 : 0) * Nh * 2);
        d.KH = take(K * Noct);
        d.LUT = take(K * 256);
        d.US = d.blur ? take(C * N) : d.Hg;
        d.SUMS = d.blur ? take(N) : d.S;
        d.BH = d.blur ? take(Noct) : d.KH;
        d.AL = d.need_alpha ? take(C * N) : d.Hg;
        hw.dev = d;
        ShapeClass &c = e->classes[hw.cls];
        c.maxC = std::max(c.maxC, C);
        c.maxNkf = std::max(c.maxNkf, d.nkf);
        c.maxK = std::max(c.maxK, K);
        c.any_blur |= d.blur != 0;
        c.any_osc |= (d.osc_la || d.osc_flow || d.osc_div_flow);
        c.all_plain = c.all_plain && plain_local_world(d);
        if (sweep_world(d)) c.any_sweep = true;
        else if (srt_local_world(d)) c.any_srt_local = true;
        else c.any_other = true;
    }
    e->pool_floats = (size_t)off;
    HIPCHK(hipMalloc(&e->pool, e->pool_floats * sizeof(float)));
    HIPCHK(hipMemsetAsync(e->pool, 0, e->pool_floats * sizeof(float), e->stream));
    // twiddle tables
    std::vector<float2> tw;
    for (ShapeClass &c : e->classes)
        for (int n : {c.W, c.H, c.D})
            if (!e->tw_off.count(n)) {
                e->tw_off[n] = tw.size();
                tw.resize(tw.size() + n);
                twiddles(n, tw.data() + e->tw_off[n]);
            }
    HIPCHK(hipMalloc(&e->d_tw, tw.size() * sizeof(float2)));
    HIPCHK(hipMemcpy(e->d_tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    std::vector<DevWorld> devs;
    for (HostWorld &hw : e->worlds) devs.push_back(hw.dev);
    HIPCHK(hipMalloc(&e->d_worlds, devs.size() * sizeof(DevWorld)));
    HIPCHK(hipMemcpy(e->d_worlds, devs.data(), devs.size() * sizeof(DevWorld), hipMemcpyHostToDevice));
    for (ShapeClass &c : e->classes) {
        HIPCHK(hipMalloc(&c.d_ids, c.ids.size() * sizeof(int)));
        HIPCHK(hipMemcpy(c.d_ids, c.ids.data(), c.ids.size() * sizeof(int), hipMemcpyHostToDevice));
        if (c.gops) {
            std::vector<uint16_t> mats;
            lh_build_dft_mats(c.W, c.conv, mats);
            HIPCHK(hipMalloc(&c.d_dft, mats.size() * sizeof(uint16_t)));
            HIPCHK(hipMemcpy(c.d_dft, mats.data(), mats.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
        }
    }
    HIPCHK(hipMalloc(&e->d_mass, e->worlds.size() * sizeof(double)));
    for (ShapeClass &c : e->classes)  // default: spec_z + inv_growth; the autotuner may pick the fused launch
        c.fused_growth = false;
    if (e->opt_schedule >= 0) {
        // a forced schedule (any segment length is correct; lengths above the depth are
        // the whole depth), in place of the autotuner's pick
        for (ShapeClass &c : e->classes) {
            const int zl = e->opt_schedule & 0xff, zt = (e->opt_schedule >> 8) & 0xff;
            c.zseg_local = std::min(zl, c.D);
            c.tiled_transport = c.any_sweep && ((e->opt_schedule >> 16) & 1);
            c.zseg_transport = c.tiled_transport ? 0 : std::min(zt, c.D);
            const int g = (e->opt_schedule >> 17) & 3;  // 0 default, 1 split, 2 fused where built
            if (g) c.fused_growth = c.ops->growth_fused && !c.gops && g == 2;
        }
    } else if (e->opt_autotune) {
        lh_status t = autotune(*e);
        if (t != LH_OK) return t;
        HIPCHK(hipMemsetAsync(e->pool, 0, e->pool_floats * sizeof(float), e->stream));
    }
    lh_status s = build_kernel_spectra(*e);
    if (s != LH_OK) return s;
    e->committed = true;
    e->spectra_dirty = true;
    return LH_OK;
}

lh_status lh_set_field(void *p, int world, int field, const float *src, size_t count) {
    Engine *e = as_engine(p);
    if (!e || !src) return fail(LH_EINVAL, "null argument");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    HostWorld &hw = e->worlds[world];
    if (field != LH_F_A && field != LH_F_GA && field != LH_F_S && field != LH_F_V)
        return fail(LH_EINVAL, "only state fields (A, GA, S, V) can be set");
    if (count != field_count(hw, field)) return fail(LH_EINVAL, "count does not match the field size");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync(e->pool + field_off(*e, hw, field), src, count * sizeof(float), hipMemcpyHostToDevice,
                          e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (field == LH_F_A) e->spectra_dirty = true;
    return LH_OK;
}

lh_status lh_get_field(void *p, int world, int field, float *dst, size_t count) {
    Engine *e = as_engine(p);
    if (!e || !dst) return fail(LH_EINVAL, "null argument");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    HostWorld &hw = e->worlds[world];
    if (field < LH_F_A || field > LH_F_USUM) return fail(LH_EINVAL, "bad field");
    if (field == LH_F_USUM && !hw.dev.keep_usum) return fail(LH_ESTATE, "U_SUM needs LH_OPT_KEEP_USUM");
    if (field == LH_F_MU && srt_local_world(hw.dev) && !hw.dev.keep_mu)
        return fail(LH_ESTATE, "MU of an SRT world needs LH_OPT_KEEP_MU");
    if (count != field_count(hw, field)) return fail(LH_EINVAL, "count does not match the field size");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync(dst, e->pool + field_off(*e, hw, field), count * sizeof(float), hipMemcpyDeviceToHost,
                          e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return LH_OK;
}

lh_status lh_field_device_ptr(void *p, int world, int field, float **dptr, size_t *count) {
    Engine *e = as_engine(p);
    if (!e || !dptr) return fail(LH_EINVAL, "null argument");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    if (field < LH_F_A || field > LH_F_USUM) return fail(LH_EINVAL, "bad field");
    HostWorld &hw = e->worlds[world];
    *dptr = e->pool + field_off(*e, hw, field);
    if (count) *count = field_count(hw, field);
    return LH_OK;
}

lh_status lh_state_updated(void *p) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    e->spectra_dirty = true;
    return LH_OK;
}

lh_status lh_step(void *p, int nsteps) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (nsteps < 0) return fail(LH_EINVAL, "nsteps < 0");
    HIPCHK(hipSetDevice(e->device));
    if (e->spectra_dirty) {
        enqueue_spectra(*e);
        HIPCHK(hipGetLastError());
        e->spectra_dirty = false;
    }
    for (HostWorld &w : e->worlds) w.steps += (uint64_t)nsteps;
    int left = nsteps;
    if (e->opt_graph && !e->opt_profile) {
        // one graph launch per kMaxGraphSteps steps (a call of up to that many steps is
        // a single launch)
        while (left > 0) {
            const int n = left < kMaxGraphSteps ? left : kMaxGraphSteps;
            hipGraphExec_t ex = nullptr;
            lh_status s = step_graph(*e, e->parity, n, &ex);
            if (s != LH_OK) return s;
            HIPCHK(hipGraphLaunch(ex, e->stream));
            e->parity ^= n & 1;
            left -= n;
        }
    }
    for (; left > 0; left--) {
        enqueue_step(*e, e->parity);
        e->parity ^= 1;
    }
    HIPCHK(hipGetLastError());
    return LH_OK;
}

lh_status lh_prepare_steps(void *p, int nsteps) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (nsteps < 1 || nsteps > kMaxGraphSteps) return fail(LH_EINVAL, "nsteps outside [1, 256]");
    if (!e->opt_graph || e->opt_profile) return LH_OK;
    HIPCHK(hipSetDevice(e->device));
    hipGraphExec_t ex = nullptr;
    for (int parity = 0; parity < 2; parity++) {
        lh_status s = step_graph(*e, parity, nsteps, &ex);
        if (s != LH_OK) return s;
    }
    return LH_OK;
}

lh_status lh_synchronize(void *p) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    HIPCHK(hipStreamSynchronize(e->stream));
    return LH_OK;
}

void *lh_stream(void *p) {
    Engine *e = as_engine(p);
    return e ? (void *)e->stream : nullptr;
}

lh_status lh_world_mass(void *p, double *out, int n) {
    Engine *e = as_engine(p);
    if (!e || !out) return fail(LH_EINVAL, "null argument");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (n != (int)e->worlds.size()) return fail(LH_EINVAL, "n_worlds does not match");
    HIPCHK(hipSetDevice(e->device));
    PhaseArgs a = args_for(*e, e->classes[0], e->parity);
    lh_launch_mass(a, n, e->d_mass, e->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, e->d_mass, n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return LH_OK;
}

// ---- checkpoints ---------------------------------------------------------------------
// Image: lh_checkpoint_header, lh_params, lh_kernel[K], then A, GA, S, V as float32 (the
// state fields of lh_set_field, in that order).
static const int CKPT_FIELDS[4] = {LH_F_A, LH_F_GA, LH_F_S, LH_F_V};

static size_t ckpt_payload(const HostWorld &hw) {
    size_t n = 0;
    for (int f : CKPT_FIELDS) n += field_count(hw, f) * sizeof(float);
    return n;
}

static size_t ckpt_bytes(const HostWorld &hw) {
    return sizeof(lh_checkpoint_header) + sizeof(lh_params) + hw.k.size() * sizeof(lh_kernel) + ckpt_payload(hw);
}

lh_status lh_checkpoint_size(void *p, int world, size_t *bytes) {
    Engine *e = as_engine(p);
    if (!e || !bytes) return fail(LH_EINVAL, "null argument");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    *bytes = ckpt_bytes(e->worlds[world]);
    return LH_OK;
}

lh_status lh_checkpoint_save(void *p, int world, void *buf, size_t bytes) {
    Engine *e = as_engine(p);
    if (!e || !buf) return fail(LH_EINVAL, "null argument");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    const HostWorld &hw = e->worlds[world];
    if (bytes < ckpt_bytes(hw)) return fail(LH_EINVAL, "buffer smaller than lh_checkpoint_size");
    lh_checkpoint_header h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, "LHCK", 4);
    h.version = LH_CHECKPOINT_VERSION;
    h.header_bytes = sizeof(lh_checkpoint_header);
    h.params_bytes = sizeof(lh_params);
    h.kernel_bytes = sizeof(lh_kernel);
    h.conv = hw.conv;
    h.shape = hw.shape;
    h.steps = hw.steps;
    h.payload_bytes = ckpt_payload(hw);
    char *out = static_cast<char *>(buf);
    std::memcpy(out, &h, sizeof h);
    out += sizeof h;
    std::memcpy(out, &hw.p, sizeof(lh_params));
    out += sizeof(lh_params);
    std::memcpy(out, hw.k.data(), hw.k.size() * sizeof(lh_kernel));
    out += hw.k.size() * sizeof(lh_kernel);
    HIPCHK(hipSetDevice(e->device));
    for (int f : CKPT_FIELDS) {
        const size_t nb = field_count(hw, f) * sizeof(float);
        HIPCHK(hipMemcpyAsync(out, e->pool + field_off(*e, hw, f), nb, hipMemcpyDeviceToHost, e->stream));
        out += nb;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    return LH_OK;
}

lh_status lh_checkpoint_load(void *p, int world, const void *buf, size_t bytes) {
    Engine *e = as_engine(p);
    if (!e || !buf) return fail(LH_EINVAL, "null argument");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    HostWorld &hw = e->worlds[world];
    if (bytes < sizeof(lh_checkpoint_header)) return fail(LH_EINVAL, "checkpoint image truncated");
    lh_checkpoint_header h;
    std::memcpy(&h, buf, sizeof h);
    if (std::memcmp(h.magic, "LHCK", 4) != 0) return fail(LH_EINVAL, "not a checkpoint image (magic)");
    if (h.version != LH_CHECKPOINT_VERSION) return fail(LH_EINVAL, "checkpoint version not supported");
    if (h.header_bytes != sizeof(lh_checkpoint_header) || h.params_bytes != sizeof(lh_params) ||
        h.kernel_bytes != sizeof(lh_kernel))
        return fail(LH_EINVAL, "checkpoint written with a different layout");
    if (bytes != ckpt_bytes(hw) || h.payload_bytes != ckpt_payload(hw))
        return fail(LH_EINVAL, "checkpoint size does not match this world");
    if (std::memcmp(&h.shape, &hw.shape, sizeof(lh_shape)) != 0 || h.conv != hw.conv)
        return fail(LH_EINVAL, "checkpoint of a world with another shape or convolution");
    const char *in = static_cast<const char *>(buf) + sizeof h;
    if (std::memcmp(in, &hw.p, sizeof(lh_params)) != 0)
        return fail(LH_EINVAL, "checkpoint of a world with other parameters");
    in += sizeof(lh_params);
    if (std::memcmp(in, hw.k.data(), hw.k.size() * sizeof(lh_kernel)) != 0)
        return fail(LH_EINVAL, "checkpoint of a world with other kernels");
    in += hw.k.size() * sizeof(lh_kernel);
    HIPCHK(hipSetDevice(e->device));
    for (int f : CKPT_FIELDS) {
        const size_t nb = field_count(hw, f) * sizeof(float);
        HIPCHK(hipMemcpyAsync(e->pool + field_off(*e, hw, f), in, nb, hipMemcpyHostToDevice, e->stream));
        in += nb;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    hw.steps = h.steps;
    e->spectra_dirty = true;
    return LH_OK;
}

lh_status lh_seed_random(void *p, int first_world, int n_worlds, uint64_t seed, int64_t first_set) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (n_worlds < 0 || first_world < 0 || first_world + n_worlds > (int)e->worlds.size())
        return fail(LH_EINVAL, "world range outside the engine");
    if (n_worlds == 0) return LH_OK;
    std::vector<SeedJob> jobs(n_worlds);
    int max_cells = 0;
    for (int j = 0; j < n_worlds; j++) {
        const HostWorld &hw = e->worlds[first_world + j];
        SeedJob &jb = jobs[j];
        jb.A = field_off(*e, hw, LH_F_A);
        jb.GA = field_off(*e, hw, LH_F_GA);
        jb.S = field_off(*e, hw, LH_F_S);
        jb.V = field_off(*e, hw, LH_F_V);
        jb.W = hw.shape.W;
        jb.H = hw.shape.H;
        jb.D = hw.shape.D;
        jb.C = hw.shape.C;
        max_cells = std::max(max_cells, jb.W * jb.H * jb.D);
    }
    HIPCHK(hipSetDevice(e->device));
    SeedJob *dj = nullptr;
    HIPCHK(hipMalloc(&dj, jobs.size() * sizeof(SeedJob)));
    hipError_t r = hipMemcpyAsync(dj, jobs.data(), jobs.size() * sizeof(SeedJob), hipMemcpyHostToDevice, e->stream);
    if (r == hipSuccess) {
        lh_launch_seed_random(e->pool, dj, n_worlds, max_cells, seed, (long long)first_set, e->stream);
        r = hipGetLastError();
    }
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    (void)hipFree(dj);
    if (r != hipSuccess) return fail(LH_EHIP, std::string("lh_seed_random: ") + hipGetErrorString(r));
    for (int j = 0; j < n_worlds; j++) e->worlds[first_world + j].steps = 0;
    e->spectra_dirty = true;
    return LH_OK;
}

// ---- snapshots ------------------------------------------------------------------------
static const int SNAP_FIELD[3] = {LH_F_A, LH_F_H, LH_F_USUM};  // bit b of the mask

lh_status lh_snapshot(void *p, int slot, int first_world, int n_worlds, int fields) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (slot < 0 || slot > 1) return fail(LH_EINVAL, "slot must be 0 or 1");
    if (n_worlds < 1 || first_world < 0 || first_world + n_worlds > (int)e->worlds.size())
        return fail(LH_EINVAL, "world range outside the engine");
    if (fields <= 0 || (fields & ~(LH_SNAP_A | LH_SNAP_H | LH_SNAP_USUM)))
        return fail(LH_EINVAL, "fields: a non-empty mask of LH_SNAP_A, LH_SNAP_H, LH_SNAP_USUM");
    if ((fields & LH_SNAP_USUM) && !e->opt_keep_usum) return fail(LH_ESTATE, "U_SUM needs LH_OPT_KEEP_USUM");
    HIPCHK(hipSetDevice(e->device));
    Snapshot &sn = e->snap[slot];
    std::vector<size_t> off((size_t)n_worlds * 3, SIZE_MAX);
    size_t total = 0;
    for (int j = 0; j < n_worlds; j++)
        for (int b = 0; b < 3; b++)
            if (fields & (1 << b)) {
                off[(size_t)j * 3 + b] = total;
                total += field_count(e->worlds[first_world + j], SNAP_FIELD[b]);
            }
    if (!sn.done) HIPCHK(hipEventCreateWithFlags(&sn.done, hipEventDisableTiming));
    if (total > sn.cap) {
        if (sn.d) {
            HIPCHK(hipEventSynchronize(sn.done));  // copies into the old buffer have landed
            HIPCHK(hipFree(sn.d));
            sn.d = nullptr;
            sn.cap = 0;
        }
        HIPCHK(hipMalloc(&sn.d, total * sizeof(float)));
        sn.cap = total;
    }
    for (int j = 0; j < n_worlds; j++) {
        const HostWorld &hw = e->worlds[first_world + j];
        for (int b = 0; b < 3; b++)
            if (fields & (1 << b))
                HIPCHK(hipMemcpyAsync(sn.d + off[(size_t)j * 3 + b], e->pool + field_off(*e, hw, SNAP_FIELD[b]),
                                      field_count(hw, SNAP_FIELD[b]) * sizeof(float), hipMemcpyDeviceToDevice,
                                      e->stream));
    }
    HIPCHK(hipEventRecord(sn.done, e->stream));
    sn.off.swap(off);
    sn.first = first_world;
    sn.n = n_worlds;
    sn.fields = fields;
    sn.taken = true;
    return LH_OK;
}

lh_status lh_snapshot_ready(void *p, int slot, int *ready) {
    Engine *e = as_engine(p);
    if (!e || !ready) return fail(LH_EINVAL, "null argument");
    if (slot < 0 || slot > 1) return fail(LH_EINVAL, "slot must be 0 or 1");
    const Snapshot &sn = e->snap[slot];
    if (!sn.taken) return fail(LH_ESTATE, "no snapshot in this slot");
    HIPCHK(hipSetDevice(e->device));
    const hipError_t q = hipEventQuery(sn.done);
    if (q != hipSuccess && q != hipErrorNotReady) return fail(LH_EHIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
    *ready = q == hipSuccess ? 1 : 0;
    return LH_OK;
}

lh_status lh_snapshot_read(void *p, int slot, int world, int field, float *dst, size_t count) {
    Engine *e = as_engine(p);
    if (!e || !dst) return fail(LH_EINVAL, "null argument");
    if (slot < 0 || slot > 1) return fail(LH_EINVAL, "slot must be 0 or 1");
    const Snapshot &sn = e->snap[slot];
    if (!sn.taken) return fail(LH_ESTATE, "no snapshot in this slot");
    if (world < sn.first || world >= sn.first + sn.n) return fail(LH_EINVAL, "world not in this snapshot");
    int b = -1;
    for (int q = 0; q < 3; q++)
        if (SNAP_FIELD[q] == field) b = q;
    if (b < 0 || !(sn.fields & (1 << b))) return fail(LH_EINVAL, "field not in this snapshot");
    const HostWorld &hw = e->worlds[world];
    if (count != field_count(hw, field)) return fail(LH_EINVAL, "count does not match the field size");
    HIPCHK(hipSetDevice(e->device));
    if (!e->rb_stream) HIPCHK(hipStreamCreateWithFlags(&e->rb_stream, hipStreamNonBlocking));
    HIPCHK(hipStreamWaitEvent(e->rb_stream, sn.done, 0));
    HIPCHK(hipMemcpyAsync(dst, sn.d + sn.off[(size_t)(world - sn.first) * 3 + b], count * sizeof(float),
                          hipMemcpyDeviceToHost, e->rb_stream));
    HIPCHK(hipStreamSynchronize(e->rb_stream));
    return LH_OK;
}

lh_status lh_world_steps(void *p, int world, uint64_t *steps) {
    Engine *e = as_engine(p);
    if (!e || !steps) return fail(LH_EINVAL, "null argument");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    *steps = e->worlds[world].steps;
    return LH_OK;
}

lh_status lh_world_moments(void *p, double *out, int n) {
    Engine *e = as_engine(p);
    if (!e || !out) return fail(LH_EINVAL, "null argument");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (n != (int)e->worlds.size()) return fail(LH_EINVAL, "n_worlds does not match");
    HIPCHK(hipSetDevice(e->device));
    if (!e->d_mom) HIPCHK(hipMalloc(&e->d_mom, 7 * e->worlds.size() * sizeof(double)));
    PhaseArgs a = args_for(*e, e->classes[0], e->parity);
    lh_launch_moments(a, n, e->d_mom, e->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, e->d_mom, 7 * n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return LH_OK;
}

lh_status lh_mc_case(void *p, int world, int field, int channel, float iso, int face_only, uint8_t *out) {
    Engine *e = as_engine(p);
    if (!e || !out) return fail(LH_EINVAL, "null argument");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    HostWorld &hw = e->worlds[world];
    if (field != LH_F_A && field != LH_F_H && field != LH_F_USUM) return fail(LH_EINVAL, "field must be A, H or U_SUM");
    if (field == LH_F_USUM && !hw.dev.keep_usum) return fail(LH_ESTATE, "U_SUM needs LH_OPT_KEEP_USUM");
    if (channel < 0 || channel >= hw.shape.C) return fail(LH_EINVAL, "bad channel");
    const size_t N = (size_t)hw.shape.W * hw.shape.H * hw.shape.D;
    HIPCHK(hipSetDevice(e->device));
    if (e->mc_cap < N) {
        HIPCHK(hipFree(e->d_mc));
        e->d_mc = nullptr;
        HIPCHK(hipMalloc(&e->d_mc, N));
        e->mc_cap = N;
    }
    const float *f = e->pool + field_off(*e, hw, field) + (size_t)channel * N;
    lh_launch_mc(f, hw.shape.W, hw.shape.H, hw.shape.D, iso, face_only, e->d_mc, e->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, e->d_mc, N, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return LH_OK;
}

lh_status lh_mc_case_noalloc(void *p, int world, int field, int channel, float iso, int face_only, uint8_t *out) {
    Engine *e = as_engine(p);
    if (!e || !out) return fail(LH_EINVAL, "null argument");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    HostWorld &hw = e->worlds[world];
    if (field != LH_F_A && field != LH_F_H && field != LH_F_USUM) return fail(LH_EINVAL, "field must be A, H or U_SUM");
    if (field == LH_F_USUM && !hw.dev.keep_usum) return fail(LH_ESTATE, "U_SUM needs LH_OPT_KEEP_USUM");
    if (channel < 0 || channel >= hw.shape.C) return fail(LH_EINVAL, "bad channel");
    const size_t N = (size_t)hw.shape.W * hw.shape.H * hw.shape.D;
    if (N > LH_SCRATCH_BYTES)
        return fail(LH_EUNSUPPORTED, "lh_mc_case_noalloc: world has " + std::to_string(N) +
                                         " cells, above the scratch limit of " + std::to_string(LH_SCRATCH_BYTES));
    uint8_t *stage = static_cast<uint8_t *>(scratch_ptr(e->device));
    if (!stage) return fail(LH_ESTATE, "scratch not initialised on this device");
    HIPCHK(hipSetDevice(e->device));
    const float *f = e->pool + field_off(*e, hw, field) + (size_t)channel * N;
    lh_launch_mc(f, hw.shape.W, hw.shape.H, hw.shape.D, iso, face_only, stage, e->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, stage, N, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return LH_OK;
}

lh_status lh_mc_table(int *tri) {
    if (!tri) return fail(LH_EINVAL, "null argument");
    lh_mc_derive(tri);
    return LH_OK;
}

// Checks shared by lh_mesh and lh_mesh_snapshot; fills the job's shape and parameters
// and picks the table (tri_table, or the derived one in *table).
static lh_status mesh_prepare(Engine *e, int world, int view, float iso, float cell_size, int face_only,
                              const int *&tri_table, int *table, lh_vertex *verts, int64_t max_verts,
                              uint32_t *indices, int64_t max_indices, int64_t *n_verts, int64_t *n_indices,
                              MeshJob &job) {
    if (!e || !n_verts || !n_indices) return fail(LH_EINVAL, "null argument");
    if (max_verts < 0 || max_indices < 0) return fail(LH_EINVAL, "negative capacity");
    if ((max_verts > 0 && !verts) || (max_indices > 0 && !indices)) return fail(LH_EINVAL, "null output buffer");
    if (!e->committed) return fail(LH_ESTATE, "commit first");
    if (world < 0 || world >= (int)e->worlds.size()) return fail(LH_EINVAL, "bad world id");
    if (view < 1 || view > 3) return fail(LH_EINVAL, "view must be 1 (A), 2 (A, U_SUM) or 3 (A, U_SUM, H)");
    if (!std::isfinite(iso) || !std::isfinite(cell_size)) return fail(LH_EINVAL, "iso and cell_size must be finite");
    const HostWorld &hw = e->worlds[world];
    if (view >= 2 && !hw.dev.keep_usum) return fail(LH_ESTATE, "views 2 and 3 need U_SUM: set LH_OPT_KEEP_USUM");
    if (!tri_table) {
        lh_mc_derive(table);
        tri_table = table;
    } else if (const char *why = lh_mc_invalid(tri_table)) {
        return fail(LH_EINVAL, std::string("triangulation table: ") + why);
    }
    const int64_t N = (int64_t)hw.shape.W * hw.shape.H * hw.shape.D;
    if (N * hw.shape.C * view * 15 >= ((int64_t)1 << 32))
        return fail(LH_EUNSUPPORTED, "mesh of this world exceeds 32-bit vertex and index counts");
    job.W = hw.shape.W;
    job.H = hw.shape.H;
    job.D = hw.shape.D;
    job.C = hw.shape.C;
    job.F = view;
    job.face_only = face_only ? 1 : 0;
    job.iso = iso;
    job.cell = cell_size;
    return LH_OK;
}

static lh_status mesh_finish(Engine *e, const MeshJob &job, const int *tri_table, const uint32_t *colors,
                             hipStream_t st, lh_vertex *verts, int64_t max_verts, uint32_t *indices,
                             int64_t max_indices, int64_t *n_verts, int64_t *n_indices) {
    HIPCHK(lh_mesh_run(job, tri_table, colors, e->mesh, st, verts, max_verts, indices, max_indices, n_verts,
                       n_indices));
    if (verts && (*n_verts > max_verts || *n_indices > max_indices))
        return fail(LH_EINVAL, "mesh does not fit max_verts / max_indices (counts written)");
    return LH_OK;
}

lh_status lh_mesh(void *p, int world, int view, float iso, float cell_size, int face_only, const int *tri_table,
                  const uint32_t *colors, lh_vertex *verts, int64_t max_verts, uint32_t *indices, int64_t max_indices,
                  int64_t *n_verts, int64_t *n_indices) {
    Engine *e = as_engine(p);
    int table[256 * 16];
    MeshJob job;
    const lh_status st = mesh_prepare(e, world, view, iso, cell_size, face_only, tri_table, table, verts, max_verts,
                                      indices, max_indices, n_verts, n_indices, job);
    if (st != LH_OK) return st;
    const HostWorld &hw = e->worlds[world];
    const int fields[3] = {LH_F_A, LH_F_USUM, LH_F_H};
    for (int f = 0; f < 3; f++) job.field[f] = f < view ? e->pool + field_off(*e, hw, fields[f]) : nullptr;
    HIPCHK(hipSetDevice(e->device));
    return mesh_finish(e, job, tri_table, colors, e->stream, verts, max_verts, indices, max_indices, n_verts,
                       n_indices);
}

lh_status lh_mesh_snapshot(void *p, int slot, int world, int view, float iso, float cell_size, int face_only,
                           const int *tri_table, const uint32_t *colors, lh_vertex *verts, int64_t max_verts,
                           uint32_t *indices, int64_t max_indices, int64_t *n_verts, int64_t *n_indices) {
    Engine *e = as_engine(p);
    int table[256 * 16];
    MeshJob job;
    const lh_status st = mesh_prepare(e, world, view, iso, cell_size, face_only, tri_table, table, verts, max_verts,
                                      indices, max_indices, n_verts, n_indices, job);
    if (st != LH_OK) return st;
    if (slot < 0 || slot > 1) return fail(LH_EINVAL, "slot must be 0 or 1");
    const Snapshot &sn = e->snap[slot];
    if (!sn.taken) return fail(LH_ESTATE, "no snapshot in this slot");
    if (world < sn.first || world >= sn.first + sn.n) return fail(LH_EINVAL, "world not in this snapshot");
    const int need[3] = {LH_F_A, LH_F_USUM, LH_F_H};
    for (int f = 0; f < view; f++) {
        int b = -1;
        for (int q = 0; q < 3; q++)
            if (SNAP_FIELD[q] == need[f]) b = q;
        if (!(sn.fields & (1 << b))) return fail(LH_EINVAL, "a field of this view is not in the snapshot");
        job.field[f] = sn.d + sn.off[(size_t)(world - sn.first) * 3 + b];
    }
    for (int f = view; f < 3; f++) job.field[f] = nullptr;
    HIPCHK(hipSetDevice(e->device));
    if (!e->rb_stream) HIPCHK(hipStreamCreateWithFlags(&e->rb_stream, hipStreamNonBlocking));
    HIPCHK(hipStreamWaitEvent(e->rb_stream, sn.done, 0));
    return mesh_finish(e, job, tri_table, colors, e->rb_stream, verts, max_verts, indices, max_indices, n_verts,
                       n_indices);
}

int lh_world_count(void *p) {
    Engine *e = as_engine(p);
    return e ? (int)e->worlds.size() : 0;
}

int lh_profile_count(void *p) {
    Engine *e = as_engine(p);
    return e ? (int)e->prof_names.size() : 0;
}

lh_status lh_profile_entry(void *p, int i, const char **name, long *launches, double *ms) {
    Engine *e = as_engine(p);
    if (!e || i < 0 || i >= (int)e->prof_names.size()) return fail(LH_EINVAL, "bad profile index");
    const std::string &k = e->prof_names[i];
    if (name) *name = k.c_str();
    if (launches) *launches = e->prof[k].launches;
    if (ms) *ms = e->prof[k].ms;
    return LH_OK;
}

lh_status lh_profile_reset(void *p) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    for (auto &kv : e->prof) kv.second = Prof{};
    return LH_OK;
}

lh_status lh_calibration_copy(void *p, size_t bytes) {
    Engine *e = as_engine(p);
    if (!e) return fail(LH_EINVAL, "null engine");
    bytes &= ~(size_t)15;
    if (bytes == 0) return fail(LH_EINVAL, "bytes must be >= 16");
    HIPCHK(hipSetDevice(e->device));
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, bytes));
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(LH_ENOMEM, "hipMalloc failed");
    }
    hipError_t err = hipMemsetAsync(a, 0, bytes, e->stream);
    if (err == hipSuccess) {
        lh_launch_copy(a, b, bytes, e->stream);
        err = hipStreamSynchronize(e->stream);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return fail(LH_EHIP, hipGetErrorString(err));
    return LH_OK;
}

int lh_shape_supported(int W, int H, int D) {
    if (lh_shape_builtin(W, H, D)) return 1;
    return lh_shape_module_exists(W, H, D) ? 2 : 0;
}

const char *lh_last_error(void) { return g_err.c_str(); }

const char *lh_version(void) { return "lenia_hip 0.1 gfx950"; }

}  // extern "C"

// Error reporting for the other translation units of the library (lh_mlp.hip).
lh_status lh_report_error(lh_status s, const std::string &msg) { return fail(s, msg); }
