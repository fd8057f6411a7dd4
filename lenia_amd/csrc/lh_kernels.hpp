// lh_kernels.hpp -- the Flow Lenia step as HIP phase kernels for gfx950 (templates).
//
// Each kernel is specialised on the world shape (W, H, D) and processes every world of
// that shape in one launch (one workgroup never straddles two worlds).  A step is
//
//   spec_z      z-FFT of the channel half spectra, x kernel spectrum, inverse z-FFT
//   inv_growth  inverse y-FFT, c2r along x, growth bell -> H, B (, U_SUM)   per z-plane
//   [blur]      fwd_xy(H,S) -> spec_z(blur) -> inv_blur -> U_S, SUM_S       central diff
//   local       gradients of S and H, alpha, F, mu, V                       per cell
//   transport   SRT / RT gather or advection, end-of-step TRM               per cell
//   fwd_xy      r2c along x + y-FFT of the new A per z-plane (next step's spectrum)
//
// Reference phases replaced: fft_step/mt_fft_step lenia.cpp:435-594, h_sum_blur
// reintegration.cpp:89-188, transpose_*_borders transposition.cpp:6-876 (halos are
// resolved on the fly from interior values), sobel_convolve[_h]/compute_f lenia.cpp:597-719,
// compute_*_mu reintegration.cpp:321-602, reintegration_step lenia.cpp:727-918 with
// srt/rt_*/adv_* reintegration.cpp:607-1410, transpose_to_row_major lenia.cpp:924-1046.
//
// All arithmetic is fp32 in a fixed order (built with -ffp-contract=off): results are
// bit-identical to the CPU restatement used as the checker.
#pragma once
#include "lh_device.hpp"

// ------------------------------------------------------------------------------------
// fwd_xy: r2c along x (rows y, y+1 packed as one complex line) and the y-FFT, for NP
// z-planes of one real field per workgroup.  Output: half spectrum planes [z][y][kx].
// ------------------------------------------------------------------------------------
template <int W, int H, int D>
struct FwdCfg {
    static constexpr int NP = (H <= 32 ? (D % 8 == 0 ? 8 : (D % 4 == 0 ? 4 : (D % 2 == 0 ? 2 : 1)))
                                       : (D % 4 == 0 ? 4 : (D % 2 == 0 ? 2 : 1)));
};
// Threads of a k_fwd_xy workgroup: one per row pair of the x transforms and one per column
// of the y transforms, in whole waves, so that neither phase takes a second pass (32^3,
// eight planes: 136 y-columns on 192 threads instead of 128, 728 -> 691 us on 4,096 worlds).
template <int W, int H, int NP>
constexpr int fwd_threads() {
    const int rows = NP * (H / 2), cols = NP * (W / 2 + 1), m = rows > cols ? rows : cols;
    return (m + 63) / 64 * 64;
}

// NP: z-planes per workgroup (FwdCfg::NP, halved for batches of few workgroups).
template <int W, int H, int D, int NP>
__global__ __launch_bounds__((fwd_threads<W, H, NP>())) void k_fwd_xy(PhaseArgs a, int arg) {
    using S = Shape<W, H, D>;
    constexpr int Wh = S::Wh, T = fwd_threads<W, H, NP>();
    __shared__ __attribute__((aligned(16))) float2 sspec[NP * H * Wh];
    const int src = arg & 0xFF, nfmax = arg >> 8;
    const int groups = D / NP;
    const int per_world = nfmax * groups;
    const int wi = blockIdx.x / per_world;
    const int f = (blockIdx.x % per_world) / groups;
    const int z0 = (blockIdx.x % groups) * NP;
    const DevWorld &w = a.worlds[a.ids[wi]];
    int nf;
    const float *in;
    float2 *out;
    if (src == SRC_A) {
        nf = w.C;
        in = A_cur(a, w) + (long long)f * S::N;
        out = reinterpret_cast<float2 *>(a.pool + w.Fsp) + (long long)f * S::Nh;
    } else if (src == SRC_HBLUR) {
        nf = w.C + 1;
        in = f < w.C ? a.pool + w.Hg + (long long)f * S::N : S_cur(a, w);
        out = reinterpret_cast<float2 *>(a.pool + w.Fsp) + (long long)f * S::Nh;
    } else {
        nf = w.nkf;
        in = a.pool2 + w.KST + (long long)f * S::N;
        out = reinterpret_cast<float2 *>(a.pool2 + w.KSP) + (long long)f * S::Nh;
    }
    if (f >= nf) return;
    const int tid = threadIdx.x;
    // the workgroup's NP planes (contiguous in HBM) staged in LDS by 16-byte loads of
    // consecutive words, rows at a pitch of W + 1 floats, in the space the spectrum takes
    // after them (NP*H*(W+1) <= NP*H*(W+2) floats)
    {
        constexpr int RP = W + 1, N4 = NP * H * W / 4, J4 = (N4 + T - 1) / T;
        const float4 *src4 = reinterpret_cast<const float4 *>(in + (long long)z0 * S::PL);
        float4 u[J4];
#pragma unroll
        for (int j = 0; j < J4; j++) u[j] = tid + j * T < N4 ? src4[tid + j * T] : make_float4(0.f, 0.f, 0.f, 0.f);
        float *stg = reinterpret_cast<float *>(sspec);
#pragma unroll
        for (int j = 0; j < J4; j++) {
            const int i = tid + j * T;
            if (i < N4) {
                float *d = stg + (4 * i / W) * RP + (4 * i) % W;
                d[0] = u[j].x; d[1] = u[j].y; d[2] = u[j].z; d[3] = u[j].w;
            }
        }
    }
    lds_barrier();
    // x: rows y, y+1 (one row pair per thread at most: T >= NP * H / 2), read from the
    // staged planes before any spectrum row is written over them
    {
        constexpr int RP = W + 1;
        static_assert(T >= NP * (H / 2), "one row pair per thread");
        const bool rows = tid < NP * (H / 2);
        const int p = tid / (H / 2), y = 2 * (tid % (H / 2));
        float2 line[W];
        if (rows) {
            const float *r0 = reinterpret_cast<const float *>(sspec) + (p * H + y) * RP, *r1 = r0 + RP;
#pragma unroll
            for (int x = 0; x < W; x++) line[x] = make_float2(r0[x], r1[x]);
        }
        lds_barrier();
        if (rows) {
            fft_line<W, false>(line, a.twW);
            float2 *o0 = sspec + (p * H + y) * Wh, *o1 = o0 + Wh;
#pragma unroll
            for (int k = 0; k < Wh; k++) {
                const float2 Z = line[k], P = line[(W - k) % W];
                o0[k] = make_float2((Z.x + P.x) * 0.5f, (Z.y - P.y) * 0.5f);
                o1[k] = make_float2((Z.y + P.y) * 0.5f, (P.x - Z.x) * 0.5f);
            }
        }
    }
    lds_barrier();
    for (int t = tid; t < NP * Wh; t += T) {
        const int p = t / Wh, kx = t % Wh;
        float2 line[H];
#pragma unroll
        for (int y = 0; y < H; y++) line[y] = sspec[(p * H + y) * Wh + kx];
        fft_line<H, false>(line, a.twH);
#pragma unroll
        for (int y = 0; y < H; y++) sspec[(p * H + y) * Wh + kx] = line[y];
    }
    lds_barrier();
    // the workgroup's planes are contiguous in the output (and H is even, so a plane is a
    // whole number of 16-byte words): written out as 16-byte words in order
    const float4 *s4 = reinterpret_cast<const float4 *>(sspec);
    float4 *o4 = reinterpret_cast<float4 *>(out + (long long)z0 * H * Wh);
    for (int i = tid; i < NP * H * Wh / 2; i += T) o4[i] = s4[i];
}

// Below this many columns in a batch (256 CUs x 256 lanes), spec_z runs one thread per
// column and kernel rather than one per column.
constexpr long long SPEC_SPLIT_COLUMNS = 65536;
// Below this many workgroups, fwd_xy gives each workgroup half as many z-planes and
// inv_growth splits each plane over two workgroups (below a quarter of it: a quarter of
// the planes, four workgroups per plane).
constexpr int FWD_SPLIT_WORKGROUPS = 1024;
constexpr int INV_SPLIT_WORKGROUPS = 1024;

// ------------------------------------------------------------------------------------
// spec_z: per (kx, ky) column.  mode 0: for each channel used as a kernel input, z-FFT,
// then for each kernel reading it: x real kernel octant, inverse z-FFT -> G_k.
// mode 1 (blur): fields 0..C (H_c, S) x blur octant -> G_f.
// ------------------------------------------------------------------------------------
template <int W, int H, int D>
__global__ __launch_bounds__(128) void k_spec_z(PhaseArgs a, int mode) {
    using S = Shape<W, H, D>;
    constexpr int Wh = S::Wh, Hh = S::Hh, COLS = Wh * H;
    constexpr int BPW = (COLS + 127) / 128;
    const int wi = blockIdx.x / BPW;
    const int col = (blockIdx.x % BPW) * 128 + threadIdx.x;
    if (col >= COLS) return;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const int kx = col % Wh, ky = col / Wh;
    const int kyo = ky <= H / 2 ? ky : H - ky;
    const float2 *Fsp = reinterpret_cast<const float2 *>(a.pool + w.Fsp);
    float2 *Gsp = reinterpret_cast<float2 *>(a.pool + w.Gsp);
    const int nsrc = mode == 0 ? w.C : w.C + 1;
    for (int c = 0; c < nsrc; c++) {
        if (mode == 0) {
            bool used = false;
            for (int k = 0; k < w.K; k++) used |= (w.kern[k].c0 == c);
            if (!used) continue;
        }
        float2 t[D];
        const float2 *src = Fsp + (long long)c * S::Nh + ky * Wh + kx;
#pragma unroll
        for (int z = 0; z < D; z++) t[z] = src[(long long)z * H * Wh];
        const int nk = mode == 0 ? w.K : 1;
        // the octant values of the kernels reading this channel, each kernel's loaded
        // before the previous kernel's stores (a load issued after a store waits for it:
        // loads and stores share one completion counter)
        auto next_k = [&](int k) {
            for (k = k + 1; k < nk; k++)
                if (mode != 0 || w.kern[k].c0 == c) return k;
            return nk;
        };
        float kv[S::Dh];
        auto load_oct = [&](int k) {
            if (k >= nk) return;
            const float *oct = mode == 0 ? a.pool + w.KH + (long long)k * S::Noct : a.pool + w.BH;
#pragma unroll
            for (int zo = 0; zo < S::Dh; zo++) kv[zo] = oct[(zo * Hh + kyo) * Wh + kx];
        };
        int k = next_k(-1);
        load_oct(k);
        fft_line<D, false>(t, a.twD);
        while (k < nk) {
            const int kn = next_k(k);
            float2 y[D];
#pragma unroll
            for (int z = 0; z < D; z++) {
                const int zo = z <= D / 2 ? z : D - z;
                y[z] = make_float2(t[z].x * kv[zo], t[z].y * kv[zo]);
            }
            load_oct(kn);
            fft_line<D, true>(y, a.twD);
            float2 *dst = Gsp + (long long)(mode == 0 ? k : c) * S::Nh + ky * Wh + kx;
#pragma unroll
            for (int z = 0; z < D; z++) dst[(long long)z * H * Wh] = y[z];
            k = kn;
        }
    }
}

// spec_z (mode 0) for batches too small to fill the GPU with one thread per column:
// one thread per (column, kernel).  Each thread repeats the forward z-FFT of its
// kernel's source column -- the same arithmetic as k_spec_z, so the same bits -- and
// its chain shrinks from 1 + K transforms to 2.
template <int W, int H, int D>
__global__ __launch_bounds__(128) void k_spec_z_split(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    constexpr int Wh = S::Wh, Hh = S::Hh, COLS = Wh * H;
    constexpr int BPW = (COLS + 127) / 128;
    const int k = blockIdx.y;
    const int wi = blockIdx.x / BPW;
    const int col = (blockIdx.x % BPW) * 128 + threadIdx.x;
    if (col >= COLS) return;
    const DevWorld &w = a.worlds[a.ids[wi]];
    if (k >= w.K) return;
    const int kx = col % Wh, ky = col / Wh;
    const int kyo = ky <= H / 2 ? ky : H - ky;
    const float2 *src = reinterpret_cast<const float2 *>(a.pool + w.Fsp) + (long long)w.kern[k].c0 * S::Nh +
                        ky * Wh + kx;
    float2 t[D];
#pragma unroll
    for (int z = 0; z < D; z++) t[z] = src[(long long)z * H * Wh];
    fft_line<D, false>(t, a.twD);
    const float *oct = a.pool + w.KH + (long long)k * S::Noct;
#pragma unroll
    for (int z = 0; z < D; z++) {
        const int zo = z <= D / 2 ? z : D - z;
        const float kv = oct[(zo * Hh + kyo) * Wh + kx];
        t[z] = make_float2(t[z].x * kv, t[z].y * kv);
    }
    fft_line<D, true>(t, a.twD);
    float2 *dst = reinterpret_cast<float2 *>(a.pool + w.Gsp) + (long long)k * S::Nh + ky * Wh + kx;
#pragma unroll
    for (int z = 0; z < D; z++) dst[(long long)z * H * Wh] = t[z];
}

// kspec_z: forward z-FFT of staged kernel spectra; keeps the real part of the octant.
template <int W, int H, int D>
__global__ __launch_bounds__(128) void k_kspec_z(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    constexpr int Wh = S::Wh, Hh = S::Hh, Dh = S::Dh, COLS = Wh * H;
    constexpr int BPW = (COLS + 127) / 128;
    const int wi = blockIdx.x / BPW;
    const int col = (blockIdx.x % BPW) * 128 + threadIdx.x;
    if (col >= COLS) return;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const int kx = col % Wh, ky = col / Wh;
    if (ky > H / 2) return;
    const float2 *Ksp = reinterpret_cast<const float2 *>(a.pool2 + w.KSP);
    for (int f = 0; f < w.nkf; f++) {
        float2 t[D];
        const float2 *src = Ksp + (long long)f * S::Nh + ky * Wh + kx;
#pragma unroll
        for (int z = 0; z < D; z++) t[z] = src[(long long)z * H * Wh];
        fft_line<D, false>(t, a.twD);
        float *oct = f < w.K ? a.pool + w.KH + (long long)f * S::Noct : a.pool + w.BH;
#pragma unroll
        for (int z = 0; z < Dh; z++) oct[(z * Hh + ky) * Wh + kx] = t[z].x;
    }
}

// ------------------------------------------------------------------------------------
// inv_growth: per (world, z-plane).  For kernels in batches of KB: inverse y-FFT,
// c2r along x, then per cell in kernel order  u = g/N; U_SUM += u; H += bell(u)*h
// (mt_fft_step lenia.cpp:496-538).  Finally B = clamp(A + H*dt) (lenia.cpp:542-592).
// ------------------------------------------------------------------------------------
// LDS <- nk spectrum planes (PLANE float2 each, plane kk at src + kk * stride) by 16-byte
// direct-to-LDS loads: each wave instruction moves 1 KB and holds no registers, so a
// workgroup's whole staging is in flight at once.  Plane sizes are multiples of 16 bytes
// (H even) and plane starts 16-byte aligned; the next barrier waits for the loads.
template <int PLANE, int T>
__device__ __forceinline__ void dma_planes(float2 *dst, const float2 *src, long long stride, int nk, int tid) {
    constexpr int BYTES = PLANE * 8, NCH = (BYTES + 1023) / 1024, WAVES = T / 64;
    static_assert(BYTES % 16 == 0, "plane size");
    const int lane = tid & 63, wave = tid >> 6;
    for (int t = wave; t < nk * NCH; t += WAVES) {
        const int kk = t / NCH, c = t % NCH;
        const int off = c * 1024 + lane * 16;
        if (off < BYTES) {
            const char *g = reinterpret_cast<const char *>(src + kk * stride) + off;
            char *l = reinterpret_cast<char *>(dst + kk * PLANE) + c * 1024;
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)g,
                                             (__attribute__((address_space(3))) void *)l, 16, 0, 0);
        }
    }
}

// The analytic growth bell out of line, so a loop calling it stays small enough to unroll.
__device__ __noinline__ float bell_growth_call(float u, float m, float s) { return bell_growth(u, m, s); }

template <int W, int H, int D>
struct InvCfg {
    static constexpr int PLANE = H * (W / 2 + 1);  // float2 per spectrum plane
    static constexpr int KB0 = (40 * 1024) / (PLANE * 8);
    static constexpr int KB = KB0 < 1 ? 1 : (KB0 > LHD_MAX_K ? LHD_MAX_K : KB0);
    static constexpr int THREADS = 256;
};


// CM: channels per world at most (3 or 4); KU: U_SUM is written; SPL: workgroups per
// plane (2 or 4 for batches of few workgroups: each repeats the y transforms, then
// takes its share of the rows through the x transforms and the growth).
template <int W, int H, int D, int CM, bool KU, int SPL>
__global__ __launch_bounds__(256, 4) void k_inv_growth(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    using Cfg = InvCfg<W, H, D>;
    constexpr int Wh = S::Wh, PL = S::PL, KB = Cfg::KB, T = Cfg::THREADS;
    static_assert(H % (2 * SPL) == 0, "row pairs split evenly");
    constexpr int PLS = PL / SPL, HP = H / 2 / SPL;  // cells and row pairs of this workgroup
    constexpr int CPT = (PLS + T - 1) / T;  // cells owned by each thread
    constexpr int CU = KU ? CM : 1;         // U_SUM accumulators
    __shared__ __attribute__((aligned(16))) float2 sg[KB * Cfg::PLANE];
    const int wi = blockIdx.x / (D * SPL), z = (blockIdx.x / SPL) % D, part = blockIdx.x % SPL;
    const int c0 = part * PLS;  // first cell of this workgroup in the plane
    const DevWorld &w = a.worlds[a.ids[wi]];
    const int tid = threadIdx.x, C = w.C, K = w.K;
    const bool lut = (w.quirks & LHQ_GROWTH_LUT) != 0;
    // growth tables are read through the cache (1 KB per kernel, reused by every cell)
    const float *glut = a.pool + w.LUT;
    const float *A = A_cur(a, w) + (long long)z * PL;
    // per-cell accumulators live in registers: h[j][c] and U_SUM us[j][c] for the CPT
    // cells tid + j*T of this plane; A is read through the cache where it is needed
    float h[CPT][CM], us[CPT][CU];
#pragma unroll
    for (int j = 0; j < CPT; j++) {
#pragma unroll
        for (int c = 0; c < CM; c++) h[j][c] = 0.0f;
#pragma unroll
        for (int c = 0; c < CU; c++) us[j][c] = 0.0f;
    }
    const float2 *G = reinterpret_cast<const float2 *>(a.pool + w.Gsp);
    float *sgf = reinterpret_cast<float *>(sg);
    for (int k0 = 0; k0 < K; k0 += KB) {
        const int nk = K - k0 < KB ? K - k0 : KB;
        lds_barrier();
        dma_planes<Cfg::PLANE, T>(sg, G + (long long)k0 * S::Nh + (long long)z * Cfg::PLANE, S::Nh, nk, tid);
        __syncthreads();  // (global memory: the direct-to-LDS loads have landed)
        for (int t = tid; t < nk * Wh; t += T) {
            const int kk = t / Wh, kx = t % Wh;
            float2 line[H];
            float2 *p = sg + kk * Cfg::PLANE + kx;
#pragma unroll
            for (int y = 0; y < H; y++) line[y] = p[y * Wh];
            fft_line<H, true>(line, a.twH);
#pragma unroll
            for (int y = 0; y < H; y++) p[y * Wh] = line[y];
        }
        lds_barrier();
        for (int t = tid; t < nk * HP; t += T) {
            const int kk = t / HP, y = 2 * (t % HP) + part * (H / SPL);
            const float2 *ra = sg + kk * Cfg::PLANE + y * Wh, *rb = ra + Wh;
            float2 line[W];
            line[0] = make_float2(ra[0].x, rb[0].x);
            line[W / 2] = make_float2(ra[W / 2].x, rb[W / 2].x);
#pragma unroll
            for (int k = 1; k < W / 2; k++) {
                const float2 av2 = ra[k], bv = rb[k];
                line[k] = make_float2(av2.x - bv.y, av2.y + bv.x);
                line[W - k] = make_float2(av2.x + bv.y, bv.x - av2.y);
            }
            fft_line<W, true>(line, a.twW);
            float *u0 = sgf + 2 * (kk * Cfg::PLANE + y * Wh), *u1 = u0 + 2 * Wh;
#pragma unroll
            for (int x = 0; x < W; x++) { u0[x] = line[x].x; u1[x] = line[x].y; }
        }
        lds_barrier();
        // growth, mt_fft_step lenia.cpp:501-538: per cell and output channel, kernels in order
        for (int kk = 0; kk < nk; kk++) {
            const DevKernel &kp = w.kern[k0 + kk];
            const int c1 = kp.c1;
            const bool asym = (w.asym_mask >> c1) & 1;
#pragma unroll
            for (int j = 0; j < CPT; j++) {
                const int i = c0 + tid + j * T;
                if (tid + j * T >= PLS) break;
                const int y = i / W, x = i % W;
                const float u = sgf[2 * (kk * Cfg::PLANE + y * Wh) + x] / (float)S::N;
                float g;
                if (lut) {
                    int iu = (int)(u * 255.0f);
                    iu = iu < 0 ? 0 : (iu > 255 ? 255 : iu);
                    g = glut[(k0 + kk) * 256 + iu];
                } else {
                    g = bell_growth(u, kp.m, kp.s);
                }
                const float hk = asym ? (g - A[(long long)c1 * S::N + i]) * kp.h : g * kp.h;
#pragma unroll
                for (int c = 0; c < CM; c++) {
                    if (c != c1) continue;
                    if (KU) us[j][c] += u;
                    h[j][c] += hk;
                }
            }
        }
    }
    float *Hg = a.pool + w.Hg + (long long)z * PL, *B = a.pool + w.Bk + (long long)z * PL;
    float *US = a.pool + w.USUM + (long long)z * PL;
    // A for B loaded before any store (loads and stores share one completion counter: a
    // load issued after a store waits for it)
    const bool with_b = w.growth_type == 1;
    float av[CPT][CM];
    if (with_b) {
#pragma unroll
        for (int j = 0; j < CPT; j++)
#pragma unroll
            for (int c = 0; c < CM; c++)
                av[j][c] = (tid + j * T < PLS && c < C) ? A[(long long)c * S::N + c0 + tid + j * T] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        const int i = c0 + tid + j * T;
        if (tid + j * T >= PLS) break;
#pragma unroll
        for (int c = 0; c < CM; c++) {
            if (c >= C) break;
            const long long g = (long long)c * S::N + i;
            const float hv = h[j][c];
            Hg[g] = hv;
            if (KU) US[g] = us[j][c];
            if (with_b)
                B[g] = ((w.soft_mask >> c) & 1) ? soft_growth(av[j][c], hv, w.dt) : clamp_ref(av[j][c] + hv * w.dt, 0.0f, 1.0f);
        }
    }
}

// ------------------------------------------------------------------------------------
// growth_fused: spec_z and inv_growth in one kernel, for worlds whose half spectrum fits
// in LDS.  One workgroup per (world, destination channel c1).  For every kernel k with
// c1(k) = c1, in kernel order: the z-FFT of its source channel's columns, the product
// with the kernel octant and the inverse z-FFT fill the whole half spectrum in LDS; the
// inverse y-FFTs and the c2r x-FFTs follow in LDS; each thread adds the growth of its two
// rows to accumulators it keeps in registers across kernels.  Every line's arithmetic is
// that of k_spec_z and k_inv_growth in the same order, and each cell's kernels are summed
// in kernel order, so H, B and U_SUM are bit-identical to that pair; the per-kernel
// spectra never reach HBM (2 x 8 bytes per half-spectrum bin and kernel saved).
// LDS: planes at a pitch of 1 mod 16 float2, so the z-columns' writes, the y-lines and
// the row pairs (taken with z fastest) all read 16 lanes from 16 distinct bank pairs.
// ------------------------------------------------------------------------------------
template <int W, int H, int D>
struct FusedCfg {
    static constexpr int Wh = W / 2 + 1;
    static constexpr int COLS = Wh * H;       // z-columns (kx, ky)
    static constexpr int YL = Wh * D;         // y-lines (kx, z)
    static constexpr int RPS = H / 2 * D;     // row pairs (y, y+1) of every plane
    static constexpr int PL = H * Wh;         // float2 per spectrum plane
    static constexpr int PP = PL + (17 - PL % 16) % 16;  // plane pitch, = 1 mod 16
    static constexpr int ZP = W * H + 1;      // float pitch of the staged output planes
    static constexpr int SPEC_BYTES = D * PP * 8;
    static constexpr int LUT_BYTES = LHD_MAX_K * 256 * 4;
    static constexpr bool FITS = SPEC_BYTES + LUT_BYTES <= 160 * 1024 && D * ZP * 4 <= SPEC_BYTES;
    static constexpr int MAXW = COLS > YL ? (COLS > RPS ? COLS : RPS) : (YL > RPS ? YL : RPS);
    static constexpr int T0 = (MAXW + 63) / 64 * 64;
    // at most 512 threads: 8 waves, two per SIMD (a ninth wave would put three on one SIMD
    // and leave the others waiting at every barrier; measured 1.3x slower at 32^3), with
    // 256 VGPRs each, room for U_SUM's 2W accumulators; the few columns and y-lines past
    // 512 take a second pass
    static constexpr int threads(bool) { return T0 > 512 ? 512 : T0; }
};

template <int W, int H, int D, int CM, bool KU>
__global__ __launch_bounds__((FusedCfg<W, H, D>::threads(KU)), 1) void k_growth_fused(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    using F = FusedCfg<W, H, D>;
    constexpr int T = F::threads(KU), Wh = F::Wh, PP = F::PP;
    constexpr int RPT = (F::RPS + T - 1) / T;  // row pairs per thread
    extern __shared__ __attribute__((aligned(16))) float2 fused_lds[];
    float2 *spec = fused_lds;
    float *slut = reinterpret_cast<float *>(fused_lds + D * PP);
    // XCD-aware order: workgroups are dealt to the 8 XCDs in turn, so the CM workgroups of
    // one world are placed 8 apart -- on one XCD, sharing its L2 for the source spectra
    constexpr int NX = 8;
    const int b = blockIdx.x, r = b % NX, m = (b / NX) % CM, g = b / (NX * CM);
    const int wi = g * NX + r, c1 = m;
    if (wi >= a.nworlds) return;
    // the world's scalars, read once (the kernels' parameters are read per kernel through
    // the pointer: a copy of the whole descriptor indexed by kernel would live in scratch)
    const DevWorld *wp = a.worlds + a.ids[wi];
    const int C = wp->C, K = wp->K, growth_type = wp->growth_type;
    if (c1 >= C) return;
    const int tid = threadIdx.x;
    const bool lut = (wp->quirks & LHQ_GROWTH_LUT) != 0;
    const bool asym = (wp->asym_mask >> c1) & 1, soft = (wp->soft_mask >> c1) & 1;
    const float dt = wp->dt;
    const float *Ac = a.pool + (a.parity ? wp->A2 : wp->A) + (long long)c1 * S::N;
    const long long kh_off = wp->KH, lut_off = wp->LUT;
    float *Hg = a.pool + wp->Hg + (long long)c1 * S::N, *B = a.pool + wp->Bk + (long long)c1 * S::N;
    float *US = a.pool + wp->USUM + (long long)c1 * S::N;
    const float2 *Fsp = reinterpret_cast<const float2 *>(a.pool + wp->Fsp);
    if (lut)
        for (int i = tid; i < K * 256; i += T) slut[i] = a.pool[lut_off + i];
    float h[RPT][2 * W], us[RPT][KU ? 2 * W : 1];
#pragma unroll
    for (int j = 0; j < RPT; j++)
#pragma unroll
        for (int x = 0; x < 2 * W; x++) {
            h[j][x] = 0.0f;
            if (KU) us[j][x] = 0.0f;
        }
    for (int k = 0; k < K; k++) {
        const DevKernel kp = wp->kern[k];
        if (kp.c1 != c1) continue;
        const float *oct = a.pool + kh_off + (long long)k * S::Noct;
        lds_barrier();  // the previous kernel's spectrum (and the table) are no longer read
        // z-columns: forward z-FFT of the source channel, x kernel octant, inverse z-FFT
        for (int col = tid; col < F::COLS; col += T) {
            const int kx = col % Wh, ky = col / Wh;
            const int kyo = ky <= H / 2 ? ky : H - ky;
            const float2 *src = Fsp + (long long)kp.c0 * S::Nh + col;
            float2 t[D];
#pragma unroll
            for (int z = 0; z < D; z++) t[z] = src[(long long)z * F::PL];
            fft_line<D, false>(t, a.twD);
#pragma unroll
            for (int z = 0; z < D; z++) {
                const int zo = z <= D / 2 ? z : D - z;
                const float kv = oct[(zo * S::Hh + kyo) * Wh + kx];
                t[z] = make_float2(t[z].x * kv, t[z].y * kv);
            }
            fft_line<D, true>(t, a.twD);
#pragma unroll
            for (int z = 0; z < D; z++) spec[z * PP + col] = t[z];
        }
        lds_barrier();
        // y-lines: inverse y-FFT in place
        for (int ln = tid; ln < F::YL; ln += T) {
            const int kx = ln % Wh, z = ln / Wh;
            float2 *p = spec + z * PP + kx;
            float2 line[H];
#pragma unroll
            for (int y = 0; y < H; y++) line[y] = p[y * Wh];
            fft_line<H, true>(line, a.twH);
#pragma unroll
            for (int y = 0; y < H; y++) p[y * Wh] = line[y];
        }
        lds_barrier();
        // row pairs (z fastest): c2r along x, then the growth of the pair's 2W cells,
        // mt_fft_step lenia.cpp:501-538
#pragma unroll
        for (int j = 0; j < RPT; j++) {
            const int rp = tid + j * T;
            if (rp >= F::RPS) break;
            const int z = rp % D, y = 2 * (rp / D);
            const float2 *ra = spec + z * PP + y * Wh, *rb = ra + Wh;
            float2 line[W];
            line[0] = make_float2(ra[0].x, rb[0].x);
            line[W / 2] = make_float2(ra[W / 2].x, rb[W / 2].x);
#pragma unroll
            for (int q = 1; q < W / 2; q++) {
                const float2 av = ra[q], bv = rb[q];
                line[q] = make_float2(av.x - bv.y, av.y + bv.x);
                line[W - q] = make_float2(av.x + bv.y, bv.x - av.y);
            }
            fft_line<W, true>(line, a.twW);
            const int i0 = (z * H + y) * W;
            // fully unrolled loops without branches inside (the accumulators must stay
            // in registers, and a branch per cell would serialise the table reads): the
            // table lookup inline, A of the two rows loaded ahead for an asymptotic
            // channel, the analytic bell as a call
            if (lut && !asym) {
#pragma unroll
                for (int e = 0; e < 2 * W; e++) {
                    const int x = e % W, r = e / W;  // cell (x, y + r)
                    const float u = (r == 0 ? line[x].x : line[x].y) / (float)S::N;
                    int iu = (int)(u * 255.0f);
                    iu = iu < 0 ? 0 : (iu > 255 ? 255 : iu);
                    const float g = slut[k * 256 + iu];
                    const float hk = g * kp.h;
                    if (KU) us[j][e] += u;
                    h[j][e] += hk;
                }
            } else if (lut) {
                float av[2 * W];
                const float4 *a4 = reinterpret_cast<const float4 *>(Ac + i0);
#pragma unroll
                for (int q = 0; q < W / 2; q++) {
                    const float4 v = a4[q];
                    av[4 * q] = v.x; av[4 * q + 1] = v.y; av[4 * q + 2] = v.z; av[4 * q + 3] = v.w;
                }
#pragma unroll
                for (int e = 0; e < 2 * W; e++) {
                    const int x = e % W, r = e / W;
                    const float u = (r == 0 ? line[x].x : line[x].y) / (float)S::N;
                    int iu = (int)(u * 255.0f);
                    iu = iu < 0 ? 0 : (iu > 255 ? 255 : iu);
                    const float g = slut[k * 256 + iu];
                    const float hk = (g - av[e]) * kp.h;
                    if (KU) us[j][e] += u;
                    h[j][e] += hk;
                }
            } else {
#pragma unroll
                for (int e = 0; e < 2 * W; e++) {
                    const int x = e % W, r = e / W;
                    const float u = (r == 0 ? line[x].x : line[x].y) / (float)S::N;
                    const float g = bell_growth_call(u, kp.m, kp.s);
                    const float hk = asym ? (g - Ac[i0 + r * W + x]) * kp.h : g * kp.h;
                    if (KU) us[j][e] += u;
                    h[j][e] += hk;
                }
            }
        }
    }
    // H, B and U_SUM through LDS (planes at a pitch of W*H + 1 floats) so the HBM writes
    // are whole rows; the spectrum is no longer needed
    float *stage = reinterpret_cast<float *>(fused_lds);
#pragma unroll
    for (int pass = 0; pass < (KU ? 2 : 1); pass++) {
        lds_barrier();
#pragma unroll
        for (int j = 0; j < RPT; j++) {
            const int rp = tid + j * T;
            if (rp >= F::RPS) break;
            const int z = rp % D, y = 2 * (rp / D);
#pragma unroll
            for (int e = 0; e < 2 * W; e++) stage[z * F::ZP + y * W + e] = pass == 0 ? h[j][e] : us[j][KU ? e : 0];
        }
        lds_barrier();
        // this thread's A values are loaded first, all in flight before any store: loads
        // and stores share one completion counter, so a load issued after a store would
        // wait for that store too (one memory round trip per cell instead of one in all)
        constexpr int PER = (S::N + T - 1) / T;
        const bool with_b = pass == 0 && growth_type == 1;
        float av[PER];
        if (with_b) {
#pragma unroll
            for (int j = 0; j < PER; j++) {
                const int i = tid + j * T;
                av[j] = i < S::N ? Ac[i] : 0.0f;
            }
        }
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int i = tid + j * T;
            if (i >= S::N) break;
            const int z = i / S::PL, r = i % S::PL;
            const float v = stage[z * F::ZP + r];
            if (pass == 0) {
                Hg[i] = v;
                if (with_b) B[i] = soft ? soft_growth(av[j], v, dt) : clamp_ref(av[j] + v * dt, 0.0f, 1.0f);
            } else {
                US[i] = v;
            }
        }
    }
}

// inv_blur: inverse y-FFT + c2r of the C+1 blurred planes; U_S = g/N, SUM_S = g/N
// (h_sum_blur reintegration.cpp:156-177).
template <int W, int H, int D>
__global__ __launch_bounds__(256) void k_inv_blur(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    using Cfg = InvCfg<W, H, D>;
    constexpr int Wh = S::Wh, PL = S::PL, KB = Cfg::KB, T = Cfg::THREADS;
    __shared__ __attribute__((aligned(16))) float2 sg[KB * Cfg::PLANE];
    const int wi = blockIdx.x / D, z = blockIdx.x % D;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const int tid = threadIdx.x, NF = w.C + 1;
    const float2 *G = reinterpret_cast<const float2 *>(a.pool + w.Gsp);
    float *sgf = reinterpret_cast<float *>(sg);
    for (int f0 = 0; f0 < NF; f0 += KB) {
        const int nk = NF - f0 < KB ? NF - f0 : KB;
        lds_barrier();
        dma_planes<Cfg::PLANE, T>(sg, G + (long long)f0 * S::Nh + (long long)z * Cfg::PLANE, S::Nh, nk, tid);
        __syncthreads();  // (global memory: the direct-to-LDS loads have landed)
        for (int t = tid; t < nk * Wh; t += T) {
            const int kk = t / Wh, kx = t % Wh;
            float2 line[H];
            float2 *p = sg + kk * Cfg::PLANE + kx;
#pragma unroll
            for (int y = 0; y < H; y++) line[y] = p[y * Wh];
            fft_line<H, true>(line, a.twH);
#pragma unroll
            for (int y = 0; y < H; y++) p[y * Wh] = line[y];
        }
        lds_barrier();
        for (int t = tid; t < nk * (H / 2); t += T) {
            const int kk = t / (H / 2), y = 2 * (t % (H / 2));
            const float2 *ra = sg + kk * Cfg::PLANE + y * Wh, *rb = ra + Wh;
            float2 line[W];
            line[0] = make_float2(ra[0].x, rb[0].x);
            line[W / 2] = make_float2(ra[W / 2].x, rb[W / 2].x);
#pragma unroll
            for (int k = 1; k < W / 2; k++) {
                const float2 av = ra[k], bv = rb[k];
                line[k] = make_float2(av.x - bv.y, av.y + bv.x);
                line[W - k] = make_float2(av.x + bv.y, bv.x - av.y);
            }
            fft_line<W, true>(line, a.twW);
            float *u0 = sgf + 2 * (kk * Cfg::PLANE + y * Wh), *u1 = u0 + 2 * Wh;
#pragma unroll
            for (int x = 0; x < W; x++) { u0[x] = line[x].x; u1[x] = line[x].y; }
        }
        lds_barrier();
        for (int i = tid; i < nk * PL; i += T) {
            const int kk = i / PL, j = i % PL, y = j / W, x = j % W;
            const int f = f0 + kk;
            const float v = sgf[2 * (kk * Cfg::PLANE + y * Wh) + x] / (float)S::N;
            float *dst = f < w.C ? a.pool + w.US + (long long)f * S::N : a.pool + w.SUMS;
            dst[(long long)z * PL + j] = v;
        }
    }
}

// ------------------------------------------------------------------------------------
// Shape table.
// ------------------------------------------------------------------------------------
template <int W, int H, int D>
struct Launch {
    // arg = source | (max fields per world) << 8
    static void fwd_xy(const PhaseArgs &a, int arg, void *st) {
        constexpr int NP = FwdCfg<W, H, D>::NP, NP2 = NP > 1 ? NP / 2 : 1, NP4 = NP > 3 ? NP / 4 : NP2;
        const int grid = a.nworlds * (arg >> 8) * (D / NP);
        if (grid < FWD_SPLIT_WORKGROUPS / 4 && NP4 < NP2)  // very few workgroups: a quarter of the planes each
            hipLaunchKernelGGL((k_fwd_xy<W, H, D, NP4>), dim3(a.nworlds * (arg >> 8) * (D / NP4)), dim3(fwd_threads<W, H, NP4>()), 0,
                               (hipStream_t)st, a, arg);
        else if (grid < FWD_SPLIT_WORKGROUPS && NP2 < NP)  // few workgroups: half the planes each, twice as many
            hipLaunchKernelGGL((k_fwd_xy<W, H, D, NP2>), dim3(a.nworlds * (arg >> 8) * (D / NP2)), dim3(fwd_threads<W, H, NP2>()), 0,
                               (hipStream_t)st, a, arg);
        else
            hipLaunchKernelGGL((k_fwd_xy<W, H, D, NP>), dim3(grid), dim3(fwd_threads<W, H, NP>()), 0, (hipStream_t)st, a, arg);
    }
    static void spec_z(const PhaseArgs &a, int mode, void *st) {
        constexpr int COLS = (W / 2 + 1) * H, BPW = (COLS + 127) / 128;
        if (mode == 0 && (long long)a.nworlds * COLS < SPEC_SPLIT_COLUMNS)
            hipLaunchKernelGGL((k_spec_z_split<W, H, D>), dim3(a.nworlds * BPW, LHD_MAX_K), dim3(128), 0,
                               (hipStream_t)st, a, 0);
        else
            hipLaunchKernelGGL((k_spec_z<W, H, D>), dim3(a.nworlds * BPW), dim3(128), 0, (hipStream_t)st, a, mode);
    }
    static void kspec_z(const PhaseArgs &a, int, void *st) {
        constexpr int BPW = ((W / 2 + 1) * H + 127) / 128;
        hipLaunchKernelGGL((k_kspec_z<W, H, D>), dim3(a.nworlds * BPW), dim3(128), 0, (hipStream_t)st, a, 0);
    }
    // arg: bit 0 U_SUM kept; maxC << 8
    static void inv_growth(const PhaseArgs &a, int arg, void *st) {
        const dim3 grid(a.nworlds * D);
        const bool ku = arg & 1, c3 = (arg >> 8) <= 3;
        if constexpr (H % 8 == 0) {
            if (a.nworlds * D < INV_SPLIT_WORKGROUPS / 4) {
                const dim3 g4(a.nworlds * D * 4);
                if (c3 && !ku) hipLaunchKernelGGL((k_inv_growth<W, H, D, 3, false, 4>), g4, dim3(256), 0, (hipStream_t)st, a, 0);
                else if (c3) hipLaunchKernelGGL((k_inv_growth<W, H, D, 3, true, 4>), g4, dim3(256), 0, (hipStream_t)st, a, 0);
                else if (!ku) hipLaunchKernelGGL((k_inv_growth<W, H, D, LHD_MAX_C, false, 4>), g4, dim3(256), 0, (hipStream_t)st, a, 0);
                else hipLaunchKernelGGL((k_inv_growth<W, H, D, LHD_MAX_C, true, 4>), g4, dim3(256), 0, (hipStream_t)st, a, 0);
                return;
            }
        }
        if constexpr (H % 4 == 0) {
            if (a.nworlds * D < INV_SPLIT_WORKGROUPS) {
                const dim3 g2(a.nworlds * D * 2);
                if (c3 && !ku) hipLaunchKernelGGL((k_inv_growth<W, H, D, 3, false, 2>), g2, dim3(256), 0, (hipStream_t)st, a, 0);
                else if (c3) hipLaunchKernelGGL((k_inv_growth<W, H, D, 3, true, 2>), g2, dim3(256), 0, (hipStream_t)st, a, 0);
                else if (!ku) hipLaunchKernelGGL((k_inv_growth<W, H, D, LHD_MAX_C, false, 2>), g2, dim3(256), 0, (hipStream_t)st, a, 0);
                else hipLaunchKernelGGL((k_inv_growth<W, H, D, LHD_MAX_C, true, 2>), g2, dim3(256), 0, (hipStream_t)st, a, 0);
                return;
            }
        }
        if (c3 && !ku) hipLaunchKernelGGL((k_inv_growth<W, H, D, 3, false, 1>), grid, dim3(256), 0, (hipStream_t)st, a, 0);
        else if (c3) hipLaunchKernelGGL((k_inv_growth<W, H, D, 3, true, 1>), grid, dim3(256), 0, (hipStream_t)st, a, 0);
        else if (!ku) hipLaunchKernelGGL((k_inv_growth<W, H, D, LHD_MAX_C, false, 1>), grid, dim3(256), 0, (hipStream_t)st, a, 0);
        else hipLaunchKernelGGL((k_inv_growth<W, H, D, LHD_MAX_C, true, 1>), grid, dim3(256), 0, (hipStream_t)st, a, 0);
    }
    static void inv_blur(const PhaseArgs &a, int, void *st) {
        hipLaunchKernelGGL((k_inv_blur<W, H, D>), dim3(a.nworlds * D), dim3(256), 0, (hipStream_t)st, a, 0);
    }
    // spec_z + inv_growth in one launch (FusedCfg::FITS shapes); arg as inv_growth's
    template <int CM, bool KU>
    static void growth_fused_cm(const PhaseArgs &a, void *st) {
        using F = FusedCfg<W, H, D>;
        constexpr size_t lds = F::SPEC_BYTES + F::LUT_BYTES;
        // dynamic LDS above 64 KB must be allowed per kernel, once
        static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void *>(k_growth_fused<W, H, D, CM, KU>),
                                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)attr;
        hipLaunchKernelGGL((k_growth_fused<W, H, D, CM, KU>), dim3((a.nworlds + 7) / 8 * 8 * CM), dim3(F::threads(KU)),
                           lds, (hipStream_t)st, a, 0);
    }
    static void growth_fused(const PhaseArgs &a, int arg, void *st) {
        if constexpr (FusedCfg<W, H, D>::FITS) {
            const bool ku = arg & 1, c3 = (arg >> 8) <= 3;
            if (c3 && !ku) growth_fused_cm<3, false>(a, st);
            else if (c3) growth_fused_cm<3, true>(a, st);
            else if (!ku) growth_fused_cm<LHD_MAX_C, false>(a, st);
            else growth_fused_cm<LHD_MAX_C, true>(a, st);
        }
    }
    // arg: bit 0 SRT worlds moved in the sweep | maxC << 8 | planes per sweep workgroup << 16
    static void local(const PhaseArgs &a, int arg, void *st) { launch_local_sweep<W, H, D>(a, arg, (hipStream_t)st); }
    // arg bits: 1 some world enables an oscillator, 2 some world is moved by the
    // z-sweep (RT, dd = 1, no oscillator), 4 some world is not (or the tiled kernel moves
    // every world); maxC << 8; planes per sweep workgroup << 16 (0: sweep_zseg)
    static void transport(const PhaseArgs &a, int arg, void *st) {
        if (arg & 2) {
            const int maxc = (arg >> 8) & 0xff;
            if (maxc <= 3) launch_transport_sweep<W, H, D, 3>(a, maxc, arg >> 16, (hipStream_t)st);
            else launch_transport_sweep<W, H, D, LHD_MAX_C>(a, maxc, arg >> 16, (hipStream_t)st);
        }
        if (arg & 4) {
            const dim3 grid(a.nworlds * TileCfg<W, H, D>::TILES);
            const int skip = (arg & 2) ? 1 : 0;
            if (arg & 1) hipLaunchKernelGGL((k_transport_tiled<W, H, D, true>), grid, dim3(256), 0, (hipStream_t)st, a, skip);
            else hipLaunchKernelGGL((k_transport_tiled<W, H, D, false>), grid, dim3(256), 0, (hipStream_t)st, a, skip);
        }
    }
    static constexpr ShapeOps ops() {
        return ShapeOps{W, H, D, &fwd_xy, &spec_z, &inv_growth, &inv_blur, &local, &transport, &kspec_z,
                        FusedCfg<W, H, D>::FITS ? &growth_fused : nullptr};
    }
};
