// lh_kernels.hip -- the Flow Lenia step as HIP phase kernels for gfx950.
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
#include <hip/hip_runtime.h>
#include "lh_common.hpp"
#include "lh_fft.hpp"

#define PADR 4  // reference PADDING (render.h:291), used by the index-quirk emulation

__device__ __forceinline__ float fmax_ref(float a, float b) { return a > b ? a : b; }  // _mm_max_ps
__device__ __forceinline__ float fmin_ref(float a, float b) { return a < b ? a : b; }  // _mm_min_ps
__device__ __forceinline__ float clamp_ref(float a, float lo, float hi) { return fmin_ref(fmax_ref(a, lo), hi); }

template <int W, int H, int D>
struct Shape {
    static constexpr int N = W * H * D;
    static constexpr int PL = W * H;
    static constexpr int Wh = W / 2 + 1, Hh = H / 2 + 1, Dh = D / 2 + 1;
    static constexpr int Nh = Wh * H * D;
    static constexpr int Noct = Wh * Hh * Dh;
    static constexpr int PW = W + 2 * PADR, PH = H + 2 * PADR, PD = D + 2 * PADR;
    static constexpr long long P = (long long)PW * PH * PD;
    // reference padded index PIDX of an interior-or-pad coordinate
    __device__ static __forceinline__ long long pidx(int x, int y, int z) {
        return (long long)(x + PADR) + (long long)(y + PADR) * PW + (long long)(z + PADR) * PW * PH;
    }
};

// Resolves a possibly-out-of-range coordinate to the interior cell whose value a
// reference halo holds there.  kind: 0 torus, 1 wall (replicate), 2 zero.  With
// faces != 0 only the six faces are filled (edges/corners stay 0), as in
// transposition.cpp.  Returns false where the halo holds 0.
template <int W, int H, int D>
__device__ __forceinline__ bool halo_src(int &x, int &y, int &z, int kind, int faces) {
    const int ox = (x < 0) | (x >= W), oy = (y < 0) | (y >= H), oz = (z < 0) | (z >= D);
    const int m = ox + oy + oz;
    if (m == 0) return true;
    if (kind == 2 || (faces && m >= 2)) return false;
    if (kind == 0) {
        x = x < 0 ? x + W : (x >= W ? x - W : x);
        y = y < 0 ? y + H : (y >= H ? y - H : y);
        z = z < 0 ? z + D : (z >= D ? z - D : z);
    } else {
        x = x < 0 ? 0 : (x >= W ? W - 1 : x);
        y = y < 0 ? 0 : (y >= H ? H - 1 : y);
        z = z < 0 ? 0 : (z >= D ? D - 1 : z);
    }
    return true;
}

__device__ __forceinline__ const float *A_cur(const PhaseArgs &a, const DevWorld &w) {
    return a.pool + (a.parity ? w.A2 : w.A);
}
__device__ __forceinline__ float *A_next(const PhaseArgs &a, const DevWorld &w) {
    return a.pool + (a.parity ? w.A : w.A2);
}

// ------------------------------------------------------------------------------------
// fwd_xy: r2c along x (rows y, y+1 packed as one complex line) and the y-FFT, for NP
// z-planes of one real field per workgroup.  Output: half spectrum planes [z][y][kx].
// ------------------------------------------------------------------------------------
template <int W, int H, int D>
struct FwdCfg {
    static constexpr int NP = (H <= 32 ? (D % 8 == 0 ? 8 : (D % 4 == 0 ? 4 : (D % 2 == 0 ? 2 : 1)))
                                       : (D % 4 == 0 ? 4 : (D % 2 == 0 ? 2 : 1)));
    static constexpr int THREADS = 128;
};

template <int W, int H, int D>
__global__ __launch_bounds__(128) void k_fwd_xy(PhaseArgs a, int arg) {
    using S = Shape<W, H, D>;
    using Cfg = FwdCfg<W, H, D>;
    constexpr int NP = Cfg::NP, Wh = S::Wh;
    __shared__ float2 sspec[NP * H * Wh];
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
        in = f < w.C ? a.pool + w.Hg + (long long)f * S::N : a.pool + w.S;
        out = reinterpret_cast<float2 *>(a.pool + w.Fsp) + (long long)f * S::Nh;
    } else {
        nf = w.nkf;
        in = a.pool2 + w.KST + (long long)f * S::N;
        out = reinterpret_cast<float2 *>(a.pool2 + w.KSP) + (long long)f * S::Nh;
    }
    if (f >= nf) return;
    const int tid = threadIdx.x;
    // x: rows y, y+1 straight from HBM (W is a multiple of 4: 16-byte row starts)
    for (int t = tid; t < NP * (H / 2); t += Cfg::THREADS) {
        const int p = t / (H / 2), y = 2 * (t % (H / 2));
        const float4 *r0 = reinterpret_cast<const float4 *>(in + (long long)(z0 + p) * S::PL + y * W);
        const float4 *r1 = r0 + W / 4;
        float2 line[W];
#pragma unroll
        for (int x = 0; x < W / 4; x++) {
            const float4 u = r0[x], v = r1[x];
            line[4 * x + 0] = make_float2(u.x, v.x);
            line[4 * x + 1] = make_float2(u.y, v.y);
            line[4 * x + 2] = make_float2(u.z, v.z);
            line[4 * x + 3] = make_float2(u.w, v.w);
        }
        fft_line<W, false>(line, a.twW);
        float2 *o0 = sspec + (p * H + y) * Wh, *o1 = o0 + Wh;
#pragma unroll
        for (int k = 0; k < Wh; k++) {
            const float2 Z = line[k], P = line[(W - k) % W];
            o0[k] = make_float2((Z.x + P.x) * 0.5f, (Z.y - P.y) * 0.5f);
            o1[k] = make_float2((Z.y + P.y) * 0.5f, (P.x - Z.x) * 0.5f);
        }
    }
    __syncthreads();
    for (int t = tid; t < NP * Wh; t += Cfg::THREADS) {
        const int p = t / Wh, kx = t % Wh;
        float2 line[H];
#pragma unroll
        for (int y = 0; y < H; y++) line[y] = sspec[(p * H + y) * Wh + kx];
        fft_line<H, false>(line, a.twH);
        float2 *o = out + (long long)(z0 + p) * H * Wh + kx;
#pragma unroll
        for (int y = 0; y < H; y++) o[y * Wh] = line[y];
    }
}

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
        fft_line<D, false>(t, a.twD);
        const int nk = mode == 0 ? w.K : 1;
        for (int k = 0; k < nk; k++) {
            if (mode == 0 && w.kern[k].c0 != c) continue;
            const float *oct = mode == 0 ? a.pool + w.KH + (long long)k * S::Noct : a.pool + w.BH;
            float2 y[D];
#pragma unroll
            for (int z = 0; z < D; z++) {
                const int zo = z <= D / 2 ? z : D - z;
                const float kv = oct[(zo * Hh + kyo) * Wh + kx];
                y[z] = make_float2(t[z].x * kv, t[z].y * kv);
            }
            fft_line<D, true>(y, a.twD);
            float2 *dst = Gsp + (long long)(mode == 0 ? k : c) * S::Nh + ky * Wh + kx;
#pragma unroll
            for (int z = 0; z < D; z++) dst[(long long)z * H * Wh] = y[z];
        }
    }
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
template <int W, int H, int D>
struct InvCfg {
    static constexpr int PLANE = H * (W / 2 + 1);  // float2 per spectrum plane
    static constexpr int KB0 = (40 * 1024) / (PLANE * 8);
    static constexpr int KB = KB0 < 1 ? 1 : (KB0 > LHD_MAX_K ? LHD_MAX_K : KB0);
    static constexpr int THREADS = 256;
};

__device__ __forceinline__ float soft_growth(float g, float h, float dt) {
    // soft clip lenia.cpp:549-571 with fastexp reintegration.h:16-37
    const float a0 = 1.0f, a1 = 1.0f, a2 = 1.0f / 2, a3 = 1.0f / 2 / 3, a4 = 1.0f / 2 / 3 / 4;
    const float a5 = 1.0f / 2 / 3 / 4 / 5, a6 = 1.0f / 2 / 3 / 4 / 5 / 6, a7 = 1.0f / 2 / 3 / 4 / 5 / 6 / 7;
    float v = g + h * dt;
    v = -4.0f * (v - 0.5f);
    float r = a7 * v + a6;
    r = r * v + a5; r = r * v + a4; r = r * v + a3; r = r * v + a2; r = r * v + a1; r = r * v + a0;
    r = 1.0f + r;
    r = 1.0f / r;
    return clamp_ref(r, 0.0f, 1.0f);
}

template <int W, int H, int D>
__global__ __launch_bounds__(256) void k_inv_growth(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    using Cfg = InvCfg<W, H, D>;
    constexpr int Wh = S::Wh, PL = S::PL, KB = Cfg::KB, T = Cfg::THREADS;
    constexpr int CPT = (PL + T - 1) / T;  // cells of the plane owned by each thread
    __shared__ float2 sg[KB * Cfg::PLANE];
    __shared__ float slut[KB * 256];
    const int wi = blockIdx.x / D, z = blockIdx.x % D;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const int tid = threadIdx.x, C = w.C, K = w.K;
    const bool lut = (w.quirks & LHQ_GROWTH_LUT) != 0;
    const float *A = A_cur(a, w) + (long long)z * PL;
    // per-cell accumulators live in registers: h[j][c], U_SUM us[j][c], and A for the
    // asymptotic form, for the CPT cells tid + j*T of this plane
    float h[CPT][LHD_MAX_C], us[CPT][LHD_MAX_C], av[CPT][LHD_MAX_C];
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        const int i = tid + j * T;
#pragma unroll
        for (int c = 0; c < LHD_MAX_C; c++) {
            h[j][c] = 0.0f;
            us[j][c] = 0.0f;
            av[j][c] = (c < C && i < PL) ? A[(long long)c * S::N + i] : 0.0f;
        }
    }
    const float2 *G = reinterpret_cast<const float2 *>(a.pool + w.Gsp);
    float *sgf = reinterpret_cast<float *>(sg);
    for (int k0 = 0; k0 < K; k0 += KB) {
        const int nk = K - k0 < KB ? K - k0 : KB;
        __syncthreads();
        for (int i = tid; i < nk * Cfg::PLANE; i += T) {
            const int kk = i / Cfg::PLANE, j = i % Cfg::PLANE;
            sg[i] = G[(long long)(k0 + kk) * S::Nh + (long long)z * Cfg::PLANE + j];
        }
        if (lut)
            for (int i = tid; i < nk * 256; i += T) slut[i] = a.pool[w.LUT + (long long)k0 * 256 + i];
        __syncthreads();
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
        __syncthreads();
        for (int t = tid; t < nk * (H / 2); t += T) {
            const int kk = t / (H / 2), y = 2 * (t % (H / 2));
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
        __syncthreads();
        // growth, mt_fft_step lenia.cpp:501-538: per cell and output channel, kernels in order
        for (int kk = 0; kk < nk; kk++) {
            const DevKernel &kp = w.kern[k0 + kk];
            const int c1 = kp.c1;
            const bool asym = (w.asym_mask >> c1) & 1;
#pragma unroll
            for (int j = 0; j < CPT; j++) {
                const int i = tid + j * T;
                if (i >= PL) break;
                const int y = i / W, x = i % W;
                const float u = sgf[2 * (kk * Cfg::PLANE + y * Wh) + x] / (float)S::N;
                float g;
                if (lut) {
                    int iu = (int)(u * 255.0f);
                    iu = iu < 0 ? 0 : (iu > 255 ? 255 : iu);
                    g = slut[kk * 256 + iu];
                } else {
                    const float q = (u - kp.m) / kp.s;
                    g = expf(-(q * q) / 2.0f);
                }
#pragma unroll
                for (int c = 0; c < LHD_MAX_C; c++) {
                    if (c != c1) continue;
                    us[j][c] += u;
                    if (asym) h[j][c] += (g - av[j][c]) * kp.h;
                    else h[j][c] += g * kp.h;
                }
            }
        }
    }
    float *Hg = a.pool + w.Hg + (long long)z * PL, *B = a.pool + w.Bk + (long long)z * PL;
    float *US = a.pool + w.USUM + (long long)z * PL;
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        const int i = tid + j * T;
        if (i >= PL) break;
#pragma unroll
        for (int c = 0; c < LHD_MAX_C; c++) {
            if (c >= C) break;
            const long long g = (long long)c * S::N + i;
            const float hv = h[j][c];
            Hg[g] = hv;
            if (w.keep_usum) US[g] = us[j][c];
            if (w.growth_type == 1)
                B[g] = ((w.soft_mask >> c) & 1) ? soft_growth(av[j][c], hv, w.dt) : clamp_ref(av[j][c] + hv * w.dt, 0.0f, 1.0f);
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
    __shared__ float2 sg[KB * Cfg::PLANE];
    const int wi = blockIdx.x / D, z = blockIdx.x % D;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const int tid = threadIdx.x, NF = w.C + 1;
    const float2 *G = reinterpret_cast<const float2 *>(a.pool + w.Gsp);
    float *sgf = reinterpret_cast<float *>(sg);
    for (int f0 = 0; f0 < NF; f0 += KB) {
        const int nk = NF - f0 < KB ? NF - f0 : KB;
        __syncthreads();
        for (int i = tid; i < nk * Cfg::PLANE; i += T) {
            const int kk = i / Cfg::PLANE, j = i % Cfg::PLANE;
            sg[i] = G[(long long)(f0 + kk) * S::Nh + (long long)z * Cfg::PLANE + j];
        }
        __syncthreads();
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
        __syncthreads();
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
        __syncthreads();
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
// local: per cell.  Gradients of S and H_c (Sobel with the reference's axis order, or
// central differences), alpha (with the PREALPHA index quirk), F, mu, V.
// ------------------------------------------------------------------------------------
// ------------------------------------------------------------------------------------
// Tiled variants of local and transport.  A workgroup owns a TX x TY x TZ block of
// cells (whole x rows) and stages the block plus a one-cell halo in LDS, with every
// halo value resolved once at load time, so the per-cell loops carry no branches.
// ------------------------------------------------------------------------------------
template <int W, int H, int D>
struct TileCfg {
    static constexpr int TX = W;
    static constexpr int TY = H % 4 == 0 ? 4 : (H % 2 == 0 ? 2 : 1);
    static constexpr int TZ = D % 4 == 0 ? 4 : (D % 2 == 0 ? 2 : 1);
    static constexpr int SX = TX + 2, SY = TY + 2, SZ = TZ + 2;
    static constexpr int NS = SX * SY * SZ;   // staged slots
    static constexpr int NT = TX * TY * TZ;   // cells per tile
    static constexpr int TILES = (H / TY) * (D / TZ);
    static constexpr int THREADS = 256;
    static constexpr int CPT = (NT + THREADS - 1) / THREADS;
    __device__ static __forceinline__ int slot(int tx, int ty, int tz) { return tx + SX * (ty + SY * tz); }
};

// Sobel from a staged tile; taps in the reference order dz, dy, dx.  Tap weights are
// +-1, +-2, +-4, so each product is exact and fma(v, c, g) rounds exactly like the
// reference's separate multiply and add.
template <int SX, int SY, bool PERM>
__device__ __forceinline__ void sobel_lds(const float *sf, int tx, int ty, int tz, float (&g)[3]) {
    constexpr int SW[3] = {1, 2, 1};
    constexpr int SD[3] = {-1, 0, 1};
    g[0] = 0.0f; g[1] = 0.0f; g[2] = 0.0f;
#pragma unroll
    for (int dz = -1; dz <= 1; dz++)
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const int gx = SD[dx + 1] * SW[dy + 1] * SW[dz + 1];
                const int gy = SW[dx + 1] * SD[dy + 1] * SW[dz + 1];
                const int gz = SW[dx + 1] * SW[dy + 1] * SD[dz + 1];
                const int c0 = PERM ? gz : gx, c1 = PERM ? gx : gy, c2 = PERM ? gy : gz;
                if (c0 == 0 && c1 == 0 && c2 == 0) continue;
                const float v = sf[(tx + dx + 1) + SX * ((ty + dy + 1) + SY * (tz + dz + 1))];
                if (c0 != 0) g[0] = __builtin_fmaf(v, (float)c0, g[0]);
                if (c1 != 0) g[1] = __builtin_fmaf(v, (float)c1, g[1]);
                if (c2 != 0) g[2] = __builtin_fmaf(v, (float)c2, g[2]);
            }
}

template <int SX, int SY>
__device__ __forceinline__ void cdiff_lds(const float *sf, int tx, int ty, int tz, float (&g)[3]) {
    auto at = [&](int x, int y, int z) { return sf[(x + 1) + SX * ((y + 1) + SY * (z + 1))]; };
    g[0] = (at(tx + 1, ty, tz) - at(tx - 1, ty, tz)) * 0.5f;
    g[1] = (at(tx, ty + 1, tz) - at(tx, ty - 1, tz)) * 0.5f;
    g[2] = (at(tx, ty, tz + 1) - at(tx, ty, tz - 1)) * 0.5f;
}

template <int SX, int SY>
__device__ __forceinline__ void gradient_lds(const DevWorld &w, const float *sf, int tx, int ty, int tz, float (&g)[3]) {
    if (w.gradient == 1) {
        if (w.quirks & LHQ_SOBEL_PERMUTE) sobel_lds<SX, SY, true>(sf, tx, ty, tz, g);
        else sobel_lds<SX, SY, false>(sf, tx, ty, tz, g);
    } else {
        cdiff_lds<SX, SY>(sf, tx, ty, tz, g);
    }
}

template <int W, int H, int D>
__global__ __launch_bounds__(256) void k_local_tiled(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    using T = TileCfg<W, H, D>;
    __shared__ float sf[(LHD_MAX_C + 1) * T::NS];
    const int wi = blockIdx.x / T::TILES, tile = blockIdx.x % T::TILES;
    const int ty0 = (tile % (H / T::TY)) * T::TY, tz0 = (tile / (H / T::TY)) * T::TZ;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const int tid = threadIdx.x, C = w.C;
    const int faces = (w.quirks & LHQ_FACE_HALO) != 0;
    // Stage S (or SUM_S) and H_c (or U_S_c) with their halos.  Each slot's halo source
    // is resolved once and shared by all fields.  The non-blur central-difference copy
    // (reintegration.cpp:179-186) leaves U_S zero from padded index 5P/2 on.
    const bool cdcopy = w.gradient == 0 && !w.blur && (w.quirks & LHQ_CD_COPY);
    const float *fbase[LHD_MAX_C + 1];
    fbase[0] = (w.gradient == 0 && w.blur) ? a.pool + w.SUMS : a.pool + w.S;
#pragma unroll
    for (int c = 0; c < LHD_MAX_C; c++)
        fbase[c + 1] = ((w.gradient == 0 && w.blur) ? a.pool + w.US : a.pool + w.Hg) + c * S::N;
    for (int q = tid; q < T::NS; q += T::THREADS) {
        const int sx = q % T::SX, sy = (q / T::SX) % T::SY, sz = q / (T::SX * T::SY);
        int x = sx - 1, y = ty0 + sy - 1, z = tz0 + sz - 1;
        const bool ok = halo_src<W, H, D>(x, y, z, w.sobel_border, faces);
        const int lin = x + W * (y + H * z);
        const int pq = (int)S::pidx(x, y, z);
#pragma unroll
        for (int f = 0; f <= LHD_MAX_C; f++) {
            if (f > C) break;
            const bool cut = cdcopy && f > 0 && (long long)pq + (long long)(f - 1) * S::P >= 5 * S::P / 2;
            sf[f * T::NS + q] = (ok && !cut) ? fbase[f][lin] : 0.0f;
        }
    }
    __syncthreads();
    const float *A = A_cur(a, w);
    const float dt = w.dt, ma = w.ma, nma = -w.ma, damp = w.damp, sig = w.sigma;
    const float lim[3] = {(float)W - sig, (float)H - sig, (float)D - sig};
    const bool wall = w.border == 1;
    const bool qa = (w.quirks & LHQ_PREALPHA_INDEX) != 0;
#pragma unroll
    for (int j = 0; j < T::CPT; j++) {
        const int cell = tid + j * T::THREADS;
        if (cell >= T::NT) break;
        const int tx = cell % T::TX, ty = (cell / T::TX) % T::TY, tz = cell / (T::TX * T::TY);
        const int x = tx, y = ty0 + ty, z = tz0 + tz;
        const long long i = x + (long long)W * (y + (long long)H * z);
        float na[3];
        gradient_lds<T::SX, T::SY>(w, sf, tx, ty, tz, na);
        const float pos[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
        for (int c = 0; c < C; c++) {
            float nu[3];
            gradient_lds<T::SX, T::SY>(w, sf + (c + 1) * T::NS, tx, ty, tz, nu);
            float al;
            {
                // PREALPHA written at c*P+i and read at c*N+i (lenia.cpp:660, 708)
                const int r = c * S::N + (int)i;
                int cs = c, is = (int)i;
                bool ok = true;
                if (qa) { cs = r / (int)S::P; is = r - cs * (int)S::P; ok = cs < C && is < S::N; }
                if (ok) {
                    float pg = A[cs * S::N + is] / w.theta_a;
                    pg = pg * pg;
                    al = clamp_ref(pg, 0.0f, 1.0f);
                } else {
                    al = 0.0f;
                }
            }
            const float al1 = 1.0f - al;
            float F[3];
#pragma unroll
            for (int d = 0; d < 3; d++) F[d] = nu[d] * al1 - na[d] * al;
            float *V = a.pool + w.V + (long long)c * 3 * S::N;
            float v[3] = {0.0f, 0.0f, 0.0f};
            if (w.v_live) {
#pragma unroll
                for (int d = 0; d < 3; d++) v[d] = V[(long long)d * S::N + i];
            }
            if (w.mu_vel) {
#pragma unroll
                for (int d = 0; d < 3; d++) v[d] = (v[d] + F[d]) * damp;
            }
            float *MU = a.pool + w.MU + (long long)c * 3 * S::N;
            if (w.mode == 2) {
                float vn[3];
#pragma unroll
                for (int d = 0; d < 3; d++) vn[d] = w.mu_vel ? v[d] : (v[d] + F[d]) * damp;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const float src = w.adv_vel ? vn[d] : F[d];
                    MU[(long long)d * S::N + i] = fmin_ref(fmax_ref(src * dt, nma), ma);
                }
            } else {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    float m = (w.mu_vel ? v[d] : F[d]) * dt;
                    m = fmax_ref(m, nma);
                    m = fmin_ref(m, ma);
                    m = pos[d] + m;
                    if (wall) { m = fmax_ref(m, sig); m = fmin_ref(m, lim[d]); }
                    MU[(long long)d * S::N + i] = m;
                }
            }
            if (!w.mu_vel) {
#pragma unroll
                for (int d = 0; d < 3; d++) v[d] = (v[d] + F[d]) * damp;
            }
            if (w.v_live) {
#pragma unroll
                for (int d = 0; d < 3; d++) V[(long long)d * S::N + i] = v[d];
            }
            if (w.need_f) {
                float *Fo = a.pool + w.Fv + (long long)c * 3 * S::N;
#pragma unroll
                for (int d = 0; d < 3; d++) Fo[(long long)d * S::N + i] = F[d];
            }
        }
    }
}

// Per-cell transport for the modes the tiled RT path does not cover (SRT, ADV, RT with
// dd > 1); the reference gather of rt_border_* / srt / adv_* for one cell.
template <int W, int H, int D>
__device__ __forceinline__ void transport_cell(const PhaseArgs &a, const DevWorld &w, int x, int y, int z,
                                               float (&tt)[LHD_MAX_C]) {
    using S = Shape<W, H, D>;
    const long long i = x + (long long)W * (y + (long long)H * z);
    const int C = w.C;
    const float *Acur = A_cur(a, w);
    const float *ARR = (w.growth_type == 0 || w.growth_type == 2) ? Acur : a.pool + w.Bk;
    const float *MU = a.pool + w.MU;
    const int faces = (w.quirks & LHQ_FACE_HALO) != 0;
    const int tkind = (w.quirks & LHQ_TRANSPORT_TORUS) ? 0 : w.border;
    const int nrt = (w.quirks & LHQ_RT_3CH) ? (C < 3 ? C : 3) : C;
    const float pos[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
    if (w.mode == 0) {
#pragma unroll
        for (int c = 0; c < LHD_MAX_C; c++) {
            if (c >= nrt) break;
            const float *mu = MU + (long long)c * 3 * S::N;
            float m[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                float dv = pos[d] - mu[(long long)d * S::N + i];
                dv = dv < 0.0f ? -dv : dv;
                dv = w.si - dv;
                m[d] = fmin_ref(fmax_ref(dv, 0.0f), w.clip);
            }
            float wt = ARR[(long long)c * S::N + i] * m[0];
            wt = wt * m[1];
            wt = wt * m[2];
            wt = wt * w.sn;
            tt[c] = tt[c] + wt;
        }
    } else if (w.mode == 1) {
        const bool torus = w.border == 0;
        const float ext[3] = {(float)W, (float)H, (float)D};
        const float half[3] = {(float)(W / 2), (float)(H / 2), (float)(D / 2)};
        const int dd = w.dd;
        for (int dz = -dd; dz <= dd; dz++)
            for (int dy = -dd; dy <= dd; dy++)
                for (int dx = -dd; dx <= dd; dx++) {
                    int sx = x + dx, sy = y + dy, sz = z + dz;
                    if (!halo_src<W, H, D>(sx, sy, sz, tkind, faces)) continue;
                    const long long s = sx + W * (sy + H * sz);
#pragma unroll
                    for (int c = 0; c < LHD_MAX_C; c++) {
                        if (c >= nrt) break;
                        const float *mu = MU + (long long)c * 3 * S::N;
                        float m[3];
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            float dv = pos[d] - mu[(long long)d * S::N + s];
                            if (torus) {
                                if (dv > half[d]) dv = dv - ext[d];
                                if (dv < -half[d]) dv = dv + ext[d];
                            }
                            dv = dv < 0.0f ? -dv : dv;
                            dv = w.si - dv;
                            dv = fmax_ref(dv, 0.0f);
                            m[d] = fmin_ref(dv, w.clip);
                        }
                        float wt = ARR[(long long)c * S::N + s] * m[0];
                        wt = wt * m[1];
                        wt = wt * m[2];
                        wt = wt * w.sn;
                        tt[c] = tt[c] + wt;
                    }
                }
    } else {
        const float ext[3] = {(float)W, (float)H, (float)D};
        const bool wrap = (w.quirks & LHQ_TRANSPORT_TORUS) || w.border == 0;
#pragma unroll
        for (int c = 0; c < LHD_MAX_C; c++) {
            if (c >= C) break;
            const float *dl = MU + (long long)c * 3 * S::N;
            int I0[3];
            float F0[3], F1[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                float p = (pos[d] - 0.5f) - dl[(long long)d * S::N + i];
                if (wrap) {
                    p = p + ext[d];
                    p = p - floorf(p / ext[d]) * ext[d];
                }
                const float fl = floorf(p);
                I0[d] = (int)fl;
                F1[d] = p - fl;
                F0[d] = 1.0f - F1[d];
            }
            const float *Ac = ARR + (long long)c * S::N;
            float acc = 0.0f;
#pragma unroll
            for (int corner = 0; corner < 8; corner++) {
                const int bx = corner & 1, by = (corner >> 1) & 1, bz = (corner >> 2) & 1;
                int cx = I0[0] + bx, cy = I0[1] + by, cz = I0[2] + bz;
                float v = 0.0f;
                if (halo_src<W, H, D>(cx, cy, cz, tkind, faces)) v = Ac[cx + W * (cy + H * cz)];
                float t = v * (bx ? F1[0] : F0[0]);
                t = t * (by ? F1[1] : F0[1]);
                t = t * (bz ? F1[2] : F0[2]);
                acc = corner == 0 ? t : acc + t;
            }
            tt[c] = acc;
        }
    }
}

// End-of-step TRM for one cell (lenia.cpp:930-1035).
template <int W, int H, int D>
__device__ __forceinline__ void trm_cell(const PhaseArgs &a, const DevWorld &w, long long i, const float (&tt)[LHD_MAX_C]) {
    using S = Shape<W, H, D>;
    const int C = w.C;
    const bool clamp01 = (w.quirks & LHQ_CLAMP01) != 0;
    const float *Acur = A_cur(a, w);
    float *An = A_next(a, w), *GA = a.pool + w.GA, *Sg = a.pool + w.S;
    const float *Hg = a.pool + w.Hg;
    float ttsum = 0.0f;
#pragma unroll
    for (int c = 0; c < LHD_MAX_C; c++) {
        if (c >= C) break;
        const long long g = (long long)c * S::N + i;
        float t = tt[c];
        if (w.growth_type == 2) {
            const bool soft = (w.soft_mask >> c) & 1;
            if (soft && (w.quirks & LHQ_SOFTCLIP_TR)) a.pool[w.Bk + g] = soft_growth(Acur[g], Hg[g], w.dt);
            else if (soft) t = soft_growth(t, Hg[g], w.dt);
            else t = clamp01 ? clamp_ref(t + Hg[g] * w.dt, 0.0f, 1.0f) : t + Hg[g] * w.dt;
        } else if (clamp01) {
            t = clamp_ref(t, 0.0f, 1.0f);
        }
        ttsum = ttsum + t;
        GA[g] = w.gtam * GA[g] + t * w.gtas;
        An[g] = t;
    }
    const float sold = (w.quirks & LHQ_SUM_RESET) ? 0.0f : Sg[i];
    Sg[i] = w.stam * sold + ttsum * w.stas;
}

// RT with dd = 1 from LDS: each staged source carries its nine per-axis overlap weights
// (toward the destinations at offsets -1, 0, +1 along x, y and z), computed once with
// exactly the expression the gather would evaluate for that (destination, source)
// pair, so each destination's 27 products and their summation order are unchanged.
// A slot's halo source is resolved once and all its global loads (mu, source mass for
// every channel) are issued before any weight is formed.
template <int W, int H, int D>
__global__ __launch_bounds__(256) void k_transport_tiled(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    using T = TileCfg<W, H, D>;
    constexpr int SPT = (T::NS + T::THREADS - 1) / T::THREADS;  // staged slots per thread
    __shared__ float s_arr[T::NS];
    __shared__ float s_w[9 * T::NS];
    const int wi = blockIdx.x / T::TILES, tile = blockIdx.x % T::TILES;
    const int ty0 = (tile % (H / T::TY)) * T::TY, tz0 = (tile / (H / T::TY)) * T::TZ;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const int tid = threadIdx.x, C = w.C;
    float tt[T::CPT][LHD_MAX_C];
#pragma unroll
    for (int j = 0; j < T::CPT; j++)
#pragma unroll
        for (int c = 0; c < LHD_MAX_C; c++) tt[j][c] = 0.0f;
    if (w.mode == 1 && w.dd == 1) {
        const float *ARR = (w.growth_type == 0 || w.growth_type == 2) ? A_cur(a, w) : a.pool + w.Bk;
        const float *MU = a.pool + w.MU;
        const int faces = (w.quirks & LHQ_FACE_HALO) != 0;
        const int tkind = (w.quirks & LHQ_TRANSPORT_TORUS) ? 0 : w.border;
        const int nrt = (w.quirks & LHQ_RT_3CH) ? (C < 3 ? C : 3) : C;
        // rt_border_torus wraps the displacement by +-extent past half the extent;
        // rt_border_other does not (an infinite threshold never triggers)
        const bool torus = w.border == 0;
        const float ext[3] = {(float)W, (float)H, (float)D};
        const float half[3] = {torus ? (float)(W / 2) : __builtin_inff(), torus ? (float)(H / 2) : __builtin_inff(),
                               torus ? (float)(D / 2) : __builtin_inff()};
        const float si = w.si, clip = w.clip, sn = w.sn;
        // gather every staged slot's inputs for all channels up front
        float mu_r[SPT][LHD_MAX_C][3], arr_r[SPT][LHD_MAX_C];
        int gpos[SPT][3];
#pragma unroll
        for (int j = 0; j < SPT; j++) {
            const int q = tid + j * T::THREADS;
            const int sx = q % T::SX, sy = (q / T::SX) % T::SY, sz = q / (T::SX * T::SY);
            gpos[j][0] = sx - 1;
            gpos[j][1] = ty0 + sy - 1;
            gpos[j][2] = tz0 + sz - 1;
            int x = gpos[j][0], y = gpos[j][1], z = gpos[j][2];
            const bool ok = q < T::NS && halo_src<W, H, D>(x, y, z, tkind, faces);
            const int s = x + W * (y + H * z);
#pragma unroll
            for (int c = 0; c < LHD_MAX_C; c++) {
                arr_r[j][c] = (ok && c < nrt) ? ARR[c * S::N + s] : 0.0f;
#pragma unroll
                for (int d = 0; d < 3; d++) mu_r[j][c][d] = (ok && c < nrt) ? MU[(c * 3 + d) * S::N + s] : 0.0f;
            }
        }
#pragma unroll
        for (int c = 0; c < LHD_MAX_C; c++) {
            if (c >= nrt) break;
            __syncthreads();
#pragma unroll
            for (int j = 0; j < SPT; j++) {
                const int q = tid + j * T::THREADS;
                if (q >= T::NS) break;
                s_arr[q] = arr_r[j][c];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const float mu = mu_r[j][c][d];
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        const float pos = (float)(gpos[j][d] - (k - 1)) + 0.5f;
                        float dv = pos - mu;
                        // the two one-sided wraps of rt_border_torus (reintegration.cpp:733-755):
                        // after dv -= ext for dv > half, dv > -half, so at most one fires and
                        // dv - (+-ext) equals the reference's dv - ext / dv + ext exactly
                        const float wr = dv - __builtin_copysignf(ext[d], dv);
                        dv = __builtin_fabsf(dv) > half[d] ? wr : dv;
                        // abs_ps clears the sign bit; si - |dv| is never -0 or NaN here, so
                        // v_max / v_min give the same bits as _mm_max_ps / _mm_min_ps
                        dv = si - __builtin_fabsf(dv);
                        s_w[(d * 3 + k) * T::NS + q] = __builtin_fminf(__builtin_fmaxf(dv, 0.0f), clip);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < T::CPT; j++) {
                const int cell = tid + j * T::THREADS;
                if (cell >= T::NT) break;
                const int tx = cell % T::TX, ty = (cell / T::TX) % T::TY, tz = cell / (T::TX * T::TY);
                float acc = tt[j][c];
#pragma unroll
                for (int dz = -1; dz <= 1; dz++)
#pragma unroll
                    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
                        for (int dx = -1; dx <= 1; dx++) {
                            const int q = T::slot(tx + dx + 1, ty + dy + 1, tz + dz + 1);
                            float wt = s_arr[q] * s_w[(dx + 1) * T::NS + q];
                            wt = wt * s_w[(3 + dy + 1) * T::NS + q];
                            wt = wt * s_w[(6 + dz + 1) * T::NS + q];
                            wt = wt * sn;
                            acc = acc + wt;
                        }
                tt[j][c] = acc;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < T::CPT; j++) {
            const int cell = tid + j * T::THREADS;
            if (cell >= T::NT) break;
            const int tx = cell % T::TX, ty = (cell / T::TX) % T::TY, tz = cell / (T::TX * T::TY);
            transport_cell<W, H, D>(a, w, tx, ty0 + ty, tz0 + tz, tt[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < T::CPT; j++) {
        const int cell = tid + j * T::THREADS;
        if (cell >= T::NT) break;
        const int tx = cell % T::TX, ty = (cell / T::TX) % T::TY, tz = cell / (T::TX * T::TY);
        trm_cell<W, H, D>(a, w, tx + (long long)W * ((ty0 + ty) + (long long)H * (tz0 + tz)), tt[j]);
    }
}

// ------------------------------------------------------------------------------------
// Shape table.
// ------------------------------------------------------------------------------------
template <int W, int H, int D>
struct Launch {
    // arg = source | (max fields per world) << 8
    static void fwd_xy(const PhaseArgs &a, int arg, void *st) {
        const int grid = a.nworlds * (arg >> 8) * (D / FwdCfg<W, H, D>::NP);
        hipLaunchKernelGGL((k_fwd_xy<W, H, D>), dim3(grid), dim3(128), 0, (hipStream_t)st, a, arg);
    }
    static void spec_z(const PhaseArgs &a, int mode, void *st) {
        constexpr int BPW = ((W / 2 + 1) * H + 127) / 128;
        hipLaunchKernelGGL((k_spec_z<W, H, D>), dim3(a.nworlds * BPW), dim3(128), 0, (hipStream_t)st, a, mode);
    }
    static void kspec_z(const PhaseArgs &a, int, void *st) {
        constexpr int BPW = ((W / 2 + 1) * H + 127) / 128;
        hipLaunchKernelGGL((k_kspec_z<W, H, D>), dim3(a.nworlds * BPW), dim3(128), 0, (hipStream_t)st, a, 0);
    }
    static void inv_growth(const PhaseArgs &a, int, void *st) {
        hipLaunchKernelGGL((k_inv_growth<W, H, D>), dim3(a.nworlds * D), dim3(256), 0, (hipStream_t)st, a, 0);
    }
    static void inv_blur(const PhaseArgs &a, int, void *st) {
        hipLaunchKernelGGL((k_inv_blur<W, H, D>), dim3(a.nworlds * D), dim3(256), 0, (hipStream_t)st, a, 0);
    }
    static void local(const PhaseArgs &a, int, void *st) {
        hipLaunchKernelGGL((k_local_tiled<W, H, D>), dim3(a.nworlds * TileCfg<W, H, D>::TILES), dim3(256), 0,
                           (hipStream_t)st, a, 0);
    }
    static void transport(const PhaseArgs &a, int, void *st) {
        hipLaunchKernelGGL((k_transport_tiled<W, H, D>), dim3(a.nworlds * TileCfg<W, H, D>::TILES), dim3(256), 0,
                           (hipStream_t)st, a, 0);
    }
    static constexpr ShapeOps ops() {
        return ShapeOps{W, H, D, &fwd_xy, &spec_z, &inv_growth, &inv_blur, &local, &transport, &kspec_z};
    }
};

#define LH_SHAPES(X) \
    X(8, 8, 8)       \
    X(8, 4, 8)       \
    X(12, 12, 12)    \
    X(16, 16, 16)    \
    X(16, 8, 16)     \
    X(24, 24, 24)    \
    X(32, 32, 32)    \
    X(36, 36, 36)    \
    X(40, 40, 40)    \
    X(44, 24, 44)    \
    X(48, 48, 48)

#define LH_OPS_ENTRY(w, h, d) Launch<w, h, d>::ops(),
static const ShapeOps g_ops[] = {LH_SHAPES(LH_OPS_ENTRY)};

const ShapeOps *lh_shape_ops(int W, int H, int D) {
    for (const ShapeOps &o : g_ops)
        if (o.W == W && o.H == H && o.D == D) return &o;
    return nullptr;
}

// ------------------------------------------------------------------------------------
// Shape-independent kernels: FP64 mass per world, marching-cubes case index.
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mass(PhaseArgs a, double *out) {
    const DevWorld &w = a.worlds[blockIdx.x];
    const long long n = (long long)w.C * w.W * w.H * w.D;
    const float *A = a.pool + (a.parity ? w.A2 : w.A);
    double s = 0.0;
    for (long long i = threadIdx.x; i < n; i += 256) s += (double)A[i];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Case index of cell (x,y,z): bit i set iff the value at corner i (mesh.cpp:19-32)
// is > iso (mesh.cpp:173-180); corners past the far faces read the torus face pad,
// edge/corner pads read 0 when face_only (transposition.cpp:881-995).
__global__ __launch_bounds__(256) void k_mc_case(const float *f, int W, int H, int D, float iso,
                                                 int face_only, unsigned char *out) {
    const long long n = (long long)W * H * D;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)(i / ((long long)W * H));
    const int DX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    const int DY[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    const int DZ[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    int v = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int cx = x + DX[k], cy = y + DY[k], cz = z + DZ[k];
        const int m = (cx >= W) + (cy >= H) + (cz >= D);
        float val = 0.0f;
        if (m <= 1 || !face_only) {
            cx = cx >= W ? cx - W : cx;
            cy = cy >= H ? cy - H : cy;
            cz = cz >= D ? cz - D : cz;
            val = f[cx + (long long)W * (cy + (long long)H * cz)];
        }
        if (val > iso) v |= 1 << k;
    }
    out[i] = (unsigned char)v;
}

__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) dst[i] = src[i];
}

void lh_launch_copy(const void *src, void *dst, size_t bytes, void *st) {
    const long long n = (long long)(bytes / 16);
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, (hipStream_t)st, (const float4 *)src, (float4 *)dst, n);
}

void lh_launch_mass(const PhaseArgs &a, int nworlds, double *out, void *st) {
    hipLaunchKernelGGL(k_mass, dim3(nworlds), dim3(256), 0, (hipStream_t)st, a, out);
}

void lh_launch_mc(const float *f, int W, int H, int D, float iso, int face_only, unsigned char *out, void *st) {
    const long long n = (long long)W * H * D;
    hipLaunchKernelGGL(k_mc_case, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)st, f, W, H, D,
                       iso, face_only, out);
}
