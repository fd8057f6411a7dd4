// lh_dftgemm.hip -- the growth convolution as DFT-as-GEMM on CDNA4 matrix cores.
//
// U_k = Re IDFT(K^_k . DFT(A_c0)) / N is evaluated as six passes of dense matrix
// products on v_mfma_f32_16x16x32_{bf16,f16}: every 1-D transform of length n is an
// n x n DFT matrix times a tile of 32 lines, with fp32 accumulation.  Operands are
// 16-bit; in the split variants each fp32 operand x is carried as hi = t(x),
// lo = t((x - hi) * 2^S) and every product as hi.hi + 2^-S (hi.lo + lo.hi), the two
// parts in separate fp32 accumulators, so the result keeps about 16 (bf16) or 22
// (fp16) significant bits; the 2^S keeps fp16 lo parts out of the subnormal range.
// Matrices are unscaled (|entries| <= 2); each forward pass multiplies its fp32 result
// by 1/n, so every stored value stays within max|A| (inside fp16 range) and the inverse
// passes, unscaled, return U exactly normalised.
//
//   fx   real rows of A_c (x)            -> half spectrum rows   (n x n matrix, packed)
//   fy   half spectrum, along y           (complex, in place)
//   fz   half spectrum, along z           (complex, in place)
//   iz   (half spectrum of c0) . K^_k     -> per-kernel spectrum along z (inverse)
//   iy   per-kernel spectrum along y      (inverse, in place)
//   ix   per-kernel rows -> u (c2r, packed n x n), growth accumulated in kernel order
//        (mt_fft_step lenia.cpp:501-538), H, B, U_SUM written once per cell.
//
// Half-spectrum row layout (n+2 floats per (z,y) row): re[0..n/2], im[0..n/2].  The
// real transforms use the packed index r in [0,n): r <= n/2 -> re[r], r > n/2 ->
// im[r-n/2], i.e. float offset r + (r > n/2); im[0] and im[n/2] are zero for real data.
//
// Matrices are built on the host in double, split once into (hi, lo), and stored in
// fragment order ([row block][k step][lane][8]) so that each lane fetches its operand
// with one 16-byte load; a wave keeps its rows of the matrix in registers for the
// whole pass.  Data tiles are split while staged into LDS as [line][j] (pitch n+8
// halves: 16 lanes reading 16 lines hit distinct banks).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "lh_common.hpp"
#include "lenia_hip.h"
#include "lh_device.hpp"

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

template <typename T>
struct MT;
template <>
struct MT<__bf16> {
    static __device__ __forceinline__ f32x4 mac(u16x8 a, u16x8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       c, 0, 0, 0);
    }
    static __device__ __forceinline__ unsigned short cvt(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
    static __device__ __forceinline__ float back(unsigned short h) { return (float)__builtin_bit_cast(__bf16, h); }
    static constexpr float LO = 256.0f, UNLO = 1.0f / 256.0f;  // lo-part scale 2^S
};
template <>
struct MT<_Float16> {
    static __device__ __forceinline__ f32x4 mac(u16x8 a, u16x8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                      0, 0, 0);
    }
    static __device__ __forceinline__ unsigned short cvt(float x) {
        return __builtin_bit_cast(unsigned short, (_Float16)x);
    }
    static __device__ __forceinline__ float back(unsigned short h) { return (float)__builtin_bit_cast(_Float16, h); }
    static constexpr float LO = 2048.0f, UNLO = 1.0f / 2048.0f;
};

template <int N>
struct GC {
    static constexpr int RB = 1;                 // 16-row blocks of the matrix per wave
    static constexpr int WAVES = N / (16 * RB);
    static constexpr int THREADS = 64 * WAVES;
    static constexpr int KS = N / 32;            // k steps of 32
    static constexpr int L = 32;                 // lines per tile
    static constexpr int LB = 2;                 // 16-line blocks per tile
    static constexpr int LDK = N + 8;            // LDS pitch per line (halves)
    static constexpr int NB = N / 2 + 1;         // half-spectrum bins
    static constexpr int RS = 2 * NB;            // floats per spectrum row
    static constexpr int NN = N * N;
    static constexpr long long N3 = (long long)N * N * N;
    static constexpr long long SPEC = (long long)N * N * RS;
    static constexpr int CTILES = N * NB / L;    // complex line tiles per field
    static constexpr int RTILES = N * N / L;     // real row tiles per field
    static constexpr int ITER = 4 * N / THREADS; // staging rounds (32 lines x N/8 chunks)
    static_assert(N % 32 == 0 && (N * NB) % L == 0 && (4 * N) % THREADS == 0, "tile geometry");
};

// matrix slots in the per-size block (each N*N halves, fragment order)
enum { M_FR = 0, M_FI = 2, M_GR = 4, M_GI = 6, M_XF = 8, M_XI = 10, M_COUNT = 12 };

__device__ __forceinline__ u16x8 ldfrag(const unsigned short *m, int idx) {
    return *reinterpret_cast<const u16x8 *>(m + (size_t)idx * 8);
}
__device__ __forceinline__ u16x8 lds8(const unsigned short *s) { return *reinterpret_cast<const u16x8 *>(s); }
__device__ __forceinline__ void sts8(unsigned short *s, u16x8 v) { *reinterpret_cast<u16x8 *>(s) = v; }

template <typename T>
__device__ __forceinline__ void split8(const float (&v)[8], u16x8 &hi, u16x8 &lo) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const unsigned short h = MT<T>::cvt(v[i]);
        hi[i] = h;
        lo[i] = MT<T>::cvt((v[i] - MT<T>::back(h)) * MT<T>::LO);
    }
}

__device__ __forceinline__ int fold(int q, int n) { return q <= n / 2 ? q : n - q; }

// Loads a wave's rows (RB blocks) of a matrix pair (hi, lo) into registers.
template <int N, bool SPLIT>
__device__ __forceinline__ void load_matrix(const unsigned short *mats, int slot, int rb0, int lane,
                                            u16x8 (&h)[GC<N>::RB][GC<N>::KS], u16x8 (&l)[GC<N>::RB][GC<N>::KS]) {
    using G = GC<N>;
#pragma unroll
    for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
        for (int ks = 0; ks < G::KS; ks++) {
            const int idx = ((rb0 + rb) * G::KS + ks) * 64 + lane;
            h[rb][ks] = ldfrag(mats + (size_t)slot * G::NN, idx);
            if (SPLIT) l[rb][ks] = ldfrag(mats + (size_t)(slot + 1) * G::NN, idx);
        }
}

// Real GEMM over a staged tile: c[rb][lb] = scale * (M . X) for the wave's rows.
template <int N, typename T, bool SPLIT>
__device__ __forceinline__ void gemm_real(const unsigned short *shi, const unsigned short *slo, int lane,
                                          const u16x8 (&mh)[GC<N>::RB][GC<N>::KS],
                                          const u16x8 (&ml)[GC<N>::RB][GC<N>::KS], float scale,
                                          f32x4 (&c)[GC<N>::RB][GC<N>::LB]) {
    using G = GC<N>;
    f32x4 x[G::RB][G::LB];
#pragma unroll
    for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
        for (int lb = 0; lb < G::LB; lb++) {
            c[rb][lb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            x[rb][lb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
    for (int ks = 0; ks < G::KS; ks++)
#pragma unroll
        for (int lb = 0; lb < G::LB; lb++) {
            const int off = (lb * 16 + (lane & 15)) * G::LDK + ks * 32 + 8 * (lane >> 4);
            const u16x8 bh = lds8(shi + off);
            u16x8 bl;
            if (SPLIT) bl = lds8(slo + off);
#pragma unroll
            for (int rb = 0; rb < G::RB; rb++) {
                c[rb][lb] = MT<T>::mac(mh[rb][ks], bh, c[rb][lb]);
                if (SPLIT) {
                    x[rb][lb] = MT<T>::mac(mh[rb][ks], bl, x[rb][lb]);
                    x[rb][lb] = MT<T>::mac(ml[rb][ks], bh, x[rb][lb]);
                }
            }
        }
#pragma unroll
    for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
        for (int lb = 0; lb < G::LB; lb++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                c[rb][lb][i] = (SPLIT ? c[rb][lb][i] + x[rb][lb][i] * MT<T>::UNLO : c[rb][lb][i]) * scale;
}

// Complex transform along y (AXIS 1) or z (AXIS 2) of 32-line tiles of a half spectrum.
// IO 0: forward, in place on channel f of Fsp.  IO 1: inverse of (Fsp[c0] . K^_f) into
// Gsp[f].  IO 2: inverse, in place on kernel f of Gsp.  arg = fields per world.
// A persistent grid: each workgroup loads its rows of the DFT matrices once and takes
// tiles blockIdx.x, + gridDim.x, ...; the next tile's elements are loaded into
// registers while the current tile's products run.  Tiles are disjoint sets of lines,
// so an in-place pass never reads a line another tile writes.
template <int N, typename T, bool SPLIT, int AXIS, int IO>
__global__ __launch_bounds__(GC<N>::THREADS) void k_dft_c2c(PhaseArgs a, int nf) {
    using G = GC<N>;
    __shared__ __attribute__((aligned(16))) unsigned short sb[4][G::L * G::LDK];
    const int total = a.nworlds * nf * G::CTILES;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // tile id -> (world, field, tile); false for fields past the world's count
    auto decode = [&](int id, const DevWorld *&w, int &f, int &tile) {
        tile = id % G::CTILES;
        f = (id / G::CTILES) % nf;
        w = &a.worlds[a.ids[id / (G::CTILES * nf)]];
        return f < (IO == 0 ? w->C : w->K);
    };
    auto next_valid = [&](int id) {
        const DevWorld *w;
        int f, tile;
        while (id < total && !decode(id, w, f, tile)) id += gridDim.x;
        return id;
    };
    // staged elements of this thread: (line l, chunk q of 8 consecutive j) per round
    float pre_re[G::ITER][8], pre_im[G::ITER][8], pre_m[G::ITER][8];
    auto fetch = [&](int id) {
        const DevWorld *w;
        int f, tile;
        decode(id, w, f, tile);
        const float *src = a.pool + (IO == 1 ? w->Fsp + w->kern[f].c0 * G::SPEC : (IO == 0 ? w->Fsp : w->Gsp) + f * G::SPEC);
        const float *kh = a.pool + w->KH + (long long)f * G::NB * G::NB * G::NB;
#pragma unroll
        for (int it = 0; it < G::ITER; it++) {
            const int l = tid & 31, q = (tid >> 5) + it * (G::THREADS / 32);
            const int line = tile * G::L + l, outer = line / G::NB, kx = line % G::NB;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int j = q * 8 + i;
                const long long row = AXIS == 1 ? (long long)outer * N + j : (long long)j * N + outer;
                pre_re[it][i] = src[row * G::RS + kx];
                pre_im[it][i] = src[row * G::RS + G::NB + kx];
                if (IO == 1) pre_m[it][i] = kh[((long long)fold(j, N) * G::NB + fold(outer, N)) * G::NB + kx];
            }
        }
    };
    int id = next_valid(blockIdx.x);
    if (id >= total) return;
    u16x8 rh[G::RB][G::KS], rl[G::RB][G::KS], ih[G::RB][G::KS], il[G::RB][G::KS];
    const int rb0 = wave * G::RB;
    load_matrix<N, SPLIT>(a.dft, IO == 0 ? M_FR : M_GR, rb0, lane, rh, rl);
    load_matrix<N, SPLIT>(a.dft, IO == 0 ? M_FI : M_GI, rb0, lane, ih, il);
    fetch(id);
    for (;;) {
        lds_barrier();  // the previous tile's products have read the staged tile (LDS only:
                        // the next tile's loads stay in flight, tiles are disjoint lines)
#pragma unroll
        for (int it = 0; it < G::ITER; it++) {
            const int l = tid & 31, q = (tid >> 5) + it * (G::THREADS / 32);
            float re[8], im[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                re[i] = IO == 1 ? pre_re[it][i] * pre_m[it][i] : pre_re[it][i];
                im[i] = IO == 1 ? pre_im[it][i] * pre_m[it][i] : pre_im[it][i];
            }
            u16x8 h, lo;
            split8<T>(re, h, lo);
            sts8(&sb[0][l * G::LDK + q * 8], h);
            if (SPLIT) sts8(&sb[1][l * G::LDK + q * 8], lo);
            split8<T>(im, h, lo);
            sts8(&sb[2][l * G::LDK + q * 8], h);
            if (SPLIT) sts8(&sb[3][l * G::LDK + q * 8], lo);
        }
        const int cur = id;
        id = next_valid(id + gridDim.x);
        if (id < total) fetch(id);  // in flight during this tile's products
        lds_barrier();
        f32x4 cr[G::RB][G::LB], ci[G::RB][G::LB], xr[G::RB][G::LB], xi[G::RB][G::LB];
#pragma unroll
        for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
            for (int lb = 0; lb < G::LB; lb++) {
                cr[rb][lb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                ci[rb][lb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                xr[rb][lb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                xi[rb][lb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            }
#pragma unroll
        for (int ks = 0; ks < G::KS; ks++)
#pragma unroll
            for (int lb = 0; lb < G::LB; lb++) {
                const int off = (lb * 16 + (lane & 15)) * G::LDK + ks * 32 + 8 * (lane >> 4);
                const u16x8 xrh = lds8(&sb[0][off]), xih = lds8(&sb[2][off]);
                const u16x8 nih = xih ^ (unsigned short)0x8000;
                u16x8 xrl, xil, nil;
                if (SPLIT) {
                    xrl = lds8(&sb[1][off]);
                    xil = lds8(&sb[3][off]);
                    nil = xil ^ (unsigned short)0x8000;
                }
#pragma unroll
                for (int rb = 0; rb < G::RB; rb++) {
                    // (Mr + i Mi)(Xr + i Xi): re = Mr Xr - Mi Xi, im = Mr Xi + Mi Xr
                    cr[rb][lb] = MT<T>::mac(rh[rb][ks], xrh, cr[rb][lb]);
                    cr[rb][lb] = MT<T>::mac(ih[rb][ks], nih, cr[rb][lb]);
                    ci[rb][lb] = MT<T>::mac(rh[rb][ks], xih, ci[rb][lb]);
                    ci[rb][lb] = MT<T>::mac(ih[rb][ks], xrh, ci[rb][lb]);
                    if (SPLIT) {
                        xr[rb][lb] = MT<T>::mac(rh[rb][ks], xrl, xr[rb][lb]);
                        xr[rb][lb] = MT<T>::mac(rl[rb][ks], xrh, xr[rb][lb]);
                        xr[rb][lb] = MT<T>::mac(ih[rb][ks], nil, xr[rb][lb]);
                        xr[rb][lb] = MT<T>::mac(il[rb][ks], nih, xr[rb][lb]);
                        xi[rb][lb] = MT<T>::mac(rh[rb][ks], xil, xi[rb][lb]);
                        xi[rb][lb] = MT<T>::mac(rl[rb][ks], xih, xi[rb][lb]);
                        xi[rb][lb] = MT<T>::mac(ih[rb][ks], xrl, xi[rb][lb]);
                        xi[rb][lb] = MT<T>::mac(il[rb][ks], xrh, xi[rb][lb]);
                    }
                }
            }
        const DevWorld *w;
        int f, tile;
        decode(cur, w, f, tile);
        float *dst = a.pool + (IO == 0 ? w->Fsp : w->Gsp) + f * G::SPEC;
        const float scale = IO == 0 ? 1.0f / N : 1.0f;
#pragma unroll
        for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
            for (int lb = 0; lb < G::LB; lb++) {
                const int line = tile * G::L + lb * 16 + (lane & 15), outer = line / G::NB, kx = line % G::NB;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int k = (rb0 + rb) * 16 + 4 * (lane >> 4) + i;
                    const long long row = AXIS == 1 ? (long long)outer * N + k : (long long)k * N + outer;
                    const float re = SPLIT ? cr[rb][lb][i] + xr[rb][lb][i] * MT<T>::UNLO : cr[rb][lb][i];
                    const float im = SPLIT ? ci[rb][lb][i] + xi[rb][lb][i] * MT<T>::UNLO : ci[rb][lb][i];
                    dst[row * G::RS + kx] = re * scale;
                    dst[row * G::RS + G::NB + kx] = im * scale;
                }
            }
        if (id >= total) break;
    }
}

// Forward real transform along x of 32 rows of A_c into half-spectrum rows of Fsp[c].
template <int N, typename T, bool SPLIT>
__global__ __launch_bounds__(GC<N>::THREADS) void k_dft_r2c_x(PhaseArgs a, int nf) {
    using G = GC<N>;
    __shared__ __attribute__((aligned(16))) unsigned short sb[2][G::L * G::LDK];
    const int tile = blockIdx.x % G::RTILES;
    const int c = (blockIdx.x / G::RTILES) % nf;
    const int wi = blockIdx.x / (G::RTILES * nf);
    const DevWorld &w = a.worlds[a.ids[wi]];
    if (c >= w.C) return;
    const float *src = A_cur(a, w) + c * G::N3;
    float *dst = a.pool + w.Fsp + c * G::SPEC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int QC = N / 8;
#pragma unroll
    for (int it = 0; it < G::ITER; it++) {
        const int q = tid % QC, l = tid / QC + it * (G::THREADS / QC);
        const float4 *p = reinterpret_cast<const float4 *>(src + (long long)(tile * G::L + l) * N + q * 8);
        const float4 v0 = p[0], v1 = p[1];
        const float v[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        u16x8 h, lo;
        split8<T>(v, h, lo);
        sts8(&sb[0][l * G::LDK + q * 8], h);
        if (SPLIT) sts8(&sb[1][l * G::LDK + q * 8], lo);
    }
    u16x8 mh[G::RB][G::KS], ml[G::RB][G::KS];
    const int rb0 = wave * G::RB;
    load_matrix<N, SPLIT>(a.dft, M_XF, rb0, lane, mh, ml);
    lds_barrier();
    f32x4 acc[G::RB][G::LB];
    gemm_real<N, T, SPLIT>(sb[0], sb[1], lane, mh, ml, 1.0f / N, acc);
#pragma unroll
    for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
        for (int lb = 0; lb < G::LB; lb++) {
            const long long line = tile * G::L + lb * 16 + (lane & 15);
            float *row = dst + line * G::RS;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = (rb0 + rb) * 16 + 4 * (lane >> 4) + i;
                row[r + (r > N / 2 ? 1 : 0)] = acc[rb][lb][i];
                if (r == 0) row[G::NB] = 0.0f;             // im[0]
                if (r == N / 2) row[G::NB + N / 2] = 0.0f;  // im[n/2]
            }
        }
}

// Inverse real transform along x of every kernel's rows, then the growth of
// mt_fft_step (lenia.cpp:501-538) in kernel order: U_SUM, H (+asymptotic), B.
template <int N, typename T, bool SPLIT>
__global__ __launch_bounds__(GC<N>::THREADS) void k_dft_c2r_growth(PhaseArgs a, int) {
    using G = GC<N>;
    __shared__ __attribute__((aligned(16))) unsigned short sb[2][G::L * G::LDK];
    __shared__ float slut[LHD_MAX_K * 256];
    const int tile = blockIdx.x % G::RTILES;
    const int wi = blockIdx.x / G::RTILES;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int C = w.C, K = w.K;
    const bool lut = (w.quirks & LHQ_GROWTH_LUT) != 0;
    if (lut)
        for (int i = tid; i < K * 256; i += G::THREADS) slut[i] = a.pool[w.LUT + i];
    const float *A = A_cur(a, w);
    float *US = a.pool + w.USUM;
    u16x8 mh[G::RB][G::KS], ml[G::RB][G::KS];
    const int rb0 = wave * G::RB;
    load_matrix<N, SPLIT>(a.dft, M_XI, rb0, lane, mh, ml);
    f32x4 hacc[LHD_MAX_C][G::RB][G::LB];
#pragma unroll
    for (int c = 0; c < LHD_MAX_C; c++)
#pragma unroll
        for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
            for (int lb = 0; lb < G::LB; lb++) hacc[c][rb][lb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // cell of (rb, lb, i): row line = tile*32 + lb*16 + lane%16, x = (rb0+rb)*16 + 4*(lane/16) + i
    auto cell = [&](int rb, int lb, int i) -> long long {
        return (long long)(tile * G::L + lb * 16 + (lane & 15)) * N + (rb0 + rb) * 16 + 4 * (lane >> 4) + i;
    };
    if (w.keep_usum)
        for (int c = 0; c < C; c++)
#pragma unroll
            for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
                for (int lb = 0; lb < G::LB; lb++)
#pragma unroll
                    for (int i = 0; i < 4; i++) US[c * G::N3 + cell(rb, lb, i)] = 0.0f;
    constexpr int QC = N / 8;
    for (int k = 0; k < K; k++) {
        const float *src = a.pool + w.Gsp + k * G::SPEC;
        lds_barrier();  // previous kernel's tile fully consumed (and the LUT staged)
#pragma unroll
        for (int it = 0; it < G::ITER; it++) {
            const int q = tid % QC, l = tid / QC + it * (G::THREADS / QC);
            const float *row = src + (long long)(tile * G::L + l) * G::RS;
            float v[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int r = q * 8 + i;
                v[i] = row[r + (r > N / 2 ? 1 : 0)];
            }
            u16x8 h, lo;
            split8<T>(v, h, lo);
            sts8(&sb[0][l * G::LDK + q * 8], h);
            if (SPLIT) sts8(&sb[1][l * G::LDK + q * 8], lo);
        }
        lds_barrier();
        f32x4 u[G::RB][G::LB];
        gemm_real<N, T, SPLIT>(sb[0], sb[1], lane, mh, ml, 1.0f, u);
        const DevKernel kp = w.kern[k];
        const int c1 = kp.c1;
        const bool asym = (w.asym_mask >> c1) & 1;
#pragma unroll
        for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
            for (int lb = 0; lb < G::LB; lb++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float uv = u[rb][lb][i];
                    float g;
                    if (lut) {
                        int iu = (int)(uv * 255.0f);
                        iu = iu < 0 ? 0 : (iu > 255 ? 255 : iu);
                        g = slut[k * 256 + iu];
                    } else {
                        g = bell_growth(uv, kp.m, kp.s);
                    }
                    const long long ci = cell(rb, lb, i);
                    const float add = asym ? (g - A[c1 * G::N3 + ci]) * kp.h : g * kp.h;
#pragma unroll
                    for (int c = 0; c < LHD_MAX_C; c++)
                        if (c == c1) hacc[c][rb][lb][i] += add;
                    if (w.keep_usum) US[c1 * G::N3 + ci] += uv;
                }
    }
    float *Hg = a.pool + w.Hg, *B = a.pool + w.Bk;
    const bool with_b = w.growth_type == 1;
#pragma unroll
    for (int c = 0; c < LHD_MAX_C; c++) {
        if (c >= C) break;
        const bool soft = (w.soft_mask >> c) & 1;
        // the channel's A values first, all in flight before its stores (a load issued
        // after a store waits for it: loads and stores share one completion counter)
        float av[G::RB][G::LB][4];
        if (with_b) {
#pragma unroll
            for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
                for (int lb = 0; lb < G::LB; lb++)
#pragma unroll
                    for (int i = 0; i < 4; i++) av[rb][lb][i] = A[c * G::N3 + cell(rb, lb, i)];
        }
#pragma unroll
        for (int rb = 0; rb < G::RB; rb++)
#pragma unroll
            for (int lb = 0; lb < G::LB; lb++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const long long gi = c * G::N3 + cell(rb, lb, i);
                    const float hv = hacc[c][rb][lb][i];
                    Hg[gi] = hv;
                    if (with_b)
                        B[gi] = soft ? soft_growth(av[rb][lb][i], hv, w.dt)
                                     : clamp_ref(av[rb][lb][i] + hv * w.dt, 0.0f, 1.0f);
                }
    }
}

// A persistent kernel over `total` tiles: as many workgroups as are resident at once (at
// most one per tile).  Blocks per CU and CUs are queried once per (kernel, block size,
// device) and cached under that key: every k_dft_c2c instantiation has the same function
// type, so the cache cannot live in a per-type static.
inline int persistent_slots(const void *fn, int threads) {
    static std::mutex mu;
    static std::map<std::tuple<const void *, int, int>, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_tuple(fn, threads, dev);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int nb = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, 0) != hipSuccess || nb < 1) nb = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    return cache[key] = nb * cus;
}

template <class K>
void launch_persistent(K *fn, int threads, int total, void *st, const PhaseArgs &a, int arg) {
    const int slots = persistent_slots(reinterpret_cast<const void *>(fn), threads);
    const int g = total < slots ? total : slots;
    if (g > 0) hipLaunchKernelGGL(fn, dim3(g), dim3(threads), 0, (hipStream_t)st, a, arg);
}

template <int N, typename T, bool SPLIT>
struct GLaunch {
    using G = GC<N>;
    static void fx(const PhaseArgs &a, int nf, void *st) {
        hipLaunchKernelGGL((k_dft_r2c_x<N, T, SPLIT>), dim3(a.nworlds * nf * G::RTILES), dim3(G::THREADS), 0,
                           (hipStream_t)st, a, nf);
    }
    static void fy(const PhaseArgs &a, int nf, void *st) {
        launch_persistent(k_dft_c2c<N, T, SPLIT, 1, 0>, G::THREADS, a.nworlds * nf * G::CTILES, st, a, nf);
    }
    static void fz(const PhaseArgs &a, int nf, void *st) {
        launch_persistent(k_dft_c2c<N, T, SPLIT, 2, 0>, G::THREADS, a.nworlds * nf * G::CTILES, st, a, nf);
    }
    static void iz(const PhaseArgs &a, int nf, void *st) {
        launch_persistent(k_dft_c2c<N, T, SPLIT, 2, 1>, G::THREADS, a.nworlds * nf * G::CTILES, st, a, nf);
    }
    static void iy(const PhaseArgs &a, int nf, void *st) {
        launch_persistent(k_dft_c2c<N, T, SPLIT, 1, 2>, G::THREADS, a.nworlds * nf * G::CTILES, st, a, nf);
    }
    static void ix(const PhaseArgs &a, int, void *st) {
        hipLaunchKernelGGL((k_dft_c2r_growth<N, T, SPLIT>), dim3(a.nworlds * G::RTILES), dim3(G::THREADS), 0,
                           (hipStream_t)st, a, 0);
    }
    static constexpr GemmOps ops(int conv) { return GemmOps{N, conv, &fx, &fy, &fz, &iz, &iy, &ix}; }
};

#define LH_GEMM_SIZES(X) X(32) X(64) X(128)
#define LH_GEMM_ENTRY(n)                                              \
    GLaunch<n, __bf16, true>::ops(LH_CONV_GEMM_BF16X3),               \
        GLaunch<n, _Float16, true>::ops(LH_CONV_GEMM_FP16X3),         \
        GLaunch<n, __bf16, false>::ops(LH_CONV_GEMM_BF16),            \
        GLaunch<n, _Float16, false>::ops(LH_CONV_GEMM_FP16),
const GemmOps g_gemm_ops[] = {LH_GEMM_SIZES(LH_GEMM_ENTRY)};

// ---- host: DFT matrices in fragment order -------------------------------------------
uint16_t bf16_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
float bf16_value(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
uint16_t f16_bits(float f) {
    const _Float16 h = (_Float16)f;
    uint16_t b;
    memcpy(&b, &h, 2);
    return b;
}
float f16_value(uint16_t b) {
    _Float16 h;
    memcpy(&h, &b, 2);
    return (float)h;
}

}  // namespace

const GemmOps *lh_gemm_ops(int n, int conv) {
    for (const GemmOps &o : g_gemm_ops)
        if (o.N == n && o.conv == conv) return &o;
    return nullptr;
}

// The twelve n x n matrices (hi, lo) of one size and operand type, unscaled, in fragment
// order; angles 2 pi ((r * j) mod n) / n are reduced exactly in integers.


void lh_build_dft_mats(int n, int conv, std::vector<uint16_t> &out) {
    const bool f16 = conv == LH_CONV_GEMM_FP16X3 || conv == LH_CONV_GEMM_FP16;
    const int NN = n * n, KS = n / 32, half = n / 2;
    out.assign((size_t)M_COUNT * NN, 0);
    const double tau = 6.283185307179586476925286766559;
    const double lo_scale = f16 ? 2048.0 : 256.0;  // the kernels' 2^S
    auto put = [&](int slot, int r, int j, double v) {
        const float fv = (float)v;
        uint16_t hi, lo;
        if (f16) {
            hi = f16_bits(fv);
            lo = f16_bits((float)((v - (double)f16_value(hi)) * lo_scale));
        } else {
            hi = bf16_bits(fv);
            lo = bf16_bits((float)((v - (double)bf16_value(hi)) * lo_scale));
        }
        // fragment order: row block r/16, k step j/32, lane (r%16) + 16*((j%32)/8), element j%8
        const int lane = (r % 16) + 16 * ((j % 32) / 8);
        const size_t idx = ((size_t)((r / 16) * KS + j / 32) * 64 + lane) * 8 + (j % 8);
        out[(size_t)slot * NN + idx] = hi;
        out[(size_t)(slot + 1) * NN + idx] = lo;
    };
    for (int r = 0; r < n; r++)
        for (int j = 0; j < n; j++) {
            const double ang = tau * (double)((long long)r * j % n) / n;
            const double cs = std::cos(ang), sn = std::sin(ang);
            put(M_FR, r, j, cs);
            put(M_FI, r, j, -sn);
            put(M_GR, r, j, cs);
            put(M_GI, r, j, sn);
            // packed forward real: rows 0..n/2 -> re bins, rows n/2+1.. -> im bins 1..n/2-1
            if (r <= half) {
                put(M_XF, r, j, std::cos(tau * (double)((long long)r * j % n) / n));
            } else {
                const int k = r - half;
                put(M_XF, r, j, -std::sin(tau * (double)((long long)k * j % n) / n));
            }
            // packed inverse real (row = x = r, column = packed index j)
            {
                const int x = r, p = j;
                double v;
                if (p == 0) v = 1.0;
                else if (p < half) v = 2.0 * std::cos(tau * (double)((long long)p * x % n) / n);
                else if (p == half) v = (x & 1) ? -1.0 : 1.0;
                else v = -2.0 * std::sin(tau * (double)((long long)(p - half) * x % n) / n);
                put(M_XI, x, p, v);
            }
        }
}
