// lh_kernels_big.hpp -- spectral kernels for large worlds (sides above 48) (templates).
//
// A line of 96 or 128 complex points no longer fits one thread's registers, so these
// kernels stage planes and columns in LDS and transform them with the block-cooperative
// Stockham (fft_lds_batch), whose per-butterfly arithmetic is that of the register
// version: the large-world step is bit-identical to the CPU checker just like the small
// one.  The gradient and transport phases reuse the tiled kernels of lh_device.hpp with
// 32-wide x tiles.
//
//   fwd_xy_big     one z-plane of one field per workgroup: x r2c, transpose, y FFT
//   spec_z_big     16 (kx,ky) columns per workgroup: z FFT, x kernel octant, inverse z
//   inv_u_big      one z-plane of one kernel (or blurred field) per workgroup (1024
//                  threads): inverse y, c2r x; the real plane replaces the spectrum plane
//                  (or, blurred, becomes U_S / SUM_S)
//   growth_sum_big per cell, kernels in order: growth into H / U_SUM, then B
#pragma once
#include <hip/hip_runtime.h>
#include "lh_device.hpp"

template <int W, int H, int D>
struct BigCfg {
    static constexpr int Wh = W / 2 + 1;
    // line pitches of the LDS-cooperative transforms: odd, so a wave's lanes (one line
    // each, fft_lds_batch) fall on distinct banks
    static constexpr int LX = W + 1, LY = H + 1, LZ = D + 1;
    static constexpr int BUF = ((H / 2) * LX > Wh * LY) ? (H / 2) * LX : Wh * LY;  // float2 per buffer
    static constexpr int CB = 16;   // columns per spec_z workgroup
    static constexpr int NT = 256;
    static constexpr int NT_INV = 1024;
};

template <int W, int H, int D>
__global__ __launch_bounds__(256) void k_fwd_xy_big(PhaseArgs a, int arg) {
    using S = Shape<W, H, D>;
    using Cfg = BigCfg<W, H, D>;
    constexpr int Wh = Cfg::Wh, NT = Cfg::NT, LX = Cfg::LX, LY = Cfg::LY;
    __shared__ float2 ba[Cfg::BUF], bb[Cfg::BUF];
    const int src = arg & 0xFF, nfmax = arg >> 8;
    const int per_world = nfmax * D;
    const int wi = blockIdx.x / per_world;
    const int f = (blockIdx.x % per_world) / D;
    const int z = blockIdx.x % D;
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
    const float *pl = in + (long long)z * S::PL;
    // rows (2l, 2l+1) packed as complex line l
    for (int t = threadIdx.x; t < (H / 2) * W; t += NT) {
        const int l = t / W, x = t % W;
        ba[l * LX + x] = make_float2(pl[(2 * l) * W + x], pl[(2 * l + 1) * W + x]);
    }
    lds_barrier();
    fft_lds_batch<W, false, H / 2, LX, NT>(ba, bb, a.twW);
    // unpack the two real spectra into [kx][y] lines for the y transform
    for (int t = threadIdx.x; t < (H / 2) * Wh; t += NT) {
        const int l = t / Wh, k = t % Wh;
        const float2 Z = ba[l * LX + k], P = ba[l * LX + (W - k) % W];
        bb[k * LY + 2 * l] = make_float2((Z.x + P.x) * 0.5f, (Z.y - P.y) * 0.5f);
        bb[k * LY + 2 * l + 1] = make_float2((Z.y + P.y) * 0.5f, (P.x - Z.x) * 0.5f);
    }
    lds_barrier();
    fft_lds_batch<H, false, Wh, LY, NT>(bb, ba, a.twH);
    float2 *o = out + (long long)z * H * Wh;
    for (int t = threadIdx.x; t < H * Wh; t += NT) {
        const int y = t / Wh, k = t % Wh;
        o[y * Wh + k] = bb[k * LY + y];
    }
}

// mode 0: growth kernels (z FFT per used channel, x octant, inverse z per kernel);
// mode 1: blur of fields 0..C; mode 2: kernel spectra (forward z, keep Re of the octant).
template <int W, int H, int D>
__global__ __launch_bounds__(256) void k_spec_z_big(PhaseArgs a, int mode) {
    using S = Shape<W, H, D>;
    using Cfg = BigCfg<W, H, D>;
    constexpr int Wh = S::Wh, Hh = S::Hh, Dh = S::Dh, CB = Cfg::CB, NT = Cfg::NT, COLS = Wh * H, LZ = Cfg::LZ;
    constexpr int BPW = (COLS + CB - 1) / CB;
    __shared__ float2 sT[CB * LZ], sY[CB * LZ], sS[CB * LZ];
    const int wi = blockIdx.x / BPW;
    const int col0 = (blockIdx.x % BPW) * CB;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const float2 *Fsp = reinterpret_cast<const float2 *>(mode == 2 ? a.pool2 + w.KSP : a.pool + w.Fsp);
    float2 *Gsp = reinterpret_cast<float2 *>(a.pool + w.Gsp);
    const int nsrc = mode == 0 ? w.C : (mode == 1 ? w.C + 1 : w.nkf);
    for (int c = 0; c < nsrc; c++) {
        if (mode == 0) {
            bool used = false;
            for (int k = 0; k < w.K; k++) used |= (w.kern[k].c0 == c);
            if (!used) continue;
        }
        lds_barrier();
        for (int t = threadIdx.x; t < CB * D; t += NT) {
            const int z = t / CB, cb = t % CB, col = col0 + cb;
            sT[cb * LZ + z] = col < COLS ? Fsp[(long long)c * S::Nh + (long long)z * COLS + col] : make_float2(0.0f, 0.0f);
        }
        lds_barrier();
        fft_lds_batch<D, false, CB, LZ, NT>(sT, sS, a.twD);
        if (mode == 2) {
            float *oct = c < w.K ? a.pool + w.KH + (long long)c * S::Noct : a.pool + w.BH;
            for (int t = threadIdx.x; t < CB * Dh; t += NT) {
                const int z = t / CB, cb = t % CB, col = col0 + cb;
                if (col >= COLS) continue;
                const int kx = col % Wh, ky = col / Wh;
                if (ky <= H / 2) oct[(z * Hh + ky) * Wh + kx] = sT[cb * LZ + z].x;
            }
            continue;
        }
        const int nk = mode == 0 ? w.K : 1;
        // each kernel's octant values (this thread's elements) are loaded before the
        // previous kernel's stores: a load issued after a store would wait for it
        constexpr int EPT = (CB * D + NT - 1) / NT;
        auto next_k = [&](int k) {
            for (k = k + 1; k < nk; k++)
                if (mode != 0 || w.kern[k].c0 == c) return k;
            return nk;
        };
        float kv[EPT];
        auto load_oct = [&](int k) {
            if (k >= nk) return;
            const float *oct = mode == 0 ? a.pool + w.KH + (long long)k * S::Noct : a.pool + w.BH;
#pragma unroll
            for (int e = 0; e < EPT; e++) {
                const int t = threadIdx.x + e * NT;
                const int z = t / CB, cb = t % CB, col = col0 + cb;
                const int kx = col % Wh, ky = col / Wh;
                const int kyo = ky <= H / 2 ? ky : H - ky, zo = z <= D / 2 ? z : D - z;
                kv[e] = (t < CB * D && col < COLS) ? oct[(zo * Hh + kyo) * Wh + kx] : 0.0f;
            }
        };
        int k = next_k(-1);
        load_oct(k);
        while (k < nk) {
            const int kn = next_k(k);
#pragma unroll
            for (int e = 0; e < EPT; e++) {
                const int t = threadIdx.x + e * NT;
                if (t >= CB * D) break;
                const int z = t / CB, cb = t % CB;
                const float2 tv = sT[cb * LZ + z];
                sY[cb * LZ + z] = make_float2(tv.x * kv[e], tv.y * kv[e]);
            }
            load_oct(kn);
            lds_barrier();
            fft_lds_batch<D, true, CB, LZ, NT>(sY, sS, a.twD);
            float2 *dst = Gsp + (long long)(mode == 0 ? k : c) * S::Nh;
            for (int t = threadIdx.x; t < CB * D; t += NT) {
                const int z = t / CB, cb = t % CB, col = col0 + cb;
                if (col < COLS) dst[(long long)z * COLS + col] = sY[cb * LZ + z];
            }
            lds_barrier();
            k = kn;
        }
    }
}

// Loads plane z of half spectrum field g into [kx][y] lines, inverse y FFT, c2r along x.
// Leaves the real plane in `lines` as u(x, 2l) = lines[l*W+x].x, u(x, 2l+1) = .y.
// Line pitches LX = W + 1, LY = H + 1 (odd, see BigCfg): the result is u(x, 2l) =
// lines[l*LX+x].x, u(x, 2l+1) = .y.
template <int W, int H, int NT>
__device__ __forceinline__ void inv_plane(const float2 *g, float2 *cols, float2 *lines, const float2 *twW,
                                          const float2 *twH) {
    constexpr int Wh = W / 2 + 1, LX = W + 1, LY = H + 1;
    for (int t = threadIdx.x; t < H * Wh; t += NT) {
        const int y = t / Wh, k = t % Wh;
        cols[k * LY + y] = g[y * Wh + k];
    }
    lds_barrier();
    fft_lds_batch<H, true, Wh, LY, NT>(cols, lines, twH);
    for (int t = threadIdx.x; t < (H / 2) * W; t += NT) {
        const int l = t / W, k = t % W;
        const float2 *A = cols + 2 * l, *B = cols + 2 * l + 1;  // row y at cols[kx*LY + y]
        float2 v;
        if (k == 0) v = make_float2(A[0].x, B[0].x);
        else if (k == W / 2) v = make_float2(A[(W / 2) * LY].x, B[(W / 2) * LY].x);
        else if (k < W / 2) {
            const float2 av = A[k * LY], bv = B[k * LY];
            v = make_float2(av.x - bv.y, av.y + bv.x);
        } else {
            const float2 av = A[(W - k) * LY], bv = B[(W - k) * LY];
            v = make_float2(av.x + bv.y, bv.x - av.y);
        }
        lines[l * LX + k] = v;
    }
    lds_barrier();
    fft_lds_batch<W, true, H / 2, LX, NT>(lines, cols, twW);
}

// inv_u_big: one workgroup per (world, z-plane, field): inverse y-FFT and c2r x of that
// plane of the field's spectrum.  mode 0 (growth kernel f): the unnormalised real plane is
// written back over the spectrum plane it came from (H*W floats fit in the H*Wh float2 it
// held; the plane was staged into LDS before the first barrier), for growth_sum_big;
// mode 1 (blur field f): U_S_f or SUM_S = plane / N (h_sum_blur reintegration.cpp:156-177).
// A plane per field rather than per world: a 96^3 world's 16 kernels give 1536 workgroups.
template <int W, int H, int D, int NF>
__global__ __launch_bounds__(1024) void k_inv_u_big(PhaseArgs a, int mode) {
    using S = Shape<W, H, D>;
    using Cfg = BigCfg<W, H, D>;
    constexpr int PL = S::PL, NT = Cfg::NT_INV, LX = Cfg::LX;
    __shared__ float2 bc[Cfg::BUF], bl[Cfg::BUF];
    const int f = blockIdx.x % NF, z = (blockIdx.x / NF) % D, wi = blockIdx.x / (NF * D);
    const DevWorld &w = a.worlds[a.ids[wi]];
    if (f >= (mode == 0 ? w.K : w.C + 1)) return;
    float2 *G = reinterpret_cast<float2 *>(a.pool + w.Gsp) + (long long)f * S::Nh + (long long)z * H * S::Wh;
    inv_plane<W, H, NT>(G, bc, bl, a.twW, a.twH);
    float *dst = mode == 0 ? reinterpret_cast<float *>(G)
                           : (f < w.C ? a.pool + w.US + (long long)f * S::N : a.pool + w.SUMS) + (long long)z * PL;
    for (int i = threadIdx.x; i < PL; i += NT) {
        const int y = i / W, x = i % W;
        const float2 lv = bl[(y >> 1) * LX + x];
        const float v = (y & 1) ? lv.y : lv.x;
        dst[i] = mode == 0 ? v : v / (float)S::N;
    }
}

// growth_sum_big: per cell, the growth of every kernel in kernel order from the real
// planes k_inv_u_big left (mt_fft_step lenia.cpp:501-538): u = plane / N, U_SUM += u,
// H += h * bell(u) (minus A for an asymptotic channel), then B (lenia.cpp:542-592).
template <int W, int H, int D>
__global__ __launch_bounds__(256) void k_growth_sum_big(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    constexpr int BPW = (S::N + 255) / 256;
    const int wi = blockIdx.x / BPW;
    const int i = (blockIdx.x % BPW) * 256 + threadIdx.x;
    if (i >= S::N) return;
    const DevWorld *wp = a.worlds + a.ids[wi];
    const int C = wp->C, K = wp->K;
    const bool lut = (wp->quirks & LHQ_GROWTH_LUT) != 0, ku = wp->keep_usum != 0;
    const float *A = a.pool + (a.parity ? wp->A2 : wp->A);
    const float *U = a.pool + wp->Gsp;  // kernel k, plane z: real values at 2 (k Nh + z H Wh)
    const int z = i / S::PL, r = i % S::PL;
    const long long zoff = 2LL * z * H * S::Wh + r;
    float h[LHD_MAX_C] = {0.0f, 0.0f, 0.0f, 0.0f}, us[LHD_MAX_C] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = 0; k < K; k++) {
        const DevKernel kp = wp->kern[k];
        const float u = U[2LL * k * S::Nh + zoff] / (float)S::N;
        float g;
        if (lut) {
            int iu = (int)(u * 255.0f);
            iu = iu < 0 ? 0 : (iu > 255 ? 255 : iu);
            g = a.pool[wp->LUT + (long long)k * 256 + iu];
        } else {
            g = bell_growth(u, kp.m, kp.s);
        }
        const bool asym = (wp->asym_mask >> kp.c1) & 1;
        const float hk = asym ? (g - A[(long long)kp.c1 * S::N + i]) * kp.h : g * kp.h;
#pragma unroll
        for (int c = 0; c < LHD_MAX_C; c++) {
            if (c != kp.c1) continue;
            us[c] += u;
            h[c] += hk;
        }
    }
    float *Hg = a.pool + wp->Hg, *B = a.pool + wp->Bk, *US = a.pool + wp->USUM;
#pragma unroll
    for (int c = 0; c < LHD_MAX_C; c++) {
        if (c >= C) break;
        const long long g = (long long)c * S::N + i;
        Hg[g] = h[c];
        if (ku) US[g] = us[c];
        if (wp->growth_type == 1)
            B[g] = ((wp->soft_mask >> c) & 1) ? soft_growth(A[g], h[c], wp->dt) : clamp_ref(A[g] + h[c] * wp->dt, 0.0f, 1.0f);
    }
}

template <int W, int H, int D>
struct LaunchBig {
    using Cfg = BigCfg<W, H, D>;
    static constexpr int SPECB = ((W / 2 + 1) * H + Cfg::CB - 1) / Cfg::CB;
    static void fwd_xy(const PhaseArgs &a, int arg, void *st) {
        hipLaunchKernelGGL((k_fwd_xy_big<W, H, D>), dim3(a.nworlds * (arg >> 8) * D), dim3(Cfg::NT), 0,
                           (hipStream_t)st, a, arg);
    }
    static void spec_z(const PhaseArgs &a, int mode, void *st) {
        hipLaunchKernelGGL((k_spec_z_big<W, H, D>), dim3(a.nworlds * SPECB), dim3(Cfg::NT), 0, (hipStream_t)st, a,
                           mode);
    }
    static void kspec_z(const PhaseArgs &a, int, void *st) {
        hipLaunchKernelGGL((k_spec_z_big<W, H, D>), dim3(a.nworlds * SPECB), dim3(Cfg::NT), 0, (hipStream_t)st, a, 2);
    }
    static void inv_growth(const PhaseArgs &a, int, void *st) {
        hipLaunchKernelGGL((k_inv_u_big<W, H, D, LHD_MAX_K>), dim3(a.nworlds * D * LHD_MAX_K), dim3(Cfg::NT_INV), 0,
                           (hipStream_t)st, a, 0);
        hipLaunchKernelGGL((k_growth_sum_big<W, H, D>), dim3(a.nworlds * ((W * H * D + 255) / 256)), dim3(256), 0,
                           (hipStream_t)st, a, 0);
    }
    static void inv_blur(const PhaseArgs &a, int, void *st) {
        hipLaunchKernelGGL((k_inv_u_big<W, H, D, LHD_MAX_C + 1>), dim3(a.nworlds * D * (LHD_MAX_C + 1)),
                           dim3(Cfg::NT_INV), 0, (hipStream_t)st, a, 1);
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
        return ShapeOps{W, H, D, &fwd_xy, &spec_z, &inv_growth, &inv_blur, &local, &transport, &kspec_z, nullptr};
    }
};
