// lh_kernels_big.hip -- spectral kernels for large worlds (sides 64..128).
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
//   inv_growth_big one z-plane per workgroup (1024 threads), kernels in order: inverse y,
//                  c2r x, growth accumulated into the world's H / U_SUM planes
//   inv_blur_big   same for the C+1 blurred fields
#include <hip/hip_runtime.h>
#include "lh_device.hpp"

template <int W, int H, int D>
struct BigCfg {
    static constexpr int Wh = W / 2 + 1;
    static constexpr int BUF = ((H / 2) * W > Wh * H) ? (H / 2) * W : Wh * H;  // float2 per buffer
    static constexpr int CB = 16;   // columns per spec_z workgroup
    static constexpr int NT = 256;
    static constexpr int NT_INV = 1024;
};

template <int W, int H, int D>
__global__ __launch_bounds__(256) void k_fwd_xy_big(PhaseArgs a, int arg) {
    using S = Shape<W, H, D>;
    using Cfg = BigCfg<W, H, D>;
    constexpr int Wh = Cfg::Wh, NT = Cfg::NT;
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
        in = f < w.C ? a.pool + w.Hg + (long long)f * S::N : a.pool + w.S;
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
        ba[l * W + x] = make_float2(pl[(2 * l) * W + x], pl[(2 * l + 1) * W + x]);
    }
    __syncthreads();
    fft_lds_batch<W, false, H / 2, W, NT>(ba, bb, a.twW);
    // unpack the two real spectra into [kx][y] lines for the y transform
    for (int t = threadIdx.x; t < (H / 2) * Wh; t += NT) {
        const int l = t / Wh, k = t % Wh;
        const float2 Z = ba[l * W + k], P = ba[l * W + (W - k) % W];
        bb[k * H + 2 * l] = make_float2((Z.x + P.x) * 0.5f, (Z.y - P.y) * 0.5f);
        bb[k * H + 2 * l + 1] = make_float2((Z.y + P.y) * 0.5f, (P.x - Z.x) * 0.5f);
    }
    __syncthreads();
    fft_lds_batch<H, false, Wh, H, NT>(bb, ba, a.twH);
    float2 *o = out + (long long)z * H * Wh;
    for (int t = threadIdx.x; t < H * Wh; t += NT) {
        const int y = t / Wh, k = t % Wh;
        o[y * Wh + k] = bb[k * H + y];
    }
}

// mode 0: growth kernels (z FFT per used channel, x octant, inverse z per kernel);
// mode 1: blur of fields 0..C; mode 2: kernel spectra (forward z, keep Re of the octant).
template <int W, int H, int D>
__global__ __launch_bounds__(256) void k_spec_z_big(PhaseArgs a, int mode) {
    using S = Shape<W, H, D>;
    using Cfg = BigCfg<W, H, D>;
    constexpr int Wh = S::Wh, Hh = S::Hh, Dh = S::Dh, CB = Cfg::CB, NT = Cfg::NT, COLS = Wh * H;
    constexpr int BPW = (COLS + CB - 1) / CB;
    __shared__ float2 sT[CB * D], sY[CB * D], sS[CB * D];
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
        __syncthreads();
        for (int t = threadIdx.x; t < CB * D; t += NT) {
            const int z = t / CB, cb = t % CB, col = col0 + cb;
            sT[cb * D + z] = col < COLS ? Fsp[(long long)c * S::Nh + (long long)z * COLS + col] : make_float2(0.0f, 0.0f);
        }
        __syncthreads();
        fft_lds_batch<D, false, CB, D, NT>(sT, sS, a.twD);
        if (mode == 2) {
            float *oct = c < w.K ? a.pool + w.KH + (long long)c * S::Noct : a.pool + w.BH;
            for (int t = threadIdx.x; t < CB * Dh; t += NT) {
                const int z = t / CB, cb = t % CB, col = col0 + cb;
                if (col >= COLS) continue;
                const int kx = col % Wh, ky = col / Wh;
                if (ky <= H / 2) oct[(z * Hh + ky) * Wh + kx] = sT[cb * D + z].x;
            }
            continue;
        }
        const int nk = mode == 0 ? w.K : 1;
        for (int k = 0; k < nk; k++) {
            if (mode == 0 && w.kern[k].c0 != c) continue;
            const float *oct = mode == 0 ? a.pool + w.KH + (long long)k * S::Noct : a.pool + w.BH;
            for (int t = threadIdx.x; t < CB * D; t += NT) {
                const int z = t / CB, cb = t % CB, col = col0 + cb;
                const int kx = col % Wh, ky = col / Wh;
                const int kyo = ky <= H / 2 ? ky : H - ky, zo = z <= D / 2 ? z : D - z;
                const float kv = col < COLS ? oct[(zo * Hh + kyo) * Wh + kx] : 0.0f;
                const float2 tv = sT[cb * D + z];
                sY[cb * D + z] = make_float2(tv.x * kv, tv.y * kv);
            }
            __syncthreads();
            fft_lds_batch<D, true, CB, D, NT>(sY, sS, a.twD);
            float2 *dst = Gsp + (long long)(mode == 0 ? k : c) * S::Nh;
            for (int t = threadIdx.x; t < CB * D; t += NT) {
                const int z = t / CB, cb = t % CB, col = col0 + cb;
                if (col < COLS) dst[(long long)z * COLS + col] = sY[cb * D + z];
            }
            __syncthreads();
        }
    }
}

// Loads plane z of half spectrum field g into [kx][y] lines, inverse y FFT, c2r along x.
// Leaves the real plane in `lines` as u(x, 2l) = lines[l*W+x].x, u(x, 2l+1) = .y.
template <int W, int H, int NT>
__device__ __forceinline__ void inv_plane(const float2 *g, float2 *cols, float2 *lines, const float2 *twW,
                                          const float2 *twH) {
    constexpr int Wh = W / 2 + 1;
    for (int t = threadIdx.x; t < H * Wh; t += NT) {
        const int y = t / Wh, k = t % Wh;
        cols[k * H + y] = g[y * Wh + k];
    }
    __syncthreads();
    fft_lds_batch<H, true, Wh, H, NT>(cols, lines, twH);
    for (int t = threadIdx.x; t < (H / 2) * W; t += NT) {
        const int l = t / W, k = t % W;
        const float2 *A = cols + 2 * l, *B = cols + 2 * l + 1;  // row y at cols[kx*H + y]
        float2 v;
        if (k == 0) v = make_float2(A[0].x, B[0].x);
        else if (k == W / 2) v = make_float2(A[(W / 2) * H].x, B[(W / 2) * H].x);
        else if (k < W / 2) {
            const float2 av = A[k * H], bv = B[k * H];
            v = make_float2(av.x - bv.y, av.y + bv.x);
        } else {
            const float2 av = A[(W - k) * H], bv = B[(W - k) * H];
            v = make_float2(av.x + bv.y, bv.x - av.y);
        }
        lines[l * W + k] = v;
    }
    __syncthreads();
    fft_lds_batch<W, true, H / 2, W, NT>(lines, cols, twW);
}

// Growth for one z-plane; H (and U_SUM) accumulate in the world's own planes in HBM:
// the workgroup owns the plane and each cell is updated by one thread, kernels in order.
template <int W, int H, int D>
__global__ __launch_bounds__(1024) void k_inv_growth_big(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    using Cfg = BigCfg<W, H, D>;
    constexpr int PL = S::PL, NT = Cfg::NT_INV;
    __shared__ float2 bc[Cfg::BUF], bl[Cfg::BUF];
    __shared__ float slut[256];
    const int wi = blockIdx.x / D, z = blockIdx.x % D;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const int tid = threadIdx.x, C = w.C, K = w.K;
    const bool lut = (w.quirks & LHQ_GROWTH_LUT) != 0;
    const float *A = A_cur(a, w) + (long long)z * PL;
    float *US = a.pool + w.USUM + (long long)z * PL;
    float *Hg = a.pool + w.Hg + (long long)z * PL, *B = a.pool + w.Bk + (long long)z * PL;
    for (int i = tid; i < PL; i += NT)
        for (int c = 0; c < C; c++) {
            Hg[(long long)c * S::N + i] = 0.0f;
            if (w.keep_usum) US[(long long)c * S::N + i] = 0.0f;
        }
    const float2 *G = reinterpret_cast<const float2 *>(a.pool + w.Gsp);
    for (int k = 0; k < K; k++) {
        __syncthreads();
        if (lut && tid < 256) slut[tid] = a.pool[w.LUT + (long long)k * 256 + tid];
        inv_plane<W, H, NT>(G + (long long)k * S::Nh + (long long)z * H * S::Wh, bc, bl, a.twW, a.twH);
        const DevKernel &kp = w.kern[k];
        const long long co = (long long)kp.c1 * S::N;
        const bool asym = (w.asym_mask >> kp.c1) & 1;
        for (int i = tid; i < PL; i += NT) {
            const int y = i / W, x = i % W;
            const float2 lv = bl[(y >> 1) * W + x];
            const float u = ((y & 1) ? lv.y : lv.x) / (float)S::N;
            float g;
            if (lut) {
                int iu = (int)(u * 255.0f);
                iu = iu < 0 ? 0 : (iu > 255 ? 255 : iu);
                g = slut[iu];
            } else {
                g = bell_growth(u, kp.m, kp.s);
            }
            if (w.keep_usum) US[co + i] += u;
            if (asym) Hg[co + i] += (g - A[co + i]) * kp.h;
            else Hg[co + i] += g * kp.h;
        }
    }
    if (w.growth_type != 1) return;
    for (int i = tid; i < PL; i += NT)
        for (int c = 0; c < C; c++) {
            const long long gi = (long long)c * S::N + i;
            const float hv = Hg[gi];
            B[gi] = ((w.soft_mask >> c) & 1) ? soft_growth(A[gi], hv, w.dt) : clamp_ref(A[gi] + hv * w.dt, 0.0f, 1.0f);
        }
}

template <int W, int H, int D>
__global__ __launch_bounds__(1024) void k_inv_blur_big(PhaseArgs a, int) {
    using S = Shape<W, H, D>;
    using Cfg = BigCfg<W, H, D>;
    constexpr int PL = S::PL, NT = Cfg::NT_INV;
    __shared__ float2 bc[Cfg::BUF], bl[Cfg::BUF];
    const int wi = blockIdx.x / D, z = blockIdx.x % D;
    const DevWorld &w = a.worlds[a.ids[wi]];
    const float2 *G = reinterpret_cast<const float2 *>(a.pool + w.Gsp);
    for (int f = 0; f <= w.C; f++) {
        __syncthreads();
        inv_plane<W, H, NT>(G + (long long)f * S::Nh + (long long)z * H * S::Wh, bc, bl, a.twW, a.twH);
        float *dst = (f < w.C ? a.pool + w.US + (long long)f * S::N : a.pool + w.SUMS) + (long long)z * PL;
        for (int i = threadIdx.x; i < PL; i += NT) {
            const int y = i / W, x = i % W;
            const float2 lv = bl[(y >> 1) * W + x];
            dst[i] = ((y & 1) ? lv.y : lv.x) / (float)S::N;
        }
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
        hipLaunchKernelGGL((k_inv_growth_big<W, H, D>), dim3(a.nworlds * D), dim3(Cfg::NT_INV), 0, (hipStream_t)st, a,
                           0);
    }
    static void inv_blur(const PhaseArgs &a, int, void *st) {
        hipLaunchKernelGGL((k_inv_blur_big<W, H, D>), dim3(a.nworlds * D), dim3(Cfg::NT_INV), 0, (hipStream_t)st, a, 0);
    }
    // arg: maxC << 8 | planes per sweep workgroup << 16 (0: sweep_zseg)
    static void local(const PhaseArgs &a, int arg, void *st) {
        if (((arg >> 8) & 0xff) <= 3) {
            using SC = SweepCfg<W, H, D, 3>;
            const int zs = (arg >> 16) ? (arg >> 16) : sweep_zseg(a.nworlds, SC::BLOCKS, D), ns = (D + zs - 1) / zs;
            hipLaunchKernelGGL((k_local_sweep<W, H, D, 3>), dim3(a.nworlds * SC::BLOCKS * ns), dim3(SC::THREADS), 0,
                               (hipStream_t)st, a, zs);
        } else {
            using SC = SweepCfg<W, H, D, LHD_MAX_C>;
            const int zs = (arg >> 16) ? (arg >> 16) : sweep_zseg(a.nworlds, SC::BLOCKS, D), ns = (D + zs - 1) / zs;
            hipLaunchKernelGGL((k_local_sweep<W, H, D, LHD_MAX_C>), dim3(a.nworlds * SC::BLOCKS * ns),
                               dim3(SC::THREADS), 0, (hipStream_t)st, a, zs);
        }
    }
    // arg bits: 1 some world enables an oscillator, 2 some world is moved by the
    // z-sweep (RT, dd = 1, no oscillator), 4 some world is not (or the tiled kernel moves
    // every world); maxC << 8; planes per sweep workgroup << 16 (0: sweep_zseg)
    static void transport(const PhaseArgs &a, int arg, void *st) {
        if (arg & 2) {
            const int maxc = (arg >> 8) & 0xff;
            if (maxc <= 3) {
                using SC = SweepCfg<W, H, D, 3, TRANSPORT_TY>;
                const int zs = (arg >> 16) ? (arg >> 16) : sweep_zseg(a.nworlds, SC::BLOCKS, D), ns = (D + zs - 1) / zs;
                hipLaunchKernelGGL((k_transport_sweep<W, H, D, 3>), dim3(a.nworlds * SC::BLOCKS * ns),
                                   dim3(SC::THREADS), (size_t)maxc * SC::LDS_FLOATS_PER_CH * sizeof(float),
                                   (hipStream_t)st, a, zs);
            } else {
                using SC = SweepCfg<W, H, D, LHD_MAX_C, TRANSPORT_TY>;
                const int zs = (arg >> 16) ? (arg >> 16) : sweep_zseg(a.nworlds, SC::BLOCKS, D), ns = (D + zs - 1) / zs;
                hipLaunchKernelGGL((k_transport_sweep<W, H, D, LHD_MAX_C>), dim3(a.nworlds * SC::BLOCKS * ns),
                                   dim3(SC::THREADS), (size_t)maxc * SC::LDS_FLOATS_PER_CH * sizeof(float),
                                   (hipStream_t)st, a, zs);
            }
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

static const ShapeOps g_ops_big[] = {LaunchBig<64, 64, 64>::ops(), LaunchBig<96, 96, 96>::ops(),
                                     LaunchBig<128, 128, 128>::ops()};

const ShapeOps *lh_shape_ops_big(int W, int H, int D) {
    for (const ShapeOps &o : g_ops_big)
        if (o.W == W && o.H == H && o.D == D) return &o;
    return nullptr;
}
