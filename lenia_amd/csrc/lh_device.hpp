// lh_device.hpp -- device helpers and the shape-generic stencil kernels (gradients,
// alpha, F, mu, V; SRT / RT / advection; TRM) shared by the small- and large-world
// kernel files.  All arithmetic follows the reference's expression order in fp32.
#pragma once
#include <hip/hip_runtime.h>
#include "lh_common.hpp"
#include "lh_fft.hpp"

#define PADR 4  // reference PADDING (render.h:291), used by the index-quirk emulation

__device__ __forceinline__ float fmax_ref(float a, float b) { return a > b ? a : b; }  // _mm_max_ps
__device__ __forceinline__ float fmin_ref(float a, float b) { return a < b ? a : b; }  // _mm_min_ps
__device__ __forceinline__ float clamp_ref(float a, float lo, float hi) { return fmin_ref(fmax_ref(a, lo), hi); }
// clamp_ref in one instruction (v_med3_f32), for lo < hi where it gives the same bits: a
// NaN gives lo in both; a zero a is returned unchanged by both unless a bound is also a
// zero of the other sign, so callers use it only where a is never -0 or no bound is 0.
__device__ __forceinline__ float clamp_med3(float a, float lo, float hi) { return __builtin_amdgcn_fmed3f(a, lo, hi); }

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
// S (the end-of-step sum the gradients read) alternates like A: a step reads S_cur and
// its end-of-step transform writes S_next, so no cell's new S can reach a neighbour's
// gradient within the step, whichever kernel runs the transform.
__device__ __forceinline__ const float *S_cur(const PhaseArgs &a, const DevWorld &w) {
    return a.pool + (a.parity ? w.S2 : w.S);
}
__device__ __forceinline__ float *S_next(const PhaseArgs &a, const DevWorld &w) {
    return a.pool + (a.parity ? w.S : w.S2);
}

// exp for the analytic growth bell: Cody-Waite reduction x = n ln2 + r (fdlibm's split of
// ln2 into a 33-bit head and a tail), the degree-13 Taylor polynomial of e^r in Horner
// order with 1/k! rounded once, then scaling by 2^n.  Built without contraction, so the
// CPU checker's identical sequence gives identical bits; relative error about 2e-16.
__device__ __forceinline__ double canon_exp(double x) {
    if (x < -700.0) return 0.0;
    const double n = floor(x * 0x1.71547652b82fep+0 + 0.5);
    const double r = (x - n * 0x1.62e42fee00000p-1) - n * 0x1.a39ef35793c76p-33;
    double p = 0x1.6124613a86d09p-33;
    p = p * r + 0x1.1eed8eff8d898p-29;
    p = p * r + 0x1.ae64567f544e4p-26;
    p = p * r + 0x1.27e4fb7789f5cp-22;
    p = p * r + 0x1.71de3a556c734p-19;
    p = p * r + 0x1.a01a01a01a01ap-16;
    p = p * r + 0x1.a01a01a01a01ap-13;
    p = p * r + 0x1.6c16c16c16c17p-10;
    p = p * r + 0x1.1111111111111p-7;
    p = p * r + 0x1.5555555555555p-5;
    p = p * r + 0x1.5555555555555p-3;
    p = p * r + 0x1.0000000000000p-1;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)n);
}

// Analytic growth bell(u, m, s) (render.h:674-677): quotient in fp32, square and exp in
// double as the reference's pow() promotes, rounded once to fp32.
__device__ __forceinline__ float bell_growth(float u, float m, float s) {
    const float q = (u - m) / s;
    const double d = (double)q;
    return (float)canon_exp(-(d * d) / 2.0);
}

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

// ------------------------------------------------------------------------------------
// local: per cell.  Gradients of S and H_c (Sobel with the reference's axis order, or
// central differences), alpha (with the PREALPHA index quirk), F, mu, V.
// ------------------------------------------------------------------------------------
// ------------------------------------------------------------------------------------
// Tiled variants of local and transport.  A workgroup owns a TX x TY x TZ block of
// cells (whole x rows) and stages the block plus a one-cell halo in LDS, with every
// halo value resolved once at load time, so the per-cell loops carry no branches.
// ------------------------------------------------------------------------------------
// x-width of a stencil kernel's tile: whole rows up to 48 cells; wider rows in tiles of
// 32 cells where the width is a multiple of 32, else in the fewest equal tiles of at most
// 48 cells.
constexpr int stencil_tx(int W) {
    if (W <= 48) return W;
    if (W % 32 == 0) return 32;
    for (int n = (W + 47) / 48; n <= W; n++)
        if (W % n == 0) return W / n;
    return 1;
}

template <int W, int H, int D>
struct TileCfg {
    static constexpr int TX = stencil_tx(W);
    static constexpr int XT = W / TX;
    static constexpr int TY = H % 4 == 0 ? 4 : (H % 2 == 0 ? 2 : 1);
    static constexpr int TZ = D % 4 == 0 ? 4 : (D % 2 == 0 ? 2 : 1);
    static constexpr int SX = TX + 2, SY = TY + 2, SZ = TZ + 2;
    static constexpr int NS = SX * SY * SZ;   // staged slots
    static constexpr int NT = TX * TY * TZ;   // cells per tile
    static constexpr int TILES = XT * (H / TY) * (D / TZ);
    static constexpr int THREADS = 256;
    static constexpr int CPT = (NT + THREADS - 1) / THREADS;
    __device__ static __forceinline__ int slot(int tx, int ty, int tz) { return tx + SX * (ty + SY * tz); }
};

// Element idx of a world array (idx < 2^30): a 32-bit byte offset from the array's
// base, so the address is base + zext(offset) -- one 32-bit add with the base in scalar
// registers, where a 64-bit index costs a 64-bit add per access.
__device__ __forceinline__ float &el(float *base, unsigned idx) {
    return *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + (idx << 2));
}
__device__ __forceinline__ const float &el(const float *base, unsigned idx) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + (idx << 2));
}

// The A value alpha reads per channel (PREALPHA index quirk applied; 0 when the quirk's
// index falls outside the channels), loaded by the sweep one plane ahead of its use.
// With SRTL, for an SRT world moved in the sweep, also the cell's moved mass per channel
// and its GA and S (the transport term's and the end-of-step transform's inputs), so that
// no load of the cell's own values waits behind the previous cell's stores.
template <int CM, bool SRTL>
struct LocalIn {
    float aal[CM];
    float arr[SRTL ? CM : 1], ga[SRTL ? CM : 1], s;
};

template <int W, int H, int D, int CM, bool SRTL>
__device__ __forceinline__ void local_load(const PhaseArgs &a, const DevWorld &w, long long i, LocalIn<CM, SRTL> &in) {
    using S = Shape<W, H, D>;
    const int C = w.C;
    const float *A = A_cur(a, w);
    const bool qa = (w.quirks & LHQ_PREALPHA_INDEX) != 0;
#pragma unroll
    for (int c = 0; c < CM; c++) {
        // PREALPHA written at c*P+i and read at c*N+i (lenia.cpp:660, 708)
        const int r = c * S::N + (int)i;
        int cs = c, is = (int)i;
        bool ok = c < C;
        if (qa) { cs = r / (int)S::P; is = r - cs * (int)S::P; ok = ok && cs < C && is < S::N; }
        in.aal[c] = ok ? el(A, (unsigned)(cs * S::N + is)) : 0.0f;
    }
    if constexpr (SRTL) {
        if (srt_local_world(w)) {
            const float *ARR = (w.growth_type == 0 || w.growth_type == 2) ? A : a.pool + w.Bk;
            const float *GA = a.pool + w.GA;
#pragma unroll
            for (int c = 0; c < CM; c++) {
                in.arr[c] = c < C ? el(ARR, (unsigned)(c * S::N + i)) : 0.0f;
                in.ga[c] = c < C ? el(GA, (unsigned)(c * S::N + i)) : 0.0f;
            }
            in.s = (w.quirks & LHQ_SUM_RESET) ? 0.0f : el(S_cur(a, w), (unsigned)i);
        }
    }
}

// Gradients -> alpha, F, mu, V for one cell (sobel_convolve_h lenia.cpp:647-672,
// compute_f lenia.cpp:679-719, compute_*_mu reintegration.cpp:321-602); grad(f, g)
// returns the gradient of staged field f (0: S or SUM_S, 1 + c: H_c or U_S_c).  A
// quirk-index alpha source outside the channels gives alpha = 0, which equals
// clamp((0 / theta)^2, 0, 1).
// For an SRT world moved here (srt_local_world) it also forms each channel's transport
// term tt[c] from the cell's own mu (srt reintegration.cpp:607-688, as transport_cell
// mode 0: the same expression on the same mu value), and writes mu only when kept.
// SRTL: compiled with that path (a launch whose shape class holds such worlds).
template <int W, int H, int D, int CM, bool SRTL, bool PLAIN, class Grad>
__device__ __forceinline__ void local_cell(const PhaseArgs &a, const DevWorld &w, int x, int y, int z,
                                           const LocalIn<CM, SRTL> &in, Grad grad, float (&tt)[CM]) {
    using S = Shape<W, H, D>;
    const int C = w.C;
    const float *A = A_cur(a, w);
    const bool srt = SRTL && srt_local_world(w), store_mu = !srt || w.keep_mu;
    const int nrt = (w.quirks & LHQ_RT_3CH) ? (C < 3 ? C : 3) : C;
    const float dt = w.dt, ma = w.ma, nma = -w.ma, damp = w.damp, sig = w.sigma;
    const float lim[3] = {(float)W - sig, (float)H - sig, (float)D - sig};
    const bool wall = !PLAIN && w.border == 1;
    const bool qa = (w.quirks & LHQ_PREALPHA_INDEX) != 0;
    const long long i = x + (long long)W * (y + (long long)H * z);
    float na[3];
    grad(0, na);
    const float pos[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
#pragma unroll
    for (int c = 0; c < CM; c++) {
        if (c >= C) break;
        // one channel's staged values live at a time: no LDS read of this channel is
        // hoisted above the previous channel's stores
        asm volatile("" ::: "memory");
        float nu[3];
        grad(c + 1, nu);
        float pg = (PLAIN || w.inv_theta != 0.0f) ? in.aal[c] * w.inv_theta : in.aal[c] / w.theta_a;
        pg = pg * pg;
        const float al = clamp_med3(pg, 0.0f, 1.0f);  // pg is a square: never -0
        if (!PLAIN && w.need_alpha) {
            // ALPHA = |A - GA[c*P+i]| (lenia.cpp:661-664), GA read before this step's TRM
            const int r = qa ? c * (int)S::P + (int)i : c * S::N + (int)i;
            const int gc = r / S::N;
            const float ga = (r < LHD_MAX_C * S::N && gc < C) ? el(a.pool + w.GA, (unsigned)r) : 0.0f;
            el(a.pool + w.AL, (unsigned)(c * S::N + i)) = __builtin_fabsf(el(A, (unsigned)(c * S::N + i)) - ga);
        }
        const float al1 = 1.0f - al;
        float F[3];
#pragma unroll
        for (int d = 0; d < 3; d++) F[d] = nu[d] * al1 - na[d] * al;
        float *V = a.pool + w.V;
        const unsigned iu = (unsigned)i;
        float v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) v[d] = (!PLAIN && w.v_live) ? el(V, (unsigned)((3 * c + d) * S::N) + iu) : 0.0f;
        if (!PLAIN && w.mu_vel) {
#pragma unroll
            for (int d = 0; d < 3; d++) v[d] = (v[d] + F[d]) * damp;
        }
        float *MU = a.pool + w.MU;
        if (!PLAIN && w.mode == 2) {
            float vn[3];
#pragma unroll
            for (int d = 0; d < 3; d++) vn[d] = w.mu_vel ? v[d] : (v[d] + F[d]) * damp;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const float src = w.adv_vel ? vn[d] : F[d];
                el(MU, (unsigned)((3 * c + d) * S::N) + iu) = clamp_med3(src * dt, nma, ma);  // -ma < 0 < ma
            }
        } else {
            float mv[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                float m = ((!PLAIN && w.mu_vel) ? v[d] : F[d]) * dt;
                m = clamp_med3(m, nma, ma);  // -ma < 0 < ma
                m = pos[d] + m;
                if (wall) m = clamp_med3(m, sig, lim[d]);  // 0 < sigma < n - sigma
                mv[d] = m;
                if (store_mu) el(MU, (unsigned)((3 * c + d) * S::N) + iu) = m;
            }
            if (srt) {
                float mm[3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    float dv = pos[d] - mv[d];
                    dv = dv < 0.0f ? -dv : dv;
                    dv = w.si - dv;
                    mm[d] = fmin_ref(fmax_ref(dv, 0.0f), w.clip);
                }
                float wt = in.arr[SRTL ? c : 0] * mm[0];
                wt = wt * mm[1];
                wt = wt * mm[2];
                wt = wt * w.sn;
                tt[c] = c < nrt ? 0.0f + wt : 0.0f;  // TT starts at 0 (srt adds to it)
            }
        }
        if (!PLAIN && !w.mu_vel && w.v_live) {  // (MU_VEL implies a live V)
#pragma unroll
            for (int d = 0; d < 3; d++) v[d] = (v[d] + F[d]) * damp;
        }
        if (!PLAIN && w.v_live) {
#pragma unroll
            for (int d = 0; d < 3; d++) el(V, (unsigned)((3 * c + d) * S::N) + iu) = v[d];
        }
        if (!PLAIN && w.need_f) {
            float *Fo = a.pool + w.Fv;
#pragma unroll
            for (int d = 0; d < 3; d++) el(Fo, (unsigned)((3 * c + d) * S::N) + iu) = F[d];
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

// Optional oscillators (lenia.cpp:803-911), added to TT after the reintegration in the
// reference's order: Laplacian of A, flow V.grad A (or F), divergence of V*ALPHA.
template <int W, int H, int D>
__device__ __forceinline__ void osc_cell(const PhaseArgs &a, const DevWorld &w, int x, int y, int z,
                                         float (&tt)[LHD_MAX_C]) {
    using S = Shape<W, H, D>;
    const int faces = (w.quirks & LHQ_FACE_HALO) != 0;
    const int tkind = (w.quirks & LHQ_TRANSPORT_TORUS) ? 0 : w.border;
    const float *G = A_cur(a, w);
    const float *VF = a.pool + (w.osc_vel ? w.V : w.Fv);
    const float *AL = a.pool + w.AL;
    const int i = x + W * (y + H * z);
    int nb[6];  // x+1, x-1, y+1, y-1, z+1, z-1 resolved through the transport halo (-1: zero)
    const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        int qx = x + off[k][0], qy = y + off[k][1], qz = z + off[k][2];
        nb[k] = halo_src<W, H, D>(qx, qy, qz, tkind, faces) ? qx + W * (qy + H * qz) : -1;
    }
#pragma unroll
    for (int c = 0; c < LHD_MAX_C; c++) {
        if (c >= w.C) break;
        const float *Gc = G + c * S::N;
        auto g = [&](int k) { return nb[k] >= 0 ? Gc[nb[k]] : 0.0f; };
        const float fx1 = g(0), fx0 = g(1), fy1 = g(2), fy0 = g(3), fz1 = g(4), fz0 = g(5);
        if (w.osc_la) {
            const float la = fx1 + fx0 + fy1 + fy0 + fz1 + fz0 - 6.0f * Gc[i];
            tt[c] = tt[c] + la * w.osc_la_amt;
        }
        const float *V0 = VF + (c * 3 + 0) * S::N, *V1 = VF + (c * 3 + 1) * S::N, *V2 = VF + (c * 3 + 2) * S::N;
        if (w.osc_flow) {
            const float flow = V0[i] * (fx1 - fx0) + V1[i] * (fy1 - fy0) + V2[i] * (fz1 - fz0);
            tt[c] = tt[c] + flow * w.osc_flow_amt;
        }
        if (w.osc_div_flow) {
            const float *ALc = AL + c * S::N;
            auto j = [&](const float *v, int k) { return nb[k] >= 0 ? v[nb[k]] * ALc[nb[k]] : 0.0f * 0.0f; };
            const float jx = j(V0, 0) - j(V0, 1), jy = j(V1, 2) - j(V1, 3), jz = j(V2, 4) - j(V2, 5);
            const float divj = jx + jy + jz;
            tt[c] = tt[c] + divj * w.osc_div_flow_amt;
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
    float *An = A_next(a, w), *GA = a.pool + w.GA, *Sn = S_next(a, w);
    const float *Hg = a.pool + w.Hg, *Sg = S_cur(a, w);
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
    Sn[i] = w.stam * sold + ttsum * w.stas;
}

// trm_cell split in two so the sweep can load a cell's inputs a step before it needs
// them: trm_load reads them, trm_store applies the same arithmetic as trm_cell.
// The growth-after-transport inputs (H, A) are read at use; GA and S are prefetched.
template <int CM>
struct TrmIn {
    float ga[CM], s;
};

template <int W, int H, int D, int CM>
__device__ __forceinline__ void trm_load(const PhaseArgs &a, const DevWorld &w, long long i, TrmIn<CM> &in) {
    using S = Shape<W, H, D>;
    const float *GA = a.pool + w.GA;
#pragma unroll
    for (int c = 0; c < CM; c++) in.ga[c] = c < w.C ? el(GA, (unsigned)(c * S::N + i)) : 0.0f;
    in.s = (w.quirks & LHQ_SUM_RESET) ? 0.0f : el(S_cur(a, w), (unsigned)i);
}

// C <= CM for every world this is used on.
template <int W, int H, int D, int CM>
__device__ __forceinline__ void trm_store(const PhaseArgs &a, const DevWorld &w, long long i, const float (&tt)[CM],
                                          const TrmIn<CM> &in) {
    using S = Shape<W, H, D>;
    const int C = w.C;
    const bool clamp01 = (w.quirks & LHQ_CLAMP01) != 0;
    const float *Acur = A_cur(a, w), *Hg = a.pool + w.Hg;
    float *An = A_next(a, w), *GA = a.pool + w.GA, *Sg = S_next(a, w);
    float ttsum = 0.0f;
#pragma unroll
    for (int c = 0; c < CM; c++) {
        if (c >= C) break;
        const unsigned g = (unsigned)(c * S::N + i);
        float t = tt[c];
        if (w.growth_type == 2) {
            const bool soft = (w.soft_mask >> c) & 1;
            if (soft && (w.quirks & LHQ_SOFTCLIP_TR)) el(a.pool + w.Bk, g) = soft_growth(el(Acur, g), el(Hg, g), w.dt);
            else if (soft) t = soft_growth(t, el(Hg, g), w.dt);
            else t = clamp01 ? clamp_med3(t + el(Hg, g) * w.dt, 0.0f, 1.0f) : t + el(Hg, g) * w.dt;
        } else if (clamp01) {
            // TT starts at +0 and only gains products, so it (and TT + dt H) is never -0
            t = clamp_med3(t, 0.0f, 1.0f);
        }
        ttsum = ttsum + t;
        el(GA, g) = w.gtam * in.ga[c] + t * w.gtas;
        el(An, g) = t;
    }
    el(Sg, (unsigned)i) = w.stam * in.s + ttsum * w.stas;
}

// Planes per sweep workgroup: the whole depth when the batch alone gives >= 2048
// workgroups, else segments of at least 4 planes until it does.
inline int sweep_zseg(int nworlds, int blocks_per_world, int depth) {
    const long long full = (long long)nworlds * blocks_per_world;
    int nseg = full >= 2048 ? 1 : (int)((2048 + full - 1) / full);
    if (nseg > depth / 4) nseg = depth / 4 > 0 ? depth / 4 : 1;
    return (depth + nseg - 1) / nseg;
}

// Geometry of the z-sweep kernels (k_local_sweep, k_transport_sweep): a workgroup owns
// a TX x TY column of cells through the whole depth and stages one z-plane at a time.
// Rows per sweep workgroup: the divisor of H with the least overhead, (TY + 2) / TY rows
// staged per row moved (the y halo) times the lanes of whole waves per lane used, for at
// most MAXT threads and the transport's plane within 64 KB of LDS, the plane counted at
// ten floats per slot and channel (it holds nine; the rows stay those chosen for ten).
constexpr int sweep_ty(int TX, int H, int CM, int MAXT, int LDSCAP) {
    int best = 1;
    long long bnum = 0, bden = 1;  // best cost as a fraction
    for (int ty = 1; ty <= H; ty++) {
        if (H % ty) continue;
        const int thr = TX * ty;
        if ((ty > 1 && thr > MAXT) || CM * 10 * (TX + 2) * (ty + 2) * 4 > LDSCAP) break;
        const long long num = (long long)(ty + 2) * ((thr + 63) / 64 * 64), den = (long long)ty * thr;
        if (ty == 1 || num * bden < bnum * den) { best = ty; bnum = num; bden = den; }
    }
    return best;
}

template <int W, int H, int D, int CM, int MAXT, int LDSCAP = 65536>
struct SweepCfg {
    static constexpr int TX = stencil_tx(W);
    static constexpr int XT = W / TX;
    static constexpr int TY = sweep_ty(TX, H, CM, MAXT, LDSCAP);
    static constexpr int YT = H / TY;
    static constexpr int THREADS = TX * TY;
    static constexpr int SX = TX + 2, SY = TY + 2, SP = SX * SY;
    static constexpr int ITEMS = (SP * CM + THREADS - 1) / THREADS;  // staged (slot, channel) per thread
    static constexpr int BLOCKS = XT * YT;
    // transport: nine weights per slot (the x weights times the mass), float2-aligned
    static constexpr int LDS_FLOATS_PER_CH = 9 * SP + (SP & 1);
    static_assert(CM * LDS_FLOATS_PER_CH * 4 <= LDSCAP, "sweep plane exceeds its LDS cap");
};

// One coordinate through the transport halo (halo_src, axis by axis): resolved index
// and whether it was outside [0, n).
__device__ __forceinline__ int halo_axis(int v, int n, int kind, int &out) {
    out = (v < 0) | (v >= n);
    if (kind == 0) return v < 0 ? v + n : (v >= n ? v - n : v);
    return v < 0 ? 0 : (v >= n ? n - 1 : v);
}

// Gradients from three staged z-planes (z-1, z, z+1).  Sobel taps in the reference order
// dz, dy, dx (lenia.h:30-71); tap weights are +-1, +-2, +-4, so each product is exact and
// fma(v, c, g) rounds exactly like the reference's separate multiply and add.  Central
// differences as reintegration.cpp:260-315.
template <int SX, bool PERM>
__device__ __forceinline__ void sobel_planes(const float *const (&pl)[3], int tx, int ty, float (&g)[3]) {
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
                const float v = pl[dz + 1][(tx + dx + 1) + SX * (ty + dy + 1)];
                if (c0 != 0) g[0] = __builtin_fmaf(v, (float)c0, g[0]);
                if (c1 != 0) g[1] = __builtin_fmaf(v, (float)c1, g[1]);
                if (c2 != 0) g[2] = __builtin_fmaf(v, (float)c2, g[2]);
            }
}

template <int SX>
__device__ __forceinline__ void cdiff_planes(const float *const (&pl)[3], int tx, int ty, float (&g)[3]) {
    auto at = [&](int x, int y, int zi) { return pl[zi][(x + 1) + SX * (y + 1)]; };
    g[0] = (at(tx + 1, ty, 1) - at(tx - 1, ty, 1)) * 0.5f;
    g[1] = (at(tx, ty + 1, 1) - at(tx, ty - 1, 1)) * 0.5f;
    g[2] = (at(tx, ty, 2) - at(tx, ty, 0)) * 0.5f;
}

// Gradients, alpha, F, mu, V as a sweep along z: a workgroup owns a TX x TY column of
// cells; S and H_c (one-cell x/y halo) are staged one z-plane per step into a ring of
// three planes, the next plane already in flight, so every source value is read from
// HBM once per column instead of once per tile it borders.
// zseg: planes per workgroup (the depth is split into segments when few worlds would
// leave the GPU idle).
// Thread cap of a gradient-sweep workgroup: four waves, as the transport sweep's.
constexpr int LOCAL_MAXT = 256;
// Waves per SIMD the gradient sweep is compiled for (its registers: at most 80 VGPRs for
// six, 96 for five): for three channels in workgroups of four waves, six waves without the
// in-sweep SRT transform (32^3: 85 -> 80 VGPRs, 2.17 -> 2.03 ms on 4,096 worlds) and five
// with it (a few registers spilled; 1,024 worlds: 32^3 797 -> 681 us, 36^3 1300 -> 1181
// us); workgroups of three waves (24^3, 44 x 24 x 44) ran slower so (372 -> 406 us,
// 1273 -> 1454 us) and keep what the compiler picks.  The plain kernels need fewer
// registers: about 60 without the SRT transform (eight waves as compiled), six waves with
// it, also in workgroups of three waves (85 and 87 -> 80 VGPRs, none spilled; 1,024 worlds:
// 24^3 286 -> 246 us, 44 x 24 x 44 847 -> 756 us).
template <int W, int H, int D, int CM, bool SRTL, bool PLAIN>
constexpr int local_sweep_waves() {
    if (CM == 3 && PLAIN && SRTL) return 6;
    if (CM != 3 || (SweepCfg<W, H, D, CM, LOCAL_MAXT>::THREADS + 63) / 64 != 4) return 1;
    return PLAIN ? 1 : (SRTL ? 5 : 6);
}
// PLAIN: every world of the launch is a plain_local_world (the optional paths compiled out).
template <int W, int H, int D, int CM, bool SRTL, bool PLAIN>
__global__ __launch_bounds__((SweepCfg<W, H, D, CM, LOCAL_MAXT>::THREADS))
__attribute__((amdgpu_waves_per_eu(local_sweep_waves<W, H, D, CM, SRTL, PLAIN>()))) void k_local_sweep(PhaseArgs a, int zseg) {
    using S = Shape<W, H, D>;
    using T = SweepCfg<W, H, D, CM, LOCAL_MAXT>;
    constexpr int F = CM + 1;  // staged fields: S, H_0 .. H_{C-1}
    constexpr int LI = (F * T::SP + T::THREADS - 1) / T::THREADS;
    __shared__ float ring[3 * F * T::SP];
    const int nseg = (D + zseg - 1) / zseg;
    const int z0 = (blockIdx.x % nseg) * zseg, z1 = z0 + zseg < D ? z0 + zseg : D;
    const int wi = blockIdx.x / nseg / T::BLOCKS, b = (blockIdx.x / nseg) % T::BLOCKS;
    const int tx0 = (b % T::XT) * T::TX, ty0 = (b / T::XT) * T::TY;
    // a copy, loaded once before any store (through a reference its fields would be
    // reloaded after every store to the pool, which the compiler must assume aliases it)
    const DevWorld w = a.worlds[a.ids[wi]];
    const int tid = threadIdx.x, C = w.C;
    const int kind = w.sobel_border, faces = (w.quirks & LHQ_FACE_HALO) != 0;
    // the non-blur central-difference copy (reintegration.cpp:179-186) leaves U_S zero
    // from padded index 5P/2 on
    const bool cdcopy = !PLAIN && w.gradient == 0 && !w.blur && (w.quirks & LHQ_CD_COPY);
    const bool blur_src = !PLAIN && w.gradient == 0 && w.blur;
    const float *srcS = blur_src ? a.pool + w.SUMS : S_cur(a, w);
    const float *srcH = a.pool + (blur_src ? w.US : w.Hg);
    const int nitems = (C + 1) * T::SP;
    // per staged item (field f, slot q): one 32-bit element offset of its (x, y) source
    // from the world's first field (every field lies after it) and two bits, readable with
    // z inside / outside the world -- kept small so the sweep holds few registers
    const float *wbase = a.pool + w.A;
    const unsigned offS = (unsigned)(srcS - wbase), offH = (unsigned)(srcH - wbase);
    unsigned off[LI], okb = 0;
#pragma unroll
    for (int k = 0; k < LI; k++) {
        const int item = tid + k * T::THREADS;
        const int f = item / T::SP, q = item % T::SP;
        int ox, oy;
        const int xr = halo_axis(tx0 + q % T::SX - 1, W, kind, ox), yr = halo_axis(ty0 + q / T::SX - 1, H, kind, oy);
        const int m = ox + oy;
        const bool valid = item < nitems;
        const bool ok_in = valid && (m == 0 || (kind != 2 && !(faces && m >= 2)));
        const bool ok_out = valid && kind != 2 && !(faces && m + 1 >= 2);
        okb |= (ok_in ? 1u : 0u) << (2 * k) | (ok_out ? 2u : 0u) << (2 * k);
        off[k] = (f == 0 ? offS : offH + (unsigned)((f > 0 ? f - 1 : 0) * S::N)) + (unsigned)(xr + W * yr);
    }
    float pf[LI];
    auto fetch = [&](int zp) {
        int oz;
        const int zr = halo_axis(zp, D, kind, oz);
        const unsigned zoff = (unsigned)(W * H * zr);
        if (cdcopy) {
            const int pqz = S::PW * S::PH * (zr + PADR);
#pragma unroll
            for (int k = 0; k < LI; k++) {
                const int item = tid + k * T::THREADS;
                const int f = item / T::SP, q = item % T::SP;
                int ox, oy;
                const int xr = halo_axis(tx0 + q % T::SX - 1, W, kind, ox), yr = halo_axis(ty0 + q / T::SX - 1, H, kind, oy);
                const int pqxy = (xr + PADR) + S::PW * (yr + PADR);
                const bool cut = f > 0 && (long long)(pqxy + pqz) + (long long)(f - 1) * S::P >= 5 * S::P / 2;
                const bool ok = ((okb >> (2 * k + (oz ? 1 : 0))) & 1u) && !cut;
                pf[k] = ok ? el(wbase, off[k] + zoff) : 0.0f;
            }
        } else {
#pragma unroll
            for (int k = 0; k < LI; k++) {
                const bool ok = (okb >> (2 * k + (oz ? 1 : 0))) & 1u;
                pf[k] = ok ? el(wbase, off[k] + zoff) : 0.0f;
            }
        }
    };
    auto store = [&](int zp) {
        float *pl = ring + ((zp + 1) % 3) * F * T::SP;
#pragma unroll
        for (int k = 0; k < LI; k++) {
            const int item = tid + k * T::THREADS;
            if (item < nitems) pl[item] = pf[k];  // item = f * SP + q
        }
    };
    fetch(z0 - 1);
    store(z0 - 1);
    fetch(z0);
    store(z0);
    fetch(z0 + 1);
    const int cx = tid % T::TX, cy = tid / T::TX, toff = cx + T::SX * cy;
    const bool perm = (w.quirks & LHQ_SOBEL_PERMUTE) != 0;
    const bool srt = SRTL && srt_local_world(w);
    const long long col = (tx0 + cx) + (long long)W * (ty0 + cy);
    LocalIn<CM, SRTL> in;
    local_load<W, H, D, CM, SRTL>(a, w, col + (long long)W * H * z0, in);
    for (int z = z0; z < z1; z++) {
        lds_barrier();  // the plane slot about to be refilled is no longer read
        store(z + 1);
        if (z + 2 <= z1) fetch(z + 2);
        lds_barrier();
        // element offsets of the three planes (z - 1, z, z + 1) in the ring
        const int r0 = (z % 3) * F * T::SP, r1 = ((z + 1) % 3) * F * T::SP, r2 = ((z + 2) % 3) * F * T::SP;
        const LocalIn<CM, SRTL> cur = in;
        if (z + 1 < z1) local_load<W, H, D, CM, SRTL>(a, w, col + (long long)W * H * (z + 1), in);  // next plane
        float tt[CM];
        local_cell<W, H, D, CM, SRTL, PLAIN>(a, w, tx0 + cx, ty0 + cy, z, cur, [&](int f, float (&g)[3]) {
            // the corner of this thread's 3 x 3 neighbourhood in each plane of field f, each
            // offset opaque to the compiler (an empty statement that claims to change it), so
            // that the nine reads of a plane pair up off one address register; without it
            // reads of different planes and fields are paired, each pair needing an address
            // register of its own (one VALU add per pair)
            int o[3] = {r0 + f * T::SP + toff, r1 + f * T::SP + toff, r2 + f * T::SP + toff};
            asm volatile("" : "+v"(o[0]), "+v"(o[1]), "+v"(o[2]));
            const float *const pl[3] = {ring + o[0], ring + o[1], ring + o[2]};
            if (PLAIN || w.gradient == 1) {
                if (perm) sobel_planes<T::SX, true>(pl, 0, 0, g);
                else sobel_planes<T::SX, false>(pl, 0, 0, g);
            } else {
                cdiff_planes<T::SX>(pl, 0, 0, g);
            }
        }, tt);
        if constexpr (SRTL) {
            if (srt) {  // the world's transport is this cell's own: the end-of-step transform here
                TrmIn<CM> ti;
#pragma unroll
                for (int c = 0; c < CM; c++) ti.ga[c] = cur.ga[c];
                ti.s = cur.s;
                trm_store<W, H, D, CM>(a, w, col + (long long)W * H * z, tt, ti);
            }
        }
    }
}

// Launches the gradient sweep of a shape class.  arg: bit 0 the class holds SRT worlds
// moved inside the sweep, bit 1 every world is plain_local_world, maxC << 8, planes per
// workgroup << 16 (0: sweep_zseg).  The plain kernels exist for three channels.
template <int W, int H, int D, int CM, bool SRTL, bool PLAIN>
void launch_local_sweep_cm(const PhaseArgs &a, int arg, hipStream_t st) {
    using SC = SweepCfg<W, H, D, CM, LOCAL_MAXT>;
    const int zs = (arg >> 16) ? (arg >> 16) : sweep_zseg(a.nworlds, SC::BLOCKS, D), ns = (D + zs - 1) / zs;
    hipLaunchKernelGGL((k_local_sweep<W, H, D, CM, SRTL, PLAIN>), dim3(a.nworlds * SC::BLOCKS * ns), dim3(SC::THREADS),
                       0, st, a, zs);
}
template <int W, int H, int D>
void launch_local_sweep(const PhaseArgs &a, int arg, hipStream_t st) {
    const bool c3 = ((arg >> 8) & 0xff) <= 3, srt = arg & 1, plain = arg & 2;
    if (c3 && plain) {
        if (srt) launch_local_sweep_cm<W, H, D, 3, true, true>(a, arg, st);
        else launch_local_sweep_cm<W, H, D, 3, false, true>(a, arg, st);
    } else if (c3 && srt) launch_local_sweep_cm<W, H, D, 3, true, false>(a, arg, st);
    else if (c3) launch_local_sweep_cm<W, H, D, 3, false, false>(a, arg, st);
    else if (srt) launch_local_sweep_cm<W, H, D, LHD_MAX_C, true, false>(a, arg, st);
    else launch_local_sweep_cm<W, H, D, LHD_MAX_C, false, false>(a, arg, st);
}

// ------------------------------------------------------------------------------------
// RT transport for dd = 1 as a sweep along z.  A workgroup owns a TX x TY column of
// destinations through the whole depth.  Step z' stages source plane z' (one-cell x/y
// halo) for every moved channel -- its mass and its nine overlap weights, each computed
// once -- and adds the plane's nine terms per channel to the three destination planes
// it touches: z'+1 (dz = -1), z' (dz = 0), z'-1 (dz = +1).  A destination therefore
// receives its 27 terms in the reference's order (dz, dy, dx), and plane z'-1 is
// complete after step z', when its TRM runs.  The product arr*wx*wy of a source is the
// same for the three destinations and is formed once, its arr*wx once per source when the
// plane is staged.
// ------------------------------------------------------------------------------------
// CM: channels moved, at most (3 or LHD_MAX_C).  Thread cap of a workgroup (sweep_ty picks
// its rows): four waves.  On batches of 1,024 worlds, 36^3 moved from 4 rows to 6 under
// this cap (835 -> 766 us, the gradient sweep 1491 -> 1300 us); caps of 320 and 448 (six
// to eight rows of 44 cells on 44 x 24 x 44, twelve rows on 36^3) ran slower.
constexpr int TRANSPORT_MAXT = 256;
// LDS bytes the transport plane may take per workgroup (dynamic: above 64 KB the launch
// allows it per kernel).
constexpr int TRANSPORT_LDS = 65536;
template <int W, int H, int D, int CM>
using TransportCfg = SweepCfg<W, H, D, CM, TRANSPORT_MAXT, TRANSPORT_LDS>;
// Three channels are compiled for four waves per SIMD (at most 128 VGPRs): left to itself
// the compiler gives 24^3, 36^3 and 44 x 24 x 44 152 VGPRs (three waves) where one spilled
// register keeps four (1,024 worlds: 44 x 24 x 44 746 -> 687 us, 24^3 239 -> 235 us, the
// others unchanged); four channels (168 VGPRs) would spill 23 or more and keep the
// compiler's choice.
template <int W, int H, int D, int CM>
__global__ __launch_bounds__((TransportCfg<W, H, D, CM>::THREADS))
__attribute__((amdgpu_waves_per_eu(CM == 3 ? 4 : 1))) void k_transport_sweep(PhaseArgs a, int zseg) {
    using S = Shape<W, H, D>;
    using T = TransportCfg<W, H, D, CM>;
    extern __shared__ float sw_dyn[];
    // destination planes [z0, z1) of this workgroup; source planes z0-1 .. z1
    const int nseg = (D + zseg - 1) / zseg;
    const int z0 = (blockIdx.x % nseg) * zseg, z1 = z0 + zseg < D ? z0 + zseg : D;
    const int wi = blockIdx.x / nseg / T::BLOCKS, b = (blockIdx.x / nseg) % T::BLOCKS;
    const int tx0 = (b % T::XT) * T::TX, ty0 = (b / T::XT) * T::TY;
    // a copy, loaded once before any store (through a reference its fields would be
    // reloaded after every store to the pool, which the compiler must assume aliases it)
    const DevWorld w = a.worlds[a.ids[wi]];
    if (!sweep_world(w)) return;
    const int tid = threadIdx.x, C = w.C;
    const float *ARR = (w.growth_type == 0 || w.growth_type == 2) ? A_cur(a, w) : a.pool + w.Bk;
    const float *MU = a.pool + w.MU;
    const int faces = (w.quirks & LHQ_FACE_HALO) != 0;
    const int tkind = (w.quirks & LHQ_TRANSPORT_TORUS) ? 0 : w.border;
    const int nrt = (w.quirks & LHQ_RT_3CH) ? (C < 3 ? C : 3) : C;
    // the torus distance wrap as a per-slot shift; see k_transport_tiled
    const bool torus = w.border == 0;
    const float si = w.si, clip = w.clip, sn = w.sn;
    const int nitems = T::SP * nrt;
    const int cx = tid % T::TX, cy = tid / T::TX;  // this thread's destination column
    // z-independent part of every staged item: offsets of its (x, y) source and whether
    // it is readable with z inside / outside the world (halo_src split per axis)
    // (few registers per item: the destination centres slot + 1 - kk are formed from the
    // slot's own centre, exact since both are small integers + 0.5, and the two flags
    // share one word)
    static_assert(2 * T::ITEMS <= 32, "readability flags exceed one word");
    unsigned off[T::ITEMS], okb = 0;
    float xc[T::ITEMS], yc[T::ITEMS], shx[T::ITEMS], shy[T::ITEMS];
#pragma unroll
    for (int k = 0; k < T::ITEMS; k++) {
        const int item = tid + k * T::THREADS;
        const int c = item / T::SP, q = item % T::SP;
        const int x = tx0 + q % T::SX - 1, y = ty0 + q / T::SX - 1;
        int ox, oy;
        const int xr = halo_axis(x, W, tkind, ox), yr = halo_axis(y, H, tkind, oy);
        const int m = ox + oy;
        const bool valid = item < nitems;
        const bool ok_in = valid && (m == 0 || (tkind != 2 && !(faces && m >= 2)));
        const bool ok_out = valid && tkind != 2 && !(faces && m + 1 >= 2);
        okb |= (ok_in ? 1u : 0u) << (2 * k) | (ok_out ? 2u : 0u) << (2 * k);
        off[k] = valid ? (unsigned)(c * 3 * S::N + xr + W * yr) : 0u;  // MU of channel c; mass at c*N
        xc[k] = (float)x + 0.5f;
        yc[k] = (float)y + 0.5f;
        shx[k] = !torus ? 0.0f : (x < 0 ? (float)W : (x >= W ? -(float)W : 0.0f));
        shy[k] = !torus ? 0.0f : (y < 0 ? (float)H : (y >= H ? -(float)H : 0.0f));
    }
    auto fetch = [&](int zp, float (&buf)[T::ITEMS][4]) {
        int oz;
        const int zr = halo_axis(zp, D, tkind, oz);
        const unsigned zoff = (unsigned)(W * H * zr);
#pragma unroll
        for (int k = 0; k < T::ITEMS; k++) {
            const bool ok = (okb >> (2 * k + (oz ? 1 : 0))) & 1u;
            const unsigned o = off[k] + zoff;
            const int c = (tid + k * T::THREADS) / T::SP;
#pragma unroll
            for (int d = 0; d < 3; d++) buf[k][d] = ok ? el(MU, o + d * S::N) : 0.0f;
            buf[k][3] = ok ? el(ARR, o - 2 * c * S::N) : 0.0f;
        }
    };
    float accP[CM], accC[CM], accN[CM];
#pragma unroll
    for (int c = 0; c < CM; c++) {
        accP[c] = 0.0f;
        accC[c] = 0.0f;
        accN[c] = 0.0f;
    }
    TrmIn<CM> pre;  // TRM inputs of destination plane z'-1, loaded during step z'-1
    // LDS per channel (9 * SP floats): the z weights of a source toward destinations z'+1
    // and z' as one float2, its z weight toward z'-1, then its three x weights each
    // multiplied by its mass (arr * wx, the first product of the reference's
    // arr * wx * wy * wz * sn, formed once per source instead of once per destination) and
    // its three y weights, as separate arrays.  The nine source slots of this
    // thread's destination each have an index opaque to the compiler (an empty statement
    // that claims to change it), so reads of neighbouring slots are not paired with each
    // other (which needs an address register per array and pair); every read is a single
    // load at a constant offset from its slot's index
    constexpr int ZA = 0, ZB = 2 * T::SP, WXO = 3 * T::SP, WYO = 6 * T::SP;
    int qs[9];
#pragma unroll
    for (int j = 0; j < 9; j++) {
        qs[j] = (cx + j % 3) + T::SX * (cy + j / 3);
        asm volatile("" : "+v"(qs[j]));
    }
    auto step = [&](int zp, const float (&cur)[T::ITEMS][4], float (&nxt)[T::ITEMS][4]) {
        if (zp < z1) fetch(zp + 1, nxt);
        lds_barrier();  // the previous plane's weights are no longer read
        const float shz = !torus ? 0.0f : (zp < 0 ? (float)D : (zp >= D ? -(float)D : 0.0f));
        const float pz[3] = {(float)(zp + 1) + 0.5f, (float)zp + 0.5f, (float)(zp - 1) + 0.5f};
#pragma unroll
        for (int k = 0; k < T::ITEMS; k++) {
            const int item = tid + k * T::THREADS;
            if (item >= nitems) break;
            const int c = item / T::SP, q = item % T::SP;
            float *Wc = sw_dyn + c * T::LDS_FLOATS_PER_CH;
            const float sh[3] = {shx[k], shy[k], shz};
            float wv[3][3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const float mu = cur[k][d];
#pragma unroll
                for (int kk = 0; kk < 3; kk++) {
                    // destination coordinate slot + 1 - kk, centre (float)(dest) + 0.5
                    const float pos = d == 0 ? xc[k] + (float)(1 - kk) : (d == 1 ? yc[k] + (float)(1 - kk) : pz[kk]);
                    float dv = pos - mu;
                    dv = dv + sh[d];
                    dv = si - __builtin_fabsf(dv);
                    // clamp(dv, 0, clip) in one instruction: for finite dv and clip > 0 the
                    // median of (dv, 0, clip) is min(max(dv, 0), clip) (a zero's sign may
                    // differ, which no sum of these products can observe)
                    wv[d][kk] = __builtin_amdgcn_fmed3f(dv, 0.0f, clip);
                }
            }
#pragma unroll
            for (int kk = 0; kk < 3; kk++) {
                Wc[WXO + kk * T::SP + q] = cur[k][3] * wv[0][kk];  // mass * x weight
                Wc[WYO + kk * T::SP + q] = wv[1][kk];
            }
            reinterpret_cast<float2 *>(Wc + ZA)[q] = make_float2(wv[2][0], wv[2][1]);
            Wc[ZB + q] = wv[2][2];
        }
        lds_barrier();
#pragma unroll
        for (int c = 0; c < CM; c++) {
            if (c >= nrt) break;
            // keep the scheduler from hoisting every channel's LDS reads at once (register
            // pressure); an empty statement with a memory clobber, no instruction
            asm volatile("" ::: "memory");
            const float *Wc = sw_dyn + c * T::LDS_FLOATS_PER_CH;
#pragma unroll
            for (int dy = -1; dy <= 1; dy++)
#pragma unroll
                for (int dx = -1; dx <= 1; dx++) {
                    const int q = qs[(dx + 1) + 3 * (dy + 1)];
                    const float2 za = reinterpret_cast<const float2 *>(Wc + ZA)[q];  // z weights: z'+1, z'
                    const float zb = Wc[ZB + q];  // z weight z'-1
                    float base = Wc[WXO + (dx + 1) * T::SP + q];  // mass * x weight
                    base = base * Wc[WYO + (dy + 1) * T::SP + q];
                    float t = base * za.x;  // destination z'+1 sees this plane at dz = -1
                    t = t * sn;
                    accN[c] = accN[c] + t;
                    t = base * za.y;        // destination z', dz = 0
                    t = t * sn;
                    accC[c] = accC[c] + t;
                    t = base * zb;          // destination z'-1, dz = +1
                    t = t * sn;
                    accP[c] = accP[c] + t;
                }
            // this channel's sums are complete here: the compiler may not sink their
            // arithmetic below the next channel's LDS reads (which would keep two
            // channels' staged values live at once)
            asm volatile("" : "+v"(accN[c]), "+v"(accC[c]), "+v"(accP[c]));
        }
        const long long col = (tx0 + cx) + (long long)W * (ty0 + cy);
        if (zp >= z0 + 1) trm_store<W, H, D, CM>(a, w, col + (long long)W * H * (zp - 1), accP, pre);
        if (zp >= z0 && zp < z1) trm_load<W, H, D, CM>(a, w, col + (long long)W * H * zp, pre);  // used next step
#pragma unroll
        for (int c = 0; c < CM; c++) {
            accP[c] = accC[c];
            accC[c] = accN[c];
            accN[c] = 0.0f;
        }
    };
    float cur[T::ITEMS][4], nxt[T::ITEMS][4];
    fetch(z0 - 1, cur);
    for (int zp = z0 - 1; zp <= z1; zp++) {
        step(zp, cur, nxt);
#pragma unroll
        for (int k = 0; k < T::ITEMS; k++)
#pragma unroll
            for (int d = 0; d < 4; d++) cur[k][d] = nxt[k][d];
    }
}

// Launches the transport sweep of a shape class for worlds of at most maxc channels
// (CM = 3 or LHD_MAX_C); zarg: planes per workgroup (0: sweep_zseg).
template <int W, int H, int D, int CM>
void launch_transport_sweep(const PhaseArgs &a, int maxc, int zarg, hipStream_t st) {
    using SC = TransportCfg<W, H, D, CM>;
    const size_t lds = (size_t)maxc * SC::LDS_FLOATS_PER_CH * sizeof(float);
    if constexpr (CM * SC::LDS_FLOATS_PER_CH * 4 > 65536) {
        static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void *>(k_transport_sweep<W, H, D, CM>),
                                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                                           CM * SC::LDS_FLOATS_PER_CH * 4);
        (void)attr;
    }
    const int zs = zarg ? zarg : sweep_zseg(a.nworlds, SC::BLOCKS, D), ns = (D + zs - 1) / zs;
    hipLaunchKernelGGL((k_transport_sweep<W, H, D, CM>), dim3(a.nworlds * SC::BLOCKS * ns), dim3(SC::THREADS), lds, st,
                       a, zs);
}

// RT with dd = 1 from LDS: each staged source carries its nine per-axis overlap weights
// (toward the destinations at offsets -1, 0, +1 along x, y and z), computed once with
// exactly the expression the gather would evaluate for that (destination, source)
// pair, so each destination's 27 products and their summation order are unchanged.
// A slot's halo source is resolved once and all its global loads (mu, source mass for
// every channel) are issued before any weight is formed.
// OSC: compiled with the oscillator terms (used when a world of the batch enables one).
template <int W, int H, int D, bool OSC>
__global__ __launch_bounds__(256) void k_transport_tiled(PhaseArgs a, int skip_sweep) {
    using S = Shape<W, H, D>;
    using T = TileCfg<W, H, D>;
    constexpr int SPT = (T::NS + T::THREADS - 1) / T::THREADS;  // staged slots per thread
    __shared__ float s_arr[T::NS];
    __shared__ float s_w[9 * T::NS];
    const int wi = blockIdx.x / T::TILES, tile = blockIdx.x % T::TILES;
    const int tx0 = (tile % T::XT) * T::TX, trest = tile / T::XT;
    const int ty0 = (trest % (H / T::TY)) * T::TY, tz0 = (trest / (H / T::TY)) * T::TZ;
    const DevWorld &w = a.worlds[a.ids[wi]];
    if (skip_sweep && sweep_world(w)) return;  // moved by k_transport_sweep
    if (srt_local_world(w)) return;            // moved by k_local_sweep
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
        // rt_border_torus (reintegration.cpp:733-755) replaces the displacement dv by
        // dv - ext when dv > ext/2 and by dv + ext when dv < -ext/2; rt_border_other never
        // does.  With dd = 1 (ma = 1 - sigma < 1) every weight a destination reads has
        // |dv| < 2 <= ext/2 when its source slot is an unwrapped neighbour and
        // |dv| > ext - 2 >= ext/2 when the slot is a periodic image across the world's
        // face (extents >= 4), so the test reduces to a per-slot, per-axis constant shift:
        // +ext for an image below 0, -ext for one at or past the extent, else +0 (which
        // leaves every dv unchanged, dv never being -0).
        const bool torus = w.border == 0;
        const float ext[3] = {(float)W, (float)H, (float)D};
        const int extent[3] = {W, H, D};
        const float si = w.si, clip = w.clip, sn = w.sn;
        // gather every staged slot's inputs for all channels up front
        float mu_r[SPT][LHD_MAX_C][3], arr_r[SPT][LHD_MAX_C], shift[SPT][3];
        int gpos[SPT][3];
#pragma unroll
        for (int j = 0; j < SPT; j++) {
            const int q = tid + j * T::THREADS;
            const int sx = q % T::SX, sy = (q / T::SX) % T::SY, sz = q / (T::SX * T::SY);
            gpos[j][0] = tx0 + sx - 1;
            gpos[j][1] = ty0 + sy - 1;
            gpos[j][2] = tz0 + sz - 1;
#pragma unroll
            for (int d = 0; d < 3; d++)
                shift[j][d] = !torus ? 0.0f : (gpos[j][d] < 0 ? ext[d] : (gpos[j][d] >= extent[d] ? -ext[d] : 0.0f));
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
            lds_barrier();
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
                        dv = dv + shift[j][d];
                        // abs_ps clears the sign bit; si - |dv| is never -0 or NaN here, so
                        // v_max / v_min give the same bits as _mm_max_ps / _mm_min_ps
                        dv = si - __builtin_fabsf(dv);
                        s_w[(d * 3 + k) * T::NS + q] = __builtin_fminf(__builtin_fmaxf(dv, 0.0f), clip);
                    }
                }
            }
            lds_barrier();
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
            transport_cell<W, H, D>(a, w, tx0 + tx, ty0 + ty, tz0 + tz, tt[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < T::CPT; j++) {
        const int cell = tid + j * T::THREADS;
        if (cell >= T::NT) break;
        const int tx = cell % T::TX, ty = (cell / T::TX) % T::TY, tz = cell / (T::TX * T::TY);
        if (OSC && (w.osc_la || w.osc_flow || w.osc_div_flow)) osc_cell<W, H, D>(a, w, tx0 + tx, ty0 + ty, tz0 + tz, tt[j]);
        trm_cell<W, H, D>(a, w, (tx0 + tx) + (long long)W * ((ty0 + ty) + (long long)H * (tz0 + tz)), tt[j]);
    }
}

