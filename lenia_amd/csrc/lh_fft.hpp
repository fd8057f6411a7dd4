// lh_fft.hpp -- register-resident 1-D complex FFT for one thread (CDNA4 / gfx950).
//
// One thread owns a whole line of N complex values in VGPRs; every index is a
// compile-time constant after unrolling, so there is no LDS traffic and no dynamic
// register indexing inside a transform.  The schedule is Stockham autosort with the
// radix plan 8,...,8,[4],[2],odd primes ascending, twiddles W_n^e = (cos, -/+sin)(2 pi e/n)
// read from a host-built table (double precision rounded once to fp32), twiddle products
// and odd radices (paired cosine and sine sums in index order, fft_butterfly) as fused
// multiply-adds.  Built with -ffp-contract=off, so each butterfly rounds exactly as
// written; the CPU checker restates the same schedule.
#pragma once
#include <hip/hip_runtime.h>

// Workgroup barrier ordering LDS only: this wave's LDS operations complete, then
// s_barrier.  __syncthreads() fences global memory as well, which waits for every
// outstanding global load and store of the wave -- draining a sweep's prefetched plane and
// its stores at every step.  Used where a workgroup's threads share data through LDS only
// (no thread reads global memory another thread of the launch writes, and no
// global-to-LDS direct loads are in flight across the barrier).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int fft_nst(int n) {
    int m = n, st = 0;
    while (m % 8 == 0) { m /= 8; st++; }
    if (m % 4 == 0) { m /= 4; st++; }
    if (m % 2 == 0) { m /= 2; st++; }
    for (int q = 3; m > 1; q += 2)
        while (m % q == 0) { m /= q; st++; }
    return st;
}

constexpr int fft_radix(int n, int s) {
    int m = n, st = 0;
    while (m % 8 == 0) { if (st == s) return 8; m /= 8; st++; }
    if (m % 4 == 0) { if (st == s) return 4; m /= 4; st++; }
    if (m % 2 == 0) { if (st == s) return 2; m /= 2; st++; }
    for (int q = 3; m > 1; q += 2)
        while (m % q == 0) { if (st == s) return q; m /= q; st++; }
    return 1;
}

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a * w with one fused multiply-add per component: re = fma(a.x, w.x, -(a.y w.y)),
// im = fma(a.x, w.y, a.y w.x) (the checker restates it with fmaf / fma)
__device__ __forceinline__ float2 c_mul(float2 a, float2 w) {
    return make_float2(__builtin_fmaf(a.x, w.x, -(a.y * w.y)), __builtin_fmaf(a.x, w.y, a.y * w.x));
}

// Twiddle tables are read-only for a kernel's lifetime: read through the constant
// address space, so a wave-uniform index (every register-resident transform's) becomes a
// scalar load instead of a vector load per lane.
typedef const __attribute__((address_space(4))) float *tw_const_ptr;
template <bool INV>
__device__ __forceinline__ float2 twiddle(const float2 *__restrict__ tw, int e) {
    const tw_const_ptr c = (tw_const_ptr)tw;
    const float re = c[2 * e], im = c[2 * e + 1];
    return make_float2(re, INV ? im : -im);
}

template <int R, int N, bool INV>
__device__ __forceinline__ void fft_butterfly(float2 (&v)[R], const float2 *__restrict__ tw) {
    if constexpr (R == 2) {
        float2 a = v[0], b = v[1];
        v[0] = c_add(a, b);
        v[1] = c_sub(a, b);
    } else if constexpr (R == 4) {
        float2 a = c_add(v[0], v[2]), b = c_sub(v[0], v[2]);
        float2 c = c_add(v[1], v[3]), d = c_sub(v[1], v[3]);
        v[0] = c_add(a, c);
        v[2] = c_sub(a, c);
        float2 p = make_float2(b.x + d.y, b.y - d.x);  // b - i d
        float2 q = make_float2(b.x - d.y, b.y + d.x);  // b + i d
        v[1] = INV ? q : p;
        v[3] = INV ? p : q;
    } else if constexpr (R == 8) {
        // two 4-point transforms of the even and odd inputs, the odd ones turned by
        // W_8^k (W_8 = (1 -/+ i) c, c = cos(pi/4) from the table; W_8^2 = -/+ i exactly)
        float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        fft_butterfly<4, N, INV>(e, tw);
        fft_butterfly<4, N, INV>(o, tw);
        const float c = twiddle<true>(tw, N / 8).x;
        const float2 t1 = INV ? make_float2((o[1].x - o[1].y) * c, (o[1].x + o[1].y) * c)
                              : make_float2((o[1].x + o[1].y) * c, (o[1].y - o[1].x) * c);
        const float2 t2 = INV ? make_float2(-o[2].y, o[2].x) : make_float2(o[2].y, -o[2].x);
        const float2 t3 = INV ? make_float2(-((o[3].x + o[3].y) * c), (o[3].x - o[3].y) * c)
                              : make_float2((o[3].y - o[3].x) * c, -((o[3].x + o[3].y) * c));
        v[0] = c_add(e[0], o[0]);
        v[4] = c_sub(e[0], o[0]);
        v[1] = c_add(e[1], t1);
        v[5] = c_sub(e[1], t1);
        v[2] = c_add(e[2], t2);
        v[6] = c_sub(e[2], t2);
        v[3] = c_add(e[3], t3);
        v[7] = c_sub(e[3], t3);
    } else {
        // odd R: inputs paired as s_j = v_j + v_{R-j}, d_j = v_j - v_{R-j} (j = 1..M,
        // M = (R-1)/2); y_0 = v_0 + s_1 + ... + s_M; for k = 1..M the cosine part
        // A_k = v_0 + sum_j s_j cos(2 pi jk/R) and the sine part C_k = sum_j d_j sin(2 pi jk/R),
        // both summed in j order by fused multiply-adds, give y_k = A_k -/+ i C_k,
        // y_{R-k} = A_k +/- i C_k (upper sign forward): about a quarter of a direct DFT's
        // products
        constexpr int M = (R - 1) / 2;
        float2 y[R], sv[M + 1], dv[M + 1];
#pragma unroll
        for (int j = 1; j <= M; j++) {
            sv[j] = c_add(v[j], v[R - j]);
            dv[j] = c_sub(v[j], v[R - j]);
        }
        y[0] = v[0];
#pragma unroll
        for (int j = 1; j <= M; j++) y[0] = c_add(y[0], sv[j]);
#pragma unroll
        for (int k = 1; k <= M; k++) {
            float ar = v[0].x, ai = v[0].y, cr = 0.0f, ci = 0.0f;
#pragma unroll
            for (int j = 1; j <= M; j++) {
                const float2 w = twiddle<true>(tw, ((j * k) % R) * (N / R));  // (cos, sin)
                ar = __builtin_fmaf(sv[j].x, w.x, ar);
                ai = __builtin_fmaf(sv[j].y, w.x, ai);
                if (j == 1) {
                    cr = dv[j].x * w.y;
                    ci = dv[j].y * w.y;
                } else {
                    cr = __builtin_fmaf(dv[j].x, w.y, cr);
                    ci = __builtin_fmaf(dv[j].y, w.y, ci);
                }
            }
            const float2 p = make_float2(ar + ci, ai - cr), q = make_float2(ar - ci, ai + cr);  // A -/+ iC
            y[k] = INV ? q : p;
            y[R - k] = INV ? p : q;
        }
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = y[k];
    }
}

template <int N, bool INV, int S, int NS>
__device__ __forceinline__ void fft_stages(float2 (&x)[N], const float2 *__restrict__ tw) {
    if constexpr (S < fft_nst(N)) {
        constexpr int R = fft_radix(N, S);
        constexpr int NR = N / R;
        constexpr int TS = N / (NS * R);
        float2 y[N];
#pragma unroll
        for (int j = 0; j < NR; j++) {
            const int k = j % NS;
            float2 v[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = x[j + r * NR];
            if (k != 0) {
#pragma unroll
                for (int r = 1; r < R; r++) v[r] = c_mul(v[r], twiddle<INV>(tw, r * k * TS));
            }
            fft_butterfly<R, N, INV>(v, tw);
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) y[base + r * NS] = v[r];
        }
#pragma unroll
        for (int i = 0; i < N; i++) x[i] = y[i];
        fft_stages<N, INV, S + 1, NS * R>(x, tw);
    }
}

// In-place unnormalised DFT of x; INV selects the +i exponent.
template <int N, bool INV>
__device__ __forceinline__ void fft_line(float2 (&x)[N], const float2 *__restrict__ tw) {
    fft_stages<N, INV, 0, 1>(x, tw);
}

// Block-cooperative Stockham over LINES lines of N points held in LDS (line l at
// a[l * LDA + i]); b is scratch of the same layout.  Every thread of the block calls it;
// each stage spreads the LINES * N/R butterflies over the block's NT threads (lines
// fastest, so with an odd LDA the lanes of a wave read and write distinct LDS banks) and
// ends with a barrier.  Per butterfly the arithmetic is fft_line's, so a line transformed
// here matches a register-resident transform bit for bit.  The result is left in a.
template <int N, bool INV, int LINES, int LDA, int NT, int S, int NS>
__device__ __forceinline__ void fft_lds_stages(float2 *src, float2 *dst, const float2 *__restrict__ tw) {
    if constexpr (S < fft_nst(N)) {
        constexpr int R = fft_radix(N, S);
        constexpr int NR = N / R;
        constexpr int TS = N / (NS * R);
        for (int t = threadIdx.x; t < LINES * NR; t += NT) {
            const int l = t % LINES, j = t / LINES;
            const int k = j % NS;
            float2 v[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = src[l * LDA + j + r * NR];
            if (k != 0) {
#pragma unroll
                for (int r = 1; r < R; r++) v[r] = c_mul(v[r], twiddle<INV>(tw, r * k * TS));
            }
            fft_butterfly<R, N, INV>(v, tw);
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) dst[l * LDA + base + r * NS] = v[r];
        }
        lds_barrier();
        fft_lds_stages<N, INV, LINES, LDA, NT, S + 1, NS * R>(dst, src, tw);
    }
}

template <int N, bool INV, int LINES, int LDA, int NT>
__device__ __forceinline__ void fft_lds_batch(float2 *a, float2 *b, const float2 *__restrict__ tw) {
    fft_lds_stages<N, INV, LINES, LDA, NT, 0, 1>(a, b, tw);
    if constexpr (fft_nst(N) & 1) {
        for (int t = threadIdx.x; t < LINES * N; t += NT) {
            const int l = t % LINES, i = t / LINES;
            a[l * LDA + i] = b[l * LDA + i];
        }
        lds_barrier();
    }
}
