/*
 * lenia_oracle_impl.h -- body of the CPU oracle, instantiated twice (float, double)
 * by lenia_oracle_f32.c / lenia_oracle_f64.c with REAL and PFX defined.
 *
 * TEST INFRASTRUCTURE ONLY (see lenia_oracle.h).  Build with -ffp-contract=off and
 * without -ffast-math: every expression below is evaluated exactly as written.
 *
 * Layout mirrors the reference (render.h:291-341): padded per-channel arrays of
 * P = (W+8)(H+8)(D+8) floats with PADDING=4, unpadded arrays of N = W*H*D, x fastest.
 * The reference grid is W x H x W; D is kept separate so non-cubic z extents work.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "lenia_oracle.h"

typedef REAL real;
typedef struct { real r, i; } cpx;

/* Host threads of the loops below (lo_set_threads, default 1).  Only loops whose
 * iterations write disjoint cells run in parallel, each cell's arithmetic in its own
 * fixed order, so results do not depend on the thread count.  Reductions (mass, the
 * kernel normaliser) stay sequential. */
extern int lo_nthreads;
#ifdef _OPENMP
#define LO_PAR _Pragma("omp parallel for schedule(static) num_threads(lo_nthreads) if(lo_nthreads > 1)")
#else
#define LO_PAR
#endif

#define PAD 4
#define NAME(n) PFX(n)

static inline real rmax(real a, real b) { return a > b ? a : b; } /* _mm_max_ps(a,b) */
static inline real rmin(real a, real b) { return a < b ? a : b; } /* _mm_min_ps(a,b) */
static inline real rclamp(real a, real lo, real hi) { return rmin(rmax(a, lo), hi); }
static inline int iemod(int a, int b) { int r = a % b; return r < 0 ? r + b : r; }

static inline cpx cadd(cpx a, cpx b) { cpx c; c.r = a.r + b.r; c.i = a.i + b.i; return c; }
static inline cpx csub(cpx a, cpx b) { cpx c; c.r = a.r - b.r; c.i = a.i - b.i; return c; }
/* a * b + c rounded once (fmaf / fma) */
static inline real rfma(real a, real b, real c) {
    return sizeof(real) == sizeof(float) ? (real)fmaf((float)a, (float)b, (float)c) : (real)fma((double)a, (double)b, (double)c);
}
/* twiddle product with one fused multiply-add per component */
static inline cpx cmul(cpx a, cpx w) {
    cpx c;
    c.r = rfma(a.r, w.r, -(a.i * w.i));
    c.i = rfma(a.r, w.i, a.i * w.r);
    return c;
}

/* ------------------------------------------------------------------------------ */
/* 1-D complex FFT: Stockham autosort, radix plan 8..8,[4],[2],odd primes ascending. */
/* Twiddles W_n^m = (cos(2 pi m/n), -/+ sin(2 pi m/n)) computed in double, rounded  */
/* once to REAL.  Twiddle products and the paired cosine and sine sums of odd radices */
/* (butterfly below) as fused multiply-adds.                                         */
/* ------------------------------------------------------------------------------ */
#define MAXN 256
#define MAXR 32
typedef struct { int n, nst, radix[16]; real c[MAXN], s[MAXN]; } fplan;
static fplan *g_plans[MAXN + 1];

static const fplan *get_plan(int n) {
    if (g_plans[n]) return g_plans[n];
    fplan *f = (fplan *)calloc(1, sizeof(fplan));
    f->n = n;
    int m = n;
    while (m % 8 == 0) { f->radix[f->nst++] = 8; m /= 8; }
    if (m % 4 == 0) { f->radix[f->nst++] = 4; m /= 4; }
    if (m % 2 == 0) { f->radix[f->nst++] = 2; m /= 2; }
    for (int q = 3; m > 1; q += 2)
        while (m % q == 0) { f->radix[f->nst++] = q; m /= q; }
    for (int e = 0; e < n; e++) {
        double a = (2.0 * 3.14159265358979323846 * (double)e) / (double)n;
        f->c[e] = (real)cos(a);
        f->s[e] = (real)sin(a);
    }
    g_plans[n] = f;
    return f;
}

static inline cpx tw(const fplan *f, int e, int inv) {
    cpx w; w.r = f->c[e]; w.i = inv ? f->s[e] : -f->s[e]; return w;
}

static void butterfly(const fplan *f, cpx *v, int R, int inv) {
    if (R == 2) {
        cpx a = v[0], b = v[1];
        v[0] = cadd(a, b); v[1] = csub(a, b);
    } else if (R == 4) {
        cpx a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        cpx c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        v[0] = cadd(a, c); v[2] = csub(a, c);
        cpx p, q; /* p = b - i d (forward), q = b + i d */
        p.r = b.r + d.i; p.i = b.i - d.r;
        q.r = b.r - d.i; q.i = b.i + d.r;
        if (inv) { v[1] = q; v[3] = p; } else { v[1] = p; v[3] = q; }
    } else if (R == 8) {
        /* two 4-point transforms of the even and odd inputs, the odd ones turned by
         * W_8^k (W_8 = (1 -/+ i) c, c = cos(pi/4) from the table; W_8^2 = -/+ i exactly) */
        cpx e[4], o[4], t1, t2, t3;
        for (int r = 0; r < 4; r++) { e[r] = v[2 * r]; o[r] = v[2 * r + 1]; }
        butterfly(f, e, 4, inv);
        butterfly(f, o, 4, inv);
        const real c = f->c[f->n / 8];
        if (inv) {
            t1.r = (o[1].r - o[1].i) * c; t1.i = (o[1].r + o[1].i) * c;
            t2.r = -o[2].i; t2.i = o[2].r;
            t3.r = -((o[3].r + o[3].i) * c); t3.i = (o[3].r - o[3].i) * c;
        } else {
            t1.r = (o[1].r + o[1].i) * c; t1.i = (o[1].i - o[1].r) * c;
            t2.r = o[2].i; t2.i = -o[2].r;
            t3.r = (o[3].i - o[3].r) * c; t3.i = -((o[3].r + o[3].i) * c);
        }
        v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], t1); v[5] = csub(e[1], t1);
        v[2] = cadd(e[2], t2); v[6] = csub(e[2], t2);
        v[3] = cadd(e[3], t3); v[7] = csub(e[3], t3);
    } else {
        /* odd R: inputs paired as s_j = v_j + v_{R-j}, d_j = v_j - v_{R-j} (j = 1..M,
         * M = (R-1)/2); y_0 = v_0 + s_1 + ... + s_M; for k = 1..M the cosine part
         * A_k = v_0 + sum_j s_j cos(2 pi jk/R) and the sine part C_k = sum_j d_j sin(2 pi jk/R),
         * both summed in j order, give y_k = A_k -/+ i C_k and y_{R-k} = A_k +/- i C_k
         * (upper sign forward) */
        cpx y[MAXR], sv[MAXR / 2 + 1], dv[MAXR / 2 + 1];
        const int step = f->n / R, M = (R - 1) / 2;
        for (int j = 1; j <= M; j++) { sv[j] = cadd(v[j], v[R - j]); dv[j] = csub(v[j], v[R - j]); }
        y[0] = v[0];
        for (int j = 1; j <= M; j++) y[0] = cadd(y[0], sv[j]);
        for (int k = 1; k <= M; k++) {
            real ar = v[0].r, ai = v[0].i, cr = 0, ci = 0;
            for (int j = 1; j <= M; j++) {
                const int e = ((j * k) % R) * step;
                const real c = f->c[e], sn = f->s[e];
                ar = rfma(sv[j].r, c, ar);
                ai = rfma(sv[j].i, c, ai);
                if (j == 1) { cr = dv[j].r * sn; ci = dv[j].i * sn; }
                else { cr = rfma(dv[j].r, sn, cr); ci = rfma(dv[j].i, sn, ci); }
            }
            cpx p, q; /* p = A - iC, q = A + iC */
            p.r = ar + ci; p.i = ai - cr;
            q.r = ar - ci; q.i = ai + cr;
            if (inv) { y[k] = q; y[R - k] = p; } else { y[k] = p; y[R - k] = q; }
        }
        for (int k = 0; k < R; k++) v[k] = y[k];
    }
}

static void fft1c(int n, cpx *x, int inv) {
    const fplan *f = get_plan(n);
    cpx buf[MAXN];
    cpx *src = x, *dst = buf;
    int Ns = 1;
    for (int st = 0; st < f->nst; st++) {
        int R = f->radix[st], nR = n / R, tstep = n / (Ns * R);
        for (int j = 0; j < nR; j++) {
            int k = j % Ns;
            cpx v[MAXR];
            for (int r = 0; r < R; r++) v[r] = src[j + r * nR];
            if (k != 0)
                for (int r = 1; r < R; r++) v[r] = cmul(v[r], tw(f, r * k * tstep, inv));
            butterfly(f, v, R, inv);
            int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) dst[base + r * Ns] = v[r];
        }
        Ns *= R;
        cpx *t = src; src = dst; dst = t;
    }
    if (src != x) memcpy(x, src, sizeof(cpx) * (size_t)n);
}

void NAME(fft1)(int n, real *re_im, int inverse) { fft1c(n, (cpx *)re_im, inverse); }

/* rfft3: real f[z][y][x] -> half spectrum F[kz][ky][kx], kx in [0, W/2].
 * x: rows y, y+1 packed as one complex line; y and z: complex lines.  Unnormalised. */
static void rfft3c(int W, int H, int D, const real *in, cpx *out) {
    int Wh = W / 2 + 1;
    LO_PAR
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y += 2) {
            cpx line[MAXN];
            const real *r0 = in + ((size_t)z * H + y) * W, *r1 = r0 + W;
            for (int x = 0; x < W; x++) { line[x].r = r0[x]; line[x].i = r1[x]; }
            fft1c(W, line, 0);
            cpx *o0 = out + ((size_t)z * H + y) * Wh, *o1 = o0 + Wh;
            for (int k = 0; k < Wh; k++) {
                cpx Z = line[k], P = line[(W - k) % W];
                o0[k].r = (Z.r + P.r) * (real)0.5; o0[k].i = (Z.i - P.i) * (real)0.5;
                o1[k].r = (Z.i + P.i) * (real)0.5; o1[k].i = (P.r - Z.r) * (real)0.5;
            }
        }
    LO_PAR
    for (int z = 0; z < D; z++)
        for (int k = 0; k < Wh; k++) {
            cpx line[MAXN];
            for (int y = 0; y < H; y++) line[y] = out[((size_t)z * H + y) * Wh + k];
            fft1c(H, line, 0);
            for (int y = 0; y < H; y++) out[((size_t)z * H + y) * Wh + k] = line[y];
        }
    LO_PAR
    for (int y = 0; y < H; y++)
        for (int k = 0; k < Wh; k++) {
            cpx line[MAXN];
            for (int z = 0; z < D; z++) line[z] = out[((size_t)z * H + y) * Wh + k];
            fft1c(D, line, 0);
            for (int z = 0; z < D; z++) out[((size_t)z * H + y) * Wh + k] = line[z];
        }
}

/* irfft3: inverse z, inverse y, then c2r along x pairing rows (y, y+1).  The DC and
 * Nyquist bins use real parts only.  Unnormalised.  `in` is consumed (overwritten). */
static void irfft3c(int W, int H, int D, cpx *in, real *out) {
    int Wh = W / 2 + 1;
    LO_PAR
    for (int y = 0; y < H; y++)
        for (int k = 0; k < Wh; k++) {
            cpx line[MAXN];
            for (int z = 0; z < D; z++) line[z] = in[((size_t)z * H + y) * Wh + k];
            fft1c(D, line, 1);
            for (int z = 0; z < D; z++) in[((size_t)z * H + y) * Wh + k] = line[z];
        }
    LO_PAR
    for (int z = 0; z < D; z++)
        for (int k = 0; k < Wh; k++) {
            cpx line[MAXN];
            for (int y = 0; y < H; y++) line[y] = in[((size_t)z * H + y) * Wh + k];
            fft1c(H, line, 1);
            for (int y = 0; y < H; y++) in[((size_t)z * H + y) * Wh + k] = line[y];
        }
    LO_PAR
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y += 2) {
            cpx line[MAXN];
            const cpx *a = in + ((size_t)z * H + y) * Wh, *b = a + Wh;
            line[0].r = a[0].r; line[0].i = b[0].r;
            line[W / 2].r = a[W / 2].r; line[W / 2].i = b[W / 2].r;
            for (int k = 1; k < W / 2; k++) {
                line[k].r = a[k].r - b[k].i;     line[k].i = a[k].i + b[k].r;
                line[W - k].r = a[k].r + b[k].i; line[W - k].i = b[k].r - a[k].i;
            }
            fft1c(W, line, 1);
            real *o0 = out + ((size_t)z * H + y) * W, *o1 = o0 + W;
            for (int x = 0; x < W; x++) { o0[x] = line[x].r; o1[x] = line[x].i; }
        }
}

void NAME(rfft3)(int W, int H, int D, const real *in, real *out) { rfft3c(W, H, D, in, (cpx *)out); }

void NAME(irfft3)(int W, int H, int D, const real *in, real *out) {
    size_t nh = (size_t)(W / 2 + 1) * H * D;
    cpx *tmp = (cpx *)malloc(sizeof(cpx) * nh);
    memcpy(tmp, in, sizeof(cpx) * nh);
    irfft3c(W, H, D, tmp, out);
    free(tmp);
}

/* ------------------------------------------------------------------------------ */
/* Kernel field and spectrum: init_kernel lenia.cpp:199-261, bell render.h:674-677.  */
/* ------------------------------------------------------------------------------ */
static real bellr(real x, real m, real s) {
    real q = (x - m) / s;
    double d = (double)q;
    return (real)exp(-(d * d) / 2.0);
}

void NAME(kernel_field)(int W, int H, int D, const lo_kernel *k, real *out) {
    size_t N = (size_t)W * H * D;
    memset(out, 0, sizeof(real) * N);
    real k_cells = 0;
    int ow = W / 2, oh = H / 2, od = D / 2;
    for (int rx = -ow; rx < ow; rx++)
        for (int ry = -oh; ry < oh; ry++)
            for (int rz = -od; rz < od; rz++) {
                real x = (real)fabs((double)rx), y = (real)fabs((double)ry), z = (real)fabs((double)rz);
                real dist = (real)sqrt((double)(x * x + y * y + z * z));
                dist = dist / (real)k->radius * (real)k->beta_size / (real)k->r_a;
                if (dist < (real)k->beta_size) {
                    int fl = (int)floor((double)dist);
                    int index = fl < k->beta_size - 1 ? fl : k->beta_size - 1;
                    real fr = (real)fmod((double)dist, 1.0);
                    real fin = (real)k->b[index] * bellr(fr, (real)0.5f, (real)0.15f);
                    out[iemod(rx, W) + (size_t)iemod(ry, H) * W + (size_t)iemod(rz, D) * W * H] = fin;
                    k_cells += fin;
                }
            }
    for (size_t i = 0; i < N; i++) out[i] /= k_cells;
}

/* Octant of Re(rfft3(field)): [kz 0..D/2][ky 0..H/2][kx 0..W/2]. */
void NAME(kernel_spectrum)(int W, int H, int D, const real *field, real *oct) {
    int Wh = W / 2 + 1, Hh = H / 2 + 1, Dh = D / 2 + 1;
    cpx *spec = (cpx *)malloc(sizeof(cpx) * (size_t)Wh * H * D);
    rfft3c(W, H, D, field, spec);
    for (int z = 0; z < Dh; z++)
        for (int y = 0; y < Hh; y++)
            for (int x = 0; x < Wh; x++)
                oct[((size_t)z * Hh + y) * Wh + x] = spec[((size_t)z * H + y) * Wh + x].r;
    free(spec);
}

/* 5^3 Gaussian blur kernel, precompute_blur_kernel reintegration.cpp:13-78. */
void NAME(blur_field)(int W, int H, int D, real *out) {
    size_t N = (size_t)W * H * D;
    memset(out, 0, sizeof(real) * N);
    real k_cells = 0;
    const float s = 1.5f;
    for (int rx = -2; rx <= 2; rx++)
        for (int ry = -2; ry <= 2; ry++)
            for (int rz = -2; rz <= 2; rz++) {
                float x = fabsf((float)rx), y = fabsf((float)ry), z = fabsf((float)rz);
                float dist = (x * x + y * y + z * z);
                real g = (real)expf(-dist / (s * s * s));
                out[iemod(rx, W) + (size_t)iemod(ry, H) * W + (size_t)iemod(rz, D) * W * H] = g;
                k_cells += g;
            }
    for (size_t i = 0; i < N; i++) out[i] /= k_cells;
}

/* The 256 LUT entries one kernel touches: BELL_LUT[(int)(m*255)][(int)(s*255)][iu],
 * lenia.cpp:93-101 and lenia.h:79-84. */
void NAME(bell_table)(float m, float s, real *out) {
    int im = (int)(m * 255.0f), is = (int)(s * 255.0f);
    for (int iu = 0; iu < 256; iu++) {
        float xf = (float)iu / 255.0f, mf = (float)im / 255.0f, sf = (float)is / 255.0f;
        float q = (xf - mf) / sf;
        double d = (double)q;
        out[iu] = (real)(float)exp(-(d * d) / 2.0);
    }
}

/* ------------------------------------------------------------------------------ */
/* World state and the step.                                                       */
/* ------------------------------------------------------------------------------ */
struct NAME(world) {
    int W, H, D, C, K, PW, PH, PD, Wh, Hh, Dh;
    long N, P, Nh, Noct;
    lo_params p;
    lo_kernel kern[LO_MAX_K];
    real *khat, *lut, *bhat;
    real *GRID, *BACK, *GH, *U_SUM, *U_S, *TT, *ALPHA, *SUM, *SUM_S;
    real *MU[3], *V[3], *F[3];
    real *GA, *PREALPHA, *NA[3], *NU[3], *POS[3], *U;
    cpx *T, *tmp;
};
typedef struct NAME(world) world;

static inline long pidx(const world *w, int x, int y, int z) {
    return (long)(x + PAD) + (long)(y + PAD) * w->PW + (long)(z + PAD) * w->PW * w->PH;
}
static inline long idx3(const world *w, int x, int y, int z) {
    return (long)x + (long)y * w->W + (long)z * w->W * w->H;
}

static real *zalloc(size_t n) { return (real *)calloc(n, sizeof(real)); }

world *NAME(create)(int W, int H, int D, int C, int K, const lo_params *p, const lo_kernel *k) {
    if (W < 2 || H < 2 || D < 2 || (W & 1) || (H & 1) || W > MAXN || H > MAXN || D > MAXN) return NULL;
    if (C < 1 || C > LO_MAX_C || K < 0 || K > LO_MAX_K) return NULL;
    world *w = (world *)calloc(1, sizeof(world));
    w->W = W; w->H = H; w->D = D; w->C = C; w->K = K;
    w->PW = W + 2 * PAD; w->PH = H + 2 * PAD; w->PD = D + 2 * PAD;
    w->Wh = W / 2 + 1; w->Hh = H / 2 + 1; w->Dh = D / 2 + 1;
    w->N = (long)W * H * D; w->P = (long)w->PW * w->PH * w->PD;
    w->Nh = (long)w->Wh * H * D; w->Noct = (long)w->Wh * w->Hh * w->Dh;
    w->p = *p;
    for (int i = 0; i < K; i++) w->kern[i] = k[i];
    (void)get_plan(W); (void)get_plan(H); (void)get_plan(D);  /* the plan cache is filled serially */
    long N = w->N, P = w->P;
    w->GRID = zalloc(4 * P); w->BACK = zalloc(4 * P); w->GH = zalloc(4 * P);
    w->U_SUM = zalloc(4 * P); w->U_S = zalloc(4 * P); w->TT = zalloc(4 * P);
    w->ALPHA = zalloc(4 * P); w->SUM = zalloc(P); w->SUM_S = zalloc(P);
    for (int a = 0; a < 3; a++) {
        w->MU[a] = zalloc(4 * P); w->V[a] = zalloc(4 * P); w->F[a] = zalloc(4 * P);
        w->NA[a] = zalloc(N); w->NU[a] = zalloc(4 * N); w->POS[a] = zalloc(N);
    }
    w->GA = zalloc(4 * N); w->PREALPHA = zalloc(4 * N);
    w->U = zalloc((size_t)(K > 0 ? K : 1) * N);
    w->T = (cpx *)calloc((size_t)C * w->Nh, sizeof(cpx));
    w->tmp = (cpx *)calloc((size_t)w->Nh, sizeof(cpx));
    w->khat = zalloc((size_t)(K > 0 ? K : 1) * w->Noct);
    w->lut = zalloc((size_t)(K > 0 ? K : 1) * 256);
    w->bhat = zalloc(w->Noct);
    for (int z = 0; z < D; z++) /* init_pos_array lenia.cpp:126-136 */
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                long i = idx3(w, x, y, z);
                w->POS[0][i] = (real)x + (real)0.5f;
                w->POS[1][i] = (real)y + (real)0.5f;
                w->POS[2][i] = (real)z + (real)0.5f;
            }
    real *field = zalloc(N);
    for (int i = 0; i < K; i++) {
        NAME(kernel_field)(W, H, D, &w->kern[i], field);
        NAME(kernel_spectrum)(W, H, D, field, w->khat + (size_t)i * w->Noct);
        NAME(bell_table)(w->kern[i].m, w->kern[i].s, w->lut + (size_t)i * 256);
    }
    NAME(blur_field)(W, H, D, field);
    NAME(kernel_spectrum)(W, H, D, field, w->bhat);
    free(field);
    return w;
}

void NAME(destroy)(world *w) {
    if (!w) return;
    free(w->GRID); free(w->BACK); free(w->GH); free(w->U_SUM); free(w->U_S); free(w->TT);
    free(w->ALPHA); free(w->SUM); free(w->SUM_S);
    for (int a = 0; a < 3; a++) { free(w->MU[a]); free(w->V[a]); free(w->F[a]); free(w->NA[a]); free(w->NU[a]); free(w->POS[a]); }
    free(w->GA); free(w->PREALPHA); free(w->U); free(w->T); free(w->tmp);
    free(w->khat); free(w->lut); free(w->bhat);
    free(w);
}

/* Scratch real field for transforms of padded arrays. */
static void gather_interior(const world *w, const real *padded, real *dst) {
    for (int z = 0; z < w->D; z++)
        for (int y = 0; y < w->H; y++)
            for (int x = 0; x < w->W; x++) dst[idx3(w, x, y, z)] = padded[pidx(w, x, y, z)];
}

static void forward_spectra(world *w) {
    real *f = zalloc(w->N);
    for (int c = 0; c < w->C; c++) {
        gather_interior(w, w->GRID + (size_t)c * w->P, f);
        rfft3c(w->W, w->H, w->D, f, w->T + (size_t)c * w->Nh);
    }
    free(f);
}

int NAME(set_field)(world *w, int field, const real *src) {
    long N = w->N, P = w->P;
    switch (field) {
    case LO_F_A:
        for (int c = 0; c < w->C; c++)
            for (int z = 0; z < w->D; z++)
                for (int y = 0; y < w->H; y++)
                    for (int x = 0; x < w->W; x++)
                        w->GRID[c * P + pidx(w, x, y, z)] = src[c * N + idx3(w, x, y, z)];
        forward_spectra(w);
        return 0;
    case LO_F_GA:
        memcpy(w->GA, src, sizeof(real) * (size_t)(w->C * N));
        return 0;
    case LO_F_S:
        for (int z = 0; z < w->D; z++)
            for (int y = 0; y < w->H; y++)
                for (int x = 0; x < w->W; x++) w->SUM[pidx(w, x, y, z)] = src[idx3(w, x, y, z)];
        return 0;
    case LO_F_V:
        for (int c = 0; c < w->C; c++)
            for (int a = 0; a < 3; a++)
                for (int z = 0; z < w->D; z++)
                    for (int y = 0; y < w->H; y++)
                        for (int x = 0; x < w->W; x++)
                            w->V[a][c * P + pidx(w, x, y, z)] = src[((long)c * 3 + a) * N + idx3(w, x, y, z)];
        return 0;
    }
    return -1;
}

static void get_padded(const world *w, const real *arr, int nch, real *dst) {
    for (int c = 0; c < nch; c++) gather_interior(w, arr + (size_t)c * w->P, dst + (size_t)c * w->N);
}

int NAME(get_field)(world *w, int field, real *dst) {
    long N = w->N;
    int C = w->C;
    switch (field) {
    case LO_F_A: get_padded(w, w->GRID, C, dst); return 0;
    case LO_F_GA: memcpy(dst, w->GA, sizeof(real) * (size_t)(C * N)); return 0;
    case LO_F_S: get_padded(w, w->SUM, 1, dst); return 0;
    case LO_F_H: get_padded(w, w->GH, C, dst); return 0;
    case LO_F_B: get_padded(w, w->BACK, C, dst); return 0;
    case LO_F_USUM: get_padded(w, w->U_SUM, C, dst); return 0;
    case LO_F_TT: get_padded(w, w->TT, C, dst); return 0; /* TT is zero after TRM; see step_debug */
    case LO_F_V: case LO_F_MU: case LO_F_F: {
        real **arr = field == LO_F_V ? w->V : field == LO_F_MU ? w->MU : w->F;
        for (int c = 0; c < C; c++)
            for (int a = 0; a < 3; a++)
                gather_interior(w, arr[a] + (size_t)c * w->P, dst + ((size_t)c * 3 + a) * N);
        return 0;
    }
    case LO_F_NA:
        for (int a = 0; a < 3; a++) memcpy(dst + (size_t)a * N, w->NA[a], sizeof(real) * (size_t)N);
        return 0;
    case LO_F_NU:
        for (int c = 0; c < C; c++)
            for (int a = 0; a < 3; a++) memcpy(dst + ((size_t)c * 3 + a) * N, w->NU[a] + (size_t)c * N, sizeof(real) * (size_t)N);
        return 0;
    case LO_F_U:
        for (int k = 0; k < w->K; k++) for (long i = 0; i < N; i++) dst[k * N + i] = w->U[k * N + i] / (real)N;
        return 0;
    }
    return -1;
}

double NAME(mass)(world *w) {
    double m = 0.0;
    for (int c = 0; c < w->C; c++)
        for (int z = 0; z < w->D; z++)
            for (int y = 0; y < w->H; y++)
                for (int x = 0; x < w->W; x++) m += (double)w->GRID[c * w->P + pidx(w, x, y, z)];
    return m;
}

/* fastexp reintegration.h:16-37: degree-7 Taylor polynomial, Horner order. */
static real fastexp(real x) {
    const real a0 = (real)1.0f, a1 = (real)1.0f;
    const real a2 = (real)(1.0f / 2), a3 = (real)(1.0f / 2 / 3), a4 = (real)(1.0f / 2 / 3 / 4);
    const real a5 = (real)(1.0f / 2 / 3 / 4 / 5), a6 = (real)(1.0f / 2 / 3 / 4 / 5 / 6);
    const real a7 = (real)(1.0f / 2 / 3 / 4 / 5 / 6 / 7);
    real r = a7 * x + a6;
    r = r * x + a5; r = r * x + a4; r = r * x + a3; r = r * x + a2; r = r * x + a1; r = r * x + a0;
    return r;
}

static real soft_growth(real g, real h, real dt) {
    real v = g + h * dt;
    v = (real)-4.0f * (v - (real)0.5f);
    v = fastexp(v);
    v = (real)1.0f + v;
    v = (real)1.0f / v;
    return rclamp(v, 0, 1);
}

/* Halo fill of one padded channel.  kind: 0 torus, 1 wall, 2 zero.  faces_only keeps
 * edges and corners untouched (the reference fills only the six faces). */
static void fill_halo(world *w, real *a, int layers, int kind, int faces_only) {
    int W = w->W, H = w->H, D = w->D;
    LO_PAR
    for (int z = -layers; z < D + layers; z++)
        for (int y = -layers; y < H + layers; y++)
            for (int x = -layers; x < W + layers; x++) {
                int ox = x < 0 || x >= W, oy = y < 0 || y >= H, oz = z < 0 || z >= D;
                int m = ox + oy + oz;
                if (m == 0 || (faces_only && m >= 2)) continue;
                real v;
                if (kind == 0) v = a[pidx(w, iemod(x, W), iemod(y, H), iemod(z, D))];
                else if (kind == 1) {
                    int cx = x < 0 ? 0 : x >= W ? W - 1 : x, cy = y < 0 ? 0 : y >= H ? H - 1 : y, cz = z < 0 ? 0 : z >= D ? D - 1 : z;
                    v = a[pidx(w, cx, cy, cz)];
                } else v = 0;
                a[pidx(w, x, y, z)] = v;
            }
}

/* exp for the analytic growth bell: Cody-Waite reduction x = n ln2 + r (fdlibm's split
 * of ln2 into a 33-bit head and a tail), the degree-13 Taylor polynomial of e^r in
 * Horner order with 1/k! rounded once, then scaling by 2^n.  Written out operation by
 * operation so that the HIP kernels evaluate the identical sequence; relative error
 * about 2e-16 (pinned against numpy.exp). */
static double canon_exp(double x) {
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

static real growth_value(const world *w, int k, real u) {
    const lo_kernel *kp = &w->kern[k];
    if (w->p.quirks & LO_Q_GROWTH_LUT) {
        int iu = (int)(u * (real)255.0f);
        iu = iu < 0 ? 0 : iu > 255 ? 255 : iu;
        return w->lut[(size_t)k * 256 + iu];
    }
    /* bell(u, m, s) render.h:674-677: the square and exp in double, as pow() promotes */
    real q = (u - (real)kp->m) / (real)kp->s;
    double d = (double)q;
    return (real)canon_exp(-(d * d) / 2.0);
}

/* Phase FFT + FFT_MT (lenia.cpp:435-594). */
static void phase_growth(world *w) {
    long N = w->N, P = w->P;
    int C = w->C;
    for (int c = 0; c < C; c++)
        for (int z = 0; z < w->D; z++)
            for (int y = 0; y < w->H; y++)
                for (int x = 0; x < w->W; x++) {
                    long q = c * P + pidx(w, x, y, z);
                    w->U_SUM[q] = 0; w->GH[q] = 0;
                }
    for (int a = 0; a < 3; a++) { memset(w->NA[a], 0, sizeof(real) * N); memset(w->NU[a], 0, sizeof(real) * 4 * N); }
    for (int k = 0; k < w->K; k++) {
        const real *T = (const real *)(w->T + (size_t)w->kern[k].c0 * w->Nh);
        const real *kh = w->khat + (size_t)k * w->Noct;
        real *G = (real *)w->tmp;
        LO_PAR
        for (int z = 0; z < w->D; z++) {
            int zz = z <= w->D / 2 ? z : w->D - z;
            for (int y = 0; y < w->H; y++) {
                int yy = y <= w->H / 2 ? y : w->H - y;
                for (int x = 0; x < w->Wh; x++) {
                    long s = ((long)z * w->H + y) * w->Wh + x;
                    real kv = kh[((long)zz * w->Hh + yy) * w->Wh + x];
                    G[2 * s] = T[2 * s] * kv;
                    G[2 * s + 1] = T[2 * s + 1] * kv;
                }
            }
        }
        irfft3c(w->W, w->H, w->D, w->tmp, w->U + (size_t)k * N);
    }
    for (int k = 0; k < w->K; k++) {
        int c = w->kern[k].c1;
        real h = (real)w->kern[k].h;
        LO_PAR
        for (int z = 0; z < w->D; z++)
            for (int y = 0; y < w->H; y++)
                for (int x = 0; x < w->W; x++) {
                    long i = idx3(w, x, y, z), q = c * P + pidx(w, x, y, z);
                    real u = w->U[k * N + i] / (real)N;
                    w->U_SUM[q] += u;
                    if (w->p.asymptotic[c]) w->GH[q] += (growth_value(w, k, u) - w->GRID[q]) * h;
                    else w->GH[q] += growth_value(w, k, u) * h;
                }
    }
    if (w->p.growth_type == LO_GROWTH_RT) {
        real dt = (real)w->p.dt;
        for (int c = 0; c < C; c++)
            LO_PAR
            for (int z = 0; z < w->D; z++)
                for (int y = 0; y < w->H; y++)
                    for (int x = 0; x < w->W; x++) {
                        long q = c * P + pidx(w, x, y, z);
                        if (w->p.soft_clip[c]) w->BACK[q] = soft_growth(w->GRID[q], w->GH[q], dt);
                        else w->BACK[q] = rclamp(w->GRID[q] + w->GH[q] * dt, 0, 1);
                    }
    }
}

/* Phase BLUR (reintegration.cpp:89-188) and the central-diff zeroing (render.cpp:399-403). */
static void phase_blur(world *w) {
    long N = w->N, P = w->P;
    int C = w->C;
    if (w->p.gradient != LO_GRAD_CENTRAL) return;
    if (w->p.sobel_border == LO_BORDER_NONE) {
        memset(w->U_S, 0, sizeof(real) * 4 * P);
        memset(w->SUM_S, 0, sizeof(real) * P);
    }
    if (w->p.blur) {
        real *f = zalloc(N), *o = zalloc(N);
        for (int c = 0; c <= C; c++) {
            const real *src = c < C ? w->GH + (size_t)c * P : w->SUM;
            real *dst = c < C ? w->U_S + (size_t)c * P : w->SUM_S;
            gather_interior(w, src, f);
            rfft3c(w->W, w->H, w->D, f, w->tmp);
            real *G = (real *)w->tmp;
            LO_PAR
            for (int z = 0; z < w->D; z++) {
                int zz = z <= w->D / 2 ? z : w->D - z;
                for (int y = 0; y < w->H; y++) {
                    int yy = y <= w->H / 2 ? y : w->H - y;
                    for (int x = 0; x < w->Wh; x++) {
                        long s = ((long)z * w->H + y) * w->Wh + x;
                        real kv = w->bhat[((long)zz * w->Hh + yy) * w->Wh + x];
                        G[2 * s] = G[2 * s] * kv;
                        G[2 * s + 1] = G[2 * s + 1] * kv;
                    }
                }
            }
            irfft3c(w->W, w->H, w->D, w->tmp, o);
            LO_PAR
            for (int z = 0; z < w->D; z++)
                for (int y = 0; y < w->H; y++)
                    for (int x = 0; x < w->W; x++) dst[pidx(w, x, y, z)] = o[idx3(w, x, y, z)] / (real)N;
        }
        free(f); free(o);
    } else if (w->p.quirks & LO_Q_CD_COPY) {
        /* Two overlapping memcpys of (P/2)*MAX_CHANNELS floats: U_S[0, 5P/2) <- H[0, 5P/2);
         * the in-bounds part of the SUM copy is SUM_S[0, P) <- SUM[0, P). */
        long n = P / 2 * 5;
        if (n > 4 * P) n = 4 * P;
        memcpy(w->U_S, w->GH, sizeof(real) * (size_t)n);
        memcpy(w->SUM_S, w->SUM, sizeof(real) * (size_t)P);
    } else {
        memcpy(w->U_S, w->GH, sizeof(real) * (size_t)(4 * P));
        memcpy(w->SUM_S, w->SUM, sizeof(real) * (size_t)P);
    }
}

/* Phase TSB (transposition.cpp:6-665). */
static void phase_tsb(world *w) {
    int faces = (w->p.quirks & LO_Q_FACE_HALO) != 0;
    int kind = w->p.sobel_border;
    if (w->p.gradient == LO_GRAD_SOBEL) {
        fill_halo(w, w->SUM, 1, kind, faces);
        for (int c = 0; c < w->C; c++) fill_halo(w, w->GH + (size_t)c * w->P, 1, kind, faces);
    } else if (kind != LO_BORDER_NONE) {
        fill_halo(w, w->SUM_S, 1, kind, faces);
        for (int c = 0; c < w->C; c++) fill_halo(w, w->U_S + (size_t)c * w->P, 1, kind, faces);
    }
}

/* 3-tap smoothing w and difference d (lenia.h:30-71). */
static const int SW[3] = {1, 2, 1};
static const int SD[3] = {-1, 0, 1};

/* Accumulates the three Sobel components of padded field f at cell (x,y,z) into out[3],
 * in tap order dz, dy, dx (reintegration.cpp:197-225).  Zero taps are skipped: the
 * accumulator starts at +0 and never becomes -0, so adding +/-0 would not change it. */
static void sobel_at(const world *w, const real *f, int x, int y, int z, real out[3]) {
    int perm = (w->p.quirks & LO_Q_SOBEL_PERMUTE) != 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                real v = f[pidx(w, x + dx, y + dy, z + dz)];
                int gx = SD[dx + 1] * SW[dy + 1] * SW[dz + 1]; /* d/dx */
                int gy = SW[dx + 1] * SD[dy + 1] * SW[dz + 1]; /* d/dy */
                int gz = SW[dx + 1] * SW[dy + 1] * SD[dz + 1]; /* d/dz */
                int c0, c1, c2;
                if (perm) { c0 = gz; c1 = gx; c2 = gy; } else { c0 = gx; c1 = gy; c2 = gz; }
                if (c0) out[0] = out[0] + v * (real)c0;
                if (c1) out[1] = out[1] + v * (real)c1;
                if (c2) out[2] = out[2] + v * (real)c2;
            }
}

static void cdiff_at(const world *w, const real *f, int x, int y, int z, real out[3]) {
    out[0] = (f[pidx(w, x + 1, y, z)] - f[pidx(w, x - 1, y, z)]) * (real)0.5f;
    out[1] = (f[pidx(w, x, y + 1, z)] - f[pidx(w, x, y - 1, z)]) * (real)0.5f;
    out[2] = (f[pidx(w, x, y, z + 1)] - f[pidx(w, x, y, z - 1)]) * (real)0.5f;
}

/* SOBEL_ONE / SOBEL_TWO (lenia.cpp:597-676) and compute_f (lenia.cpp:679-719). */
static void phase_gradients_force(world *w) {
    long N = w->N, P = w->P;
    int C = w->C, sob = w->p.gradient == LO_GRAD_SOBEL;
    const real *S = sob ? w->SUM : w->SUM_S;
    const real *Hf = sob ? w->GH : w->U_S;
    LO_PAR
    for (int z = 0; z < w->D; z++)
        for (int y = 0; y < w->H; y++)
            for (int x = 0; x < w->W; x++) {
                long i = idx3(w, x, y, z);
                real g[3] = {0, 0, 0};
                if (sob) sobel_at(w, S, x, y, z, g); else cdiff_at(w, S, x, y, z, g);
                for (int a = 0; a < 3; a++) w->NA[a][i] = g[a];
                for (int c = 0; c < C; c++) {
                    real h[3] = {0, 0, 0};
                    if (sob) sobel_at(w, Hf + (size_t)c * P, x, y, z, h); else cdiff_at(w, Hf + (size_t)c * P, x, y, z, h);
                    for (int a = 0; a < 3; a++) w->NU[a][c * N + i] = h[a];
                }
            }
    /* alpha: ALPHA = |A - GA[c*P+i]|, PREALPHA[c*P+i] = clamp((A/theta)^2, 0, 1) */
    real ta = (real)w->p.theta_a;
    int qa = (w->p.quirks & LO_Q_PREALPHA_INDEX) != 0;
    for (int c = 0; c < C; c++)
        LO_PAR
        for (int z = 0; z < w->D; z++)
            for (int y = 0; y < w->H; y++)
                for (int x = 0; x < w->W; x++) {
                    long i = idx3(w, x, y, z), q = c * P + pidx(w, x, y, z);
                    long gi = qa ? c * P + i : c * N + i;
                    real ga = gi < 4 * N ? w->GA[gi] : 0;
                    real ar = w->GRID[q] - ga;
                    w->ALPHA[q] = ar < 0 ? -ar : ar;
                    real pg = w->GRID[q] / ta;
                    pg = pg * pg;
                    if (gi < 4 * N) w->PREALPHA[gi] = rclamp(pg, 0, 1);
                }
    for (int c = 0; c < C; c++)
        LO_PAR
        for (int z = 0; z < w->D; z++)
            for (int y = 0; y < w->H; y++)
                for (int x = 0; x < w->W; x++) {
                    long i = idx3(w, x, y, z), q = c * P + pidx(w, x, y, z);
                    real al = w->PREALPHA[c * N + i];
                    real al1 = (real)1.0f - al;
                    for (int a = 0; a < 3; a++) w->F[a][q] = w->NU[a][c * N + i] * al1 - w->NA[a][i] * al;
                }
}

/* MU phase (reintegration.cpp:321-602, compute_vel reintegration.h:40-72). */
static void phase_mu(world *w) {
    long P = w->P;
    int C = w->C;
    real dt = (real)w->p.dt, sig = (real)w->p.sigma;
    real ma = (real)(w->p.dd - w->p.sigma), nma = -ma, damp = (real)w->p.v_damp;
    int wall = w->p.border == LO_BORDER_WALL;
    real lim[3] = {(real)w->W - sig, (real)w->H - sig, (real)w->D - sig};
    for (int c = 0; c < C; c++)
        LO_PAR
        for (int z = 0; z < w->D; z++)
            for (int y = 0; y < w->H; y++)
                for (int x = 0; x < w->W; x++) {
                    long i = idx3(w, x, y, z), q = c * P + pidx(w, x, y, z);
                    if (w->p.mu_vel)
                        for (int a = 0; a < 3; a++) w->V[a][q] = (w->V[a][q] + w->F[a][q]) * damp;
                    for (int a = 0; a < 3; a++) {
                        real v = (w->p.mu_vel ? w->V[a][q] : w->F[a][q]) * dt;
                        v = rmax(v, nma);
                        v = rmin(v, ma);
                        v = w->POS[a][i] + v;
                        if (wall) { v = rmax(v, sig); v = rmin(v, lim[a]); }
                        w->MU[a][q] = v;
                    }
                    if (!w->p.mu_vel)
                        for (int a = 0; a < 3; a++) w->V[a][q] = (w->V[a][q] + w->F[a][q]) * damp;
                }
}

/* TB phase (transposition.cpp:671-876): dd layers of the transport arrays. */
static void phase_tb(world *w) {
    int faces = (w->p.quirks & LO_Q_FACE_HALO) != 0;
    int kind = (w->p.quirks & LO_Q_TRANSPORT_TORUS) ? LO_BORDER_TORUS : w->p.border;
    int layers = (int)w->p.dd;
    if (layers > PAD) layers = PAD;
    for (int c = 0; c < w->C; c++) {
        size_t o = (size_t)c * w->P;
        for (int a = 0; a < 3; a++) {
            fill_halo(w, w->MU[a] + o, layers, kind, faces);
            fill_halo(w, w->V[a] + o, layers, kind, faces);
            fill_halo(w, w->F[a] + o, layers, kind, faces);
        }
        fill_halo(w, w->ALPHA + o, layers, kind, faces);
        fill_halo(w, w->GRID + o, layers, kind, faces);
        fill_halo(w, w->BACK + o, layers, kind, faces);
    }
}

static real rt_norm(const world *w) {
    real sigma = (real)w->p.sigma;
    if (!(w->p.quirks & LO_Q_RT_SNORM)) return (real)1.0f / ((real)8.0f * sigma * sigma * sigma);
    real s2 = (real)1.0f / ((real)4.0f * sigma * sigma);
    real s23 = (real)1.0f / ((real)2.0f * sigma * sigma * sigma);
    return s2 * ((real)1.0f - (real)w->p.s_norm) + ((real)w->p.s_norm * s23);
}

/* REINTEGRATION phase (lenia.cpp:727-918, reintegration.cpp:607-1410). */
static void phase_reintegration(world *w) {
    long P = w->P;
    int C = w->C, W = w->W, Hh = w->H, D = w->D;
    real sigma = (real)w->p.sigma;
    real sn = rt_norm(w);
    real clip = rmin((real)2.0f * sigma, (real)1.0f);
    real si = (real)0.5f + sigma;
    const real *ARR = (w->p.growth_type == LO_GROWTH_NONE || w->p.growth_type == LO_GROWTH_TRANSPOSITION) ? w->GRID : w->BACK;
    int nrt = (w->p.quirks & LO_Q_RT_3CH) ? (C < 3 ? C : 3) : C;
    int mode = w->p.mode;
    if (mode == LO_MODE_SRT) {
        for (int c = 0; c < nrt; c++)
            LO_PAR
            for (int z = 0; z < D; z++)
                for (int y = 0; y < Hh; y++)
                    for (int x = 0; x < W; x++) {
                        long i = idx3(w, x, y, z), q = c * P + pidx(w, x, y, z);
                        real m[3];
                        for (int a = 0; a < 3; a++) {
                            real d = w->POS[a][i] - w->MU[a][q];
                            d = d < 0 ? -d : d;
                            d = si - d;
                            m[a] = rmin(rmax(d, 0), clip);
                        }
                        real wgt = ARR[q] * m[0];
                        wgt = wgt * m[1]; wgt = wgt * m[2]; wgt = wgt * sn;
                        w->TT[q] = w->TT[q] + wgt;
                    }
    } else if (mode == LO_MODE_RT) {
        int torus = w->p.border == LO_BORDER_TORUS;
        int dd = (int)w->p.dd;
        real ext[3] = {(real)W, (real)Hh, (real)D};
        real half[3] = {(real)(W / 2), (real)(Hh / 2), (real)(D / 2)};
        for (int dz = -dd; dz <= dd; dz++)
            for (int dy = -dd; dy <= dd; dy++)
                for (int dx = -dd; dx <= dd; dx++)
                    LO_PAR
                    for (int z = 0; z < D; z++)
                        for (int y = 0; y < Hh; y++)
                            for (int x = 0; x < W; x++) {
                                long i = idx3(w, x, y, z), dp = pidx(w, x, y, z), sp = pidx(w, x + dx, y + dy, z + dz);
                                for (int c = 0; c < nrt; c++) {
                                    long dq = c * P + dp, sq = c * P + sp;
                                    real m[3];
                                    for (int a = 0; a < 3; a++) {
                                        real d = w->POS[a][i] - w->MU[a][sq];
                                        if (torus) {
                                            if (d > half[a]) d = d - ext[a];
                                            if (d < -half[a]) d = d + ext[a];
                                        }
                                        d = d < 0 ? -d : d;
                                        d = si - d;
                                        d = rmax(d, 0);
                                        m[a] = rmin(d, clip);
                                    }
                                    real wgt = ARR[sq] * m[0];
                                    wgt = wgt * m[1]; wgt = wgt * m[2]; wgt = wgt * sn;
                                    w->TT[dq] = w->TT[dq] + wgt;
                                }
                            }
    } else { /* advection: adv_vel / adv_f reintegration.cpp:1112-1410 */
        real dt = (real)w->p.dt, ma = (real)(w->p.dd - w->p.sigma), nma = -ma;
        real ext[3] = {(real)W, (real)Hh, (real)D};
        int wrap = (w->p.quirks & LO_Q_TRANSPORT_TORUS) || w->p.border == LO_BORDER_TORUS;
        real **vel = w->p.adv_vel ? w->V : w->F;
        for (int c = 0; c < C; c++)
            LO_PAR
            for (int z = 0; z < D; z++)
                for (int y = 0; y < Hh; y++)
                    for (int x = 0; x < W; x++) {
                        long i = idx3(w, x, y, z), q = c * P + pidx(w, x, y, z);
                        real X[3], F0[3], F1[3];
                        int I0[3];
                        for (int a = 0; a < 3; a++) {
                            real v = rmin(rmax(vel[a][q] * dt, nma), ma);
                            real p = (w->POS[a][i] - (real)0.5f) - v;
                            if (wrap) {
                                p = p + ext[a];
                                p = p - (real)floor((double)(p / ext[a])) * ext[a];
                            }
                            real fl = (real)floor((double)p);
                            I0[a] = (int)fl;
                            F1[a] = p - fl;
                            F0[a] = (real)1.0f - F1[a];
                            X[a] = p;
                        }
                        (void)X;
                        const real *A = ARR + (size_t)c * P;
                        real acc = 0;
                        for (int corner = 0; corner < 8; corner++) {
                            int bx = corner & 1, by = (corner >> 1) & 1, bz = (corner >> 2) & 1;
                            real v = A[pidx(w, I0[0] + bx, I0[1] + by, I0[2] + bz)];
                            real t = v * (bx ? F1[0] : F0[0]);
                            t = t * (by ? F1[1] : F0[1]);
                            t = t * (bz ? F1[2] : F0[2]);
                            acc = corner == 0 ? t : acc + t;
                        }
                        w->TT[q] = acc;
                    }
    }
    /* optional oscillators lenia.cpp:803-911 */
    if (w->p.osc_la || w->p.osc_flow || w->p.osc_div_flow) {
        for (int c = 0; c < C; c++)
            LO_PAR
            for (int z = 0; z < D; z++)
                for (int y = 0; y < Hh; y++)
                    for (int x = 0; x < W; x++) {
                        long o = (long)c * P;
                        long q = o + pidx(w, x, y, z);
                        long mx1 = o + pidx(w, x + 1, y, z), mx0 = o + pidx(w, x - 1, y, z);
                        long my1 = o + pidx(w, x, y + 1, z), my0 = o + pidx(w, x, y - 1, z);
                        long mz1 = o + pidx(w, x, y, z + 1), mz0 = o + pidx(w, x, y, z - 1);
                        const real *G = w->GRID;
                        real fx1 = G[mx1], fx0 = G[mx0], fy1 = G[my1], fy0 = G[my0], fz1 = G[mz1], fz0 = G[mz0];
                        real ddx = fx1 - fx0, ddy = fy1 - fy0, ddz = fz1 - fz0;
                        if (w->p.osc_la) {
                            real la = fx1 + fx0 + fy1 + fy0 + fz1 + fz0 - (real)6 * G[q];
                            w->TT[q] += la * (real)w->p.osc_la_amt;
                        }
                        real **VF = w->p.osc_vel ? w->V : w->F;
                        if (w->p.osc_flow) {
                            real flow = VF[0][q] * ddx + VF[1][q] * ddy + VF[2][q] * ddz;
                            w->TT[q] += flow * (real)w->p.osc_flow_amt;
                        }
                        if (w->p.osc_div_flow) {
                            real jx = VF[0][mx1] * w->ALPHA[mx1] - VF[0][mx0] * w->ALPHA[mx0];
                            real jy = VF[1][my1] * w->ALPHA[my1] - VF[1][my0] * w->ALPHA[my0];
                            real jz = VF[2][mz1] * w->ALPHA[mz1] - VF[2][mz0] * w->ALPHA[mz0];
                            real divj = jx + jy + jz;
                            w->TT[q] += divj * (real)w->p.osc_div_flow_amt;
                        }
                    }
    }
    if (w->p.quirks & LO_Q_SUM_RESET) memset(w->SUM, 0, sizeof(real) * (size_t)P);
}

/* TRM phase (lenia.cpp:924-1046). */
static void phase_trm(world *w) {
    long N = w->N, P = w->P;
    int C = w->C;
    real dt = (real)w->p.dt;
    int clamp01 = (w->p.quirks & LO_Q_CLAMP01) != 0;
    for (int c = 0; c < C; c++)
        LO_PAR
        for (int z = 0; z < w->D; z++)
            for (int y = 0; y < w->H; y++)
                for (int x = 0; x < w->W; x++) {
                    long q = c * P + pidx(w, x, y, z);
                    if (w->p.growth_type == LO_GROWTH_TRANSPOSITION) {
                        if (w->p.soft_clip[c] && (w->p.quirks & LO_Q_SOFTCLIP_TR)) w->BACK[q] = soft_growth(w->GRID[q], w->GH[q], dt);
                        else if (w->p.soft_clip[c]) w->TT[q] = soft_growth(w->TT[q], w->GH[q], dt);
                        else w->TT[q] = clamp01 ? rclamp(w->TT[q] + w->GH[q] * dt, 0, 1) : w->TT[q] + w->GH[q] * dt;
                    } else if (clamp01) {
                        w->TT[q] = rclamp(w->TT[q], 0, 1);
                    }
                }
    real gtam = (real)1.0f - (real)w->p.gta, gtas = (real)w->p.gta;
    real stam = (real)1.0f - (real)w->p.sta, stas = (real)w->p.sta;
    LO_PAR
    for (int z = 0; z < w->D; z++)
        for (int y = 0; y < w->H; y++)
            for (int x = 0; x < w->W; x++) {
                long i = idx3(w, x, y, z), pi = pidx(w, x, y, z);
                real ttsum = 0;
                for (int c = 0; c < C; c++) {
                    long q = c * P + pi;
                    real tt = w->TT[q];
                    ttsum = ttsum + tt;
                    w->TT[q] = 0;
                    w->GA[c * N + i] = gtam * w->GA[c * N + i] + tt * gtas;
                    w->GRID[q] = tt;
                }
                w->SUM[pi] = stam * w->SUM[pi] + ttsum * stas;
            }
    forward_spectra(w);
}

/* The first nphases phases of one step (1 growth .. 8 TRM); 7 leaves TT readable. */
void NAME(step_phases)(world *w, int nphases) {
    void (*const ph[8])(world *) = {phase_growth, phase_blur, phase_tsb, phase_gradients_force,
                                    phase_mu, phase_tb, phase_reintegration, phase_trm};
    for (int k = 0; k < nphases && k < 8; k++) ph[k](w);
}

void NAME(step)(world *w, int nsteps) {
    for (int s = 0; s < nsteps; s++) {
        phase_growth(w);
        phase_blur(w);
        phase_tsb(w);
        phase_gradients_force(w);
        phase_mu(w);
        phase_tb(w);
        phase_reintegration(w);
        phase_trm(w);
    }
}

/* Marching-cubes case index (mesh.cpp:19-32, 173-185) over a torus-padded field whose
 * edge/corner pads are zero when face_only (transposition.cpp:881-995). */
void NAME(mc_case)(int W, int H, int D, const real *f, real iso, int face_only, uint8_t *out) {
    static const int DX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    static const int DY[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    static const int DZ[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                int v = 0;
                for (int i = 0; i < 8; i++) {
                    int cx = x + DX[i], cy = y + DY[i], cz = z + DZ[i];
                    int m = (cx >= W) + (cy >= H) + (cz >= D);
                    real val = 0;
                    if (m == 0 || m == 1 || !face_only)
                        val = f[(size_t)(cx % W) + (size_t)(cy % H) * W + (size_t)(cz % D) * W * H];
                    if (val > iso) v |= 1 << i;
                }
                out[(size_t)x + (size_t)y * W + (size_t)z * W * H] = (uint8_t)v;
            }
}
