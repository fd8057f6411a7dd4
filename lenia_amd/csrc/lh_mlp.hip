// lh_mlp.hip -- the parameter-proposal MLP: bf16 matrix-core GEMMs with fused
// epilogues, loss, bias gradients, Adam on fp32 master weights, and a device-side
// parameter sampler that reproduces lenia_amd/proposal.py draw for draw.
//
// One GEMM kernel serves every product of training and inference:
//   forward   Z = X W + b, H = gelu(Z)          (A = X [M][K], B = W [K][N])
//   backward  dH = dZ W^T                        (B stored [N][K])
//             dZ_prev = dH * gelu'(Z_prev)       (fused epilogue)
//             dW = H^T dZ                        (A stored [K][M], reduction over the batch)
// Tiles of 64 x 64 per workgroup (four waves of 32 x 32, v_mfma_f32_16x16x32_bf16),
// k steps of 32 staged in LDS as [row][k] with a pitch of 40 halves, so each lane's
// operand fragment is one 16-byte read and 16 lanes hit distinct banks.  Transposed
// operands are transposed while staged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <string>
#include <vector>

#include "../../include/lenia_mlp.h"

lh_status lh_report_error(lh_status s, const std::string &msg);

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int IN = LM_IN, INP = LM_IN_PAD, HID = LM_HIDDEN, OUT = LM_OUT, OUTP = LM_OUT_PAD;
// parameter layout (floats)
constexpr size_t OW1 = 0, OB1 = OW1 + (size_t)INP * HID, OW2 = OB1 + HID, OB2 = OW2 + (size_t)HID * HID,
                 OW3 = OB2 + HID, OB3 = OW3 + (size_t)HID * HID, OW4 = OB3 + HID, OB4 = OW4 + (size_t)HID * OUTP,
                 NPARAM = OB4 + OUTP;

__device__ __forceinline__ unsigned short to_bf16(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }

// GELU, tanh form: 0.5 z (1 + tanh(c (z + 0.044715 z^3))), c = sqrt(2 / pi)
__device__ __forceinline__ float gelu(float z) {
    const float c = 0.7978845608028654f;
    const float t = tanhf(c * (z + 0.044715f * z * z * z));
    return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float z) {
    const float c = 0.7978845608028654f;
    const float t = tanhf(c * (z + 0.044715f * z * z * z));
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

enum { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_DGELU = 3 };

struct GemmArgs {
    const unsigned short *A, *B;
    int M, N, K, lda, ldb, ldc;
    float *C;              // fp32 output (F32, BIAS, DGELU) or saved pre-activation Z (BIAS_GELU, optional)
    const float *bias;     // BIAS_GELU, BIAS
    const float *Z;        // DGELU: pre-activation of the layer below
    unsigned short *H;     // bf16 output (BIAS_GELU: gelu(Z), DGELU: dZ)
};

// C[M][N] = op(A) op(B); TA: A stored [K][M]; TB: B stored [N][K] (else [K][N]).
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    constexpr int P = 40;
    __shared__ __attribute__((aligned(16))) unsigned short sA[64 * P], sB[64 * P];
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        if (!TA) {
            const int r = tid >> 2, c = (tid & 3) * 8;
            *reinterpret_cast<u16x8 *>(&sA[r * P + c]) =
                *reinterpret_cast<const u16x8 *>(g.A + (size_t)(m0 + r) * g.lda + k0 + c);
        } else {
            const int kk = tid >> 3, c = (tid & 7) * 8;
            const u16x8 v = *reinterpret_cast<const u16x8 *>(g.A + (size_t)(k0 + kk) * g.lda + m0 + c);
#pragma unroll
            for (int i = 0; i < 8; i++) sA[(c + i) * P + kk] = v[i];
        }
        if (TB) {
            const int r = tid >> 2, c = (tid & 3) * 8;
            *reinterpret_cast<u16x8 *>(&sB[r * P + c]) =
                *reinterpret_cast<const u16x8 *>(g.B + (size_t)(n0 + r) * g.ldb + k0 + c);
        } else {
            const int kk = tid >> 3, c = (tid & 7) * 8;
            const u16x8 v = *reinterpret_cast<const u16x8 *>(g.B + (size_t)(k0 + kk) * g.ldb + n0 + c);
#pragma unroll
            for (int i = 0; i < 8; i++) sB[(c + i) * P + kk] = v[i];
        }
        __syncthreads();
        u16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            a[i] = *reinterpret_cast<const u16x8 *>(&sA[(wm + i * 16 + (lane & 15)) * P + 8 * (lane >> 4)]);
            b[i] = *reinterpret_cast<const u16x8 *>(&sB[(wn + i * 16 + (lane & 15)) * P + 8 * (lane >> 4)]);
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[i]),
                                                                    __builtin_bit_cast(bf16x8, b[j]), acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + e, n = n0 + wn + j * 16 + (lane & 15);
                const size_t o = (size_t)m * g.ldc + n;
                const float v = acc[i][j][e];
                if (EPI == EPI_F32) {
                    g.C[o] = v;
                } else if (EPI == EPI_BIAS) {
                    g.C[o] = v + g.bias[n];
                } else if (EPI == EPI_BIAS_GELU) {
                    const float z = v + g.bias[n];
                    if (g.C) g.C[o] = z;
                    g.H[o] = to_bf16(gelu(z));
                } else {
                    const float dz = v * gelu_grad(g.Z[o]);
                    g.C[o] = dz;
                    g.H[o] = to_bf16(dz);
                }
            }
}

template <bool TA, bool TB, int EPI>
void gemm(const GemmArgs &g, hipStream_t st) {
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), dim3(g.M / 64, g.N / 64), dim3(256), 0, st, g);
}

// Loss and its gradient per row: out [M][OUTP] fp32, y [M][3]; dout fp32 and bf16
// [M][OUTP] (zero past column 2), row losses into rl[M].  scale = 1 / batch.
__global__ void k_loss(const float *out, const float *y, int M, int nvalid, float scale, float *dout,
                       unsigned short *dout16, float *rl) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    float g[3] = {0.0f, 0.0f, 0.0f}, l = 0.0f;
    if (r < nvalid) {
        const float *o = out + (size_t)r * OUTP;
        const float s = y[r * 3 + 0];
        // BCE with logits: max(x,0) - x s + log(1 + exp(-|x|))
        const float x = o[0];
        l = fmaxf(x, 0.0f) - x * s + log1pf(expf(-fabsf(x)));
        g[0] = 1.0f / (1.0f + expf(-x)) - s;
        const float d1 = o[1] - y[r * 3 + 1], d2 = o[2] - y[r * 3 + 2];
        l += d1 * d1 + d2 * d2;
        g[1] = 2.0f * d1;
        g[2] = 2.0f * d2;
    }
    rl[r] = l;
    // columns 3.. stay zero from lm_create
    for (int c = 0; c < 3; c++) {
        const float v = g[c] * scale;
        dout[(size_t)r * OUTP + c] = v;
        dout16[(size_t)r * OUTP + c] = to_bf16(v);
    }
}

// Column sums of dZ [M][N] fp32 -> db[N]: one block of 1024 threads per 64 columns, 16
// row phases per column, partial sums combined in a fixed order.
__global__ __launch_bounds__(1024) void k_colsum(const float *dz, int M, int N, float *db) {
    __shared__ float red[16][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), part = threadIdx.x >> 6;
    float s0 = 0.0f, s1 = 0.0f;
    int r = part;
    for (; r + 16 < M; r += 32) {
        s0 += dz[(size_t)r * N + c];
        s1 += dz[(size_t)(r + 16) * N + c];
    }
    if (r < M) s0 += dz[(size_t)r * N + c];
    red[part][threadIdx.x & 63] = s0 + s1;
    __syncthreads();
    if (part == 0) {
        float t = 0.0f;
        for (int k = 0; k < 16; k++) t += red[k][threadIdx.x];
        db[c] = t;
    }
}

// Sum of row losses -> *out (single block).
__global__ void k_sum(const float *v, int n, float *out) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Adam on the fp32 master parameters; refreshes their bf16 copies.
__global__ void k_adam(float *p, const float *g, float *m, float *v, unsigned short *p16, size_t n, float lr,
                       float bc1, float bc2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    const float pi = p[i] - lr * mh / (sqrtf(vh) + 1e-8f);
    p[i] = pi;
    p16[i] = to_bf16(pi);
}

__global__ void k_to_bf16(const float *src, unsigned short *dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = to_bf16(src[i]);
}

// Rows of x [*][81] (selected by idx, or consecutive when idx is null) -> X16 [M][128]
// bf16 with zero padding; targets likewise into Y [M][3].  Rows past n are zero.
__global__ void k_gather(const float *x, const float *y, const int *idx, int n, int M, unsigned short *X16,
                         float *Y) {
    const int r = blockIdx.x, c = threadIdx.x;  // INP threads
    const int src = r < n ? (idx ? idx[r] : r) : -1;
    const float v = (src >= 0 && c < IN) ? x[(size_t)src * IN + c] : 0.0f;
    X16[(size_t)r * INP + c] = to_bf16(v);
    if (Y && c < 3) Y[r * 3 + c] = src >= 0 ? y[(size_t)src * 3 + c] : 0.0f;
}

// splitmix64 draws of lenia_amd/proposal.py: u = (mix(seed*G + set*256 + slot + G) >> 40) 2^-24
__device__ __forceinline__ float draw(uint64_t seed, uint64_t set, int slot) {
    const uint64_t G = 0x9E3779B97F4A7C15ull;
    uint64_t z = seed * G + set * 256ull + (uint64_t)slot + G;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (float)(z >> 40) * (1.0f / 16777216.0f);
}

// Features of sets first .. first+n-1 (proposal.features of proposal.sample), bf16
// [M][128] and optionally fp32 [n][81].
__global__ void k_sample(uint64_t seed, long long first, int n, int M, unsigned short *X16, float *x32) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    const uint64_t set = (uint64_t)(first + r);
    float f[IN];
    for (int k = 0; k < 9; k++) {
        const int base = 16 * k;
        const int beta = 1 + min((int)((double)draw(seed, set, base) * 3.0), 2);
        f[6 * k + 0] = (float)(k / 3) / 2.0f;
        f[6 * k + 1] = (float)(k % 3) / 2.0f;
        f[6 * k + 2] = (float)(beta - 1) / 2.0f;
        for (int q = 0; q < 3; q++) {
            const float u = draw(seed, set, base + 1 + q);
            f[6 * k + 3 + q] = beta > q ? u : 0.0f;
        }
        const float m = 0.1f + draw(seed, set, base + 4) * 0.3f;
        const float s = 0.02f + draw(seed, set, base + 5) * 0.08f;
        const float h = 0.2f + draw(seed, set, base + 6) * 0.8f;
        f[54 + 3 * k + 0] = (m - 0.1f) / 0.3f;
        f[54 + 3 * k + 1] = (s - 0.02f) / 0.08f;
        f[54 + 3 * k + 2] = (h - 0.2f) / 0.8f;
    }
    const bool valid = r < n;
    for (int c = 0; c < INP; c++) X16[(size_t)r * INP + c] = to_bf16(valid && c < IN ? f[c] : 0.0f);
    if (x32 && valid)
        for (int c = 0; c < IN; c++) x32[(size_t)r * IN + c] = f[c];
}

__global__ void k_take_col0(const float *out, int n, float *dst) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) dst[r] = out[(size_t)r * OUTP];
}

__global__ void k_take_out(const float *out, int n, float *dst) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        for (int c = 0; c < 3; c++) dst[r * 3 + c] = out[(size_t)r * OUTP + c];
}

struct Mlp {
    int device = 0, cap = 0;
    hipStream_t st = nullptr;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    unsigned short *p16 = nullptr;
    long step = 0;
    // activations for cap rows
    unsigned short *X16 = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr, *dO16 = nullptr, *dA16 = nullptr,
                   *dB16 = nullptr;
    float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *O = nullptr, *dO = nullptr, *dZa = nullptr, *dZb = nullptr,
          *Y = nullptr, *rl = nullptr, *lsum = nullptr, *tmp = nullptr;
    int *idx = nullptr;
    // resident training set
    float *dx = nullptr, *dy = nullptr;
    int ndata = 0;
};

#define MCHK(x)                                                                                          \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return lh_report_error(LH_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

Mlp *as_mlp(void *p) { return static_cast<Mlp *>(p); }

// Forward over M rows (multiple of 64) of X16; saves Z when train.
void forward(Mlp &n, int M, bool train) {
    GemmArgs a{};
    a.M = M;
    a.A = n.X16; a.B = n.p16 + OW1; a.K = INP; a.N = HID; a.lda = INP; a.ldb = HID; a.ldc = HID;
    a.bias = n.p + OB1; a.C = train ? n.Z1 : nullptr; a.H = n.H1;
    gemm<false, false, EPI_BIAS_GELU>(a, n.st);
    a.A = n.H1; a.B = n.p16 + OW2; a.K = HID; a.lda = HID; a.bias = n.p + OB2; a.C = train ? n.Z2 : nullptr; a.H = n.H2;
    gemm<false, false, EPI_BIAS_GELU>(a, n.st);
    a.A = n.H2; a.B = n.p16 + OW3; a.bias = n.p + OB3; a.C = train ? n.Z3 : nullptr; a.H = n.H3;
    gemm<false, false, EPI_BIAS_GELU>(a, n.st);
    a.A = n.H3; a.B = n.p16 + OW4; a.N = OUTP; a.ldb = OUTP; a.ldc = OUTP; a.bias = n.p + OB4; a.C = n.O; a.H = nullptr;
    gemm<false, false, EPI_BIAS>(a, n.st);
}

// Loss gradient -> parameter gradient n.g for M rows (nvalid real), batch mean.
void backward(Mlp &n, int M, int nvalid) {
    hipLaunchKernelGGL(k_loss, dim3((M + 255) / 256), dim3(256), 0, n.st, n.O, n.Y, M, nvalid, 1.0f / nvalid, n.dO,
                       n.dO16, n.rl);
    GemmArgs a{};
    // layer 4: dW4 = H3^T dO, db4, dH3 = dO W4^T -> dZ3
    a.A = n.H3; a.B = n.dO16; a.M = HID; a.N = OUTP; a.K = M; a.lda = HID; a.ldb = OUTP; a.ldc = OUTP; a.C = n.g + OW4;
    gemm<true, false, EPI_F32>(a, n.st);
    hipLaunchKernelGGL(k_colsum, dim3(OUTP / 64), dim3(1024), 0, n.st, n.dO, M, OUTP, n.g + OB4);
    a = GemmArgs{};
    a.A = n.dO16; a.B = n.p16 + OW4; a.M = M; a.N = HID; a.K = OUTP; a.lda = OUTP; a.ldb = OUTP; a.ldc = HID;
    a.Z = n.Z3; a.C = n.dZa; a.H = n.dA16;
    gemm<false, true, EPI_DGELU>(a, n.st);
    // layer 3
    const unsigned short *Hs[3] = {n.X16, n.H1, n.H2};
    const float *Zs[3] = {nullptr, n.Z1, n.Z2};
    const size_t OW[3] = {OW1, OW2, OW3}, OB[3] = {OB1, OB2, OB3};
    float *dz = n.dZa, *dz_next = n.dZb;
    unsigned short *dz16 = n.dA16, *dz16_next = n.dB16;
    for (int l = 2; l >= 0; l--) {
        const int kin = l == 0 ? INP : HID;
        a = GemmArgs{};
        a.A = Hs[l]; a.B = dz16; a.M = kin; a.N = HID; a.K = M; a.lda = kin; a.ldb = HID; a.ldc = HID; a.C = n.g + OW[l];
        gemm<true, false, EPI_F32>(a, n.st);
        hipLaunchKernelGGL(k_colsum, dim3(HID / 64), dim3(1024), 0, n.st, dz, M, HID, n.g + OB[l]);
        if (l == 0) break;
        a = GemmArgs{};
        a.A = dz16; a.B = n.p16 + OW[l]; a.M = M; a.N = HID; a.K = HID; a.lda = HID; a.ldb = HID; a.ldc = HID;
        a.Z = Zs[l]; a.C = dz_next; a.H = dz16_next;
        gemm<false, true, EPI_DGELU>(a, n.st);
        std::swap(dz, dz_next);
        std::swap(dz16, dz16_next);
    }
}

int round64(int v) { return (v + 63) / 64 * 64; }

}  // namespace

extern "C" {

size_t lm_param_count(void) { return NPARAM; }

lh_status lm_create(int device, int max_batch, void **out) {
    if (!out || max_batch < 1) return lh_report_error(LH_EINVAL, "bad argument");
    Mlp *n = new Mlp();
    n->device = device;
    n->cap = round64(max_batch);
    const size_t c = n->cap;
    MCHK(hipSetDevice(device));
    MCHK(hipStreamCreate(&n->st));
    MCHK(hipMalloc(&n->p, NPARAM * 4));
    MCHK(hipMalloc(&n->g, NPARAM * 4));
    MCHK(hipMalloc(&n->m, NPARAM * 4));
    MCHK(hipMalloc(&n->v, NPARAM * 4));
    MCHK(hipMalloc(&n->p16, NPARAM * 2));
    MCHK(hipMemset(n->p, 0, NPARAM * 4));
    MCHK(hipMemset(n->g, 0, NPARAM * 4));
    MCHK(hipMemset(n->m, 0, NPARAM * 4));
    MCHK(hipMemset(n->v, 0, NPARAM * 4));
    MCHK(hipMemset(n->p16, 0, NPARAM * 2));
    MCHK(hipMalloc(&n->X16, c * INP * 2));
    MCHK(hipMalloc(&n->H1, c * HID * 2));
    MCHK(hipMalloc(&n->H2, c * HID * 2));
    MCHK(hipMalloc(&n->H3, c * HID * 2));
    MCHK(hipMalloc(&n->dA16, c * HID * 2));
    MCHK(hipMalloc(&n->dB16, c * HID * 2));
    MCHK(hipMalloc(&n->dO16, c * OUTP * 2));
    MCHK(hipMemset(n->dO16, 0, c * OUTP * 2));
    MCHK(hipMalloc(&n->Z1, c * HID * 4));
    MCHK(hipMalloc(&n->Z2, c * HID * 4));
    MCHK(hipMalloc(&n->Z3, c * HID * 4));
    MCHK(hipMalloc(&n->dZa, c * HID * 4));
    MCHK(hipMalloc(&n->dZb, c * HID * 4));
    MCHK(hipMalloc(&n->O, c * OUTP * 4));
    MCHK(hipMalloc(&n->dO, c * OUTP * 4));
    MCHK(hipMemset(n->dO, 0, c * OUTP * 4));
    MCHK(hipMalloc(&n->Y, c * 3 * 4));
    MCHK(hipMalloc(&n->rl, c * 4));
    MCHK(hipMalloc(&n->lsum, 4));
    MCHK(hipMalloc(&n->tmp, c * 3 * 4));
    MCHK(hipMalloc(&n->idx, c * 4));
    *out = n;
    return LH_OK;
}

lh_status lm_destroy(void *p) {
    Mlp *n = as_mlp(p);
    if (!n) return LH_OK;
    hipSetDevice(n->device);
    hipStreamSynchronize(n->st);
    void *bufs[] = {n->p, n->g, n->m, n->v, n->p16, n->X16, n->H1, n->H2, n->H3, n->dA16, n->dB16, n->dO16, n->Z1,
                    n->Z2, n->Z3, n->dZa, n->dZb, n->O, n->dO, n->Y, n->rl, n->lsum, n->tmp, n->idx, n->dx, n->dy};
    for (void *b : bufs) hipFree(b);
    hipStreamDestroy(n->st);
    delete n;
    return LH_OK;
}

lh_status lm_set_params(void *p, const float *src, size_t count) {
    Mlp *n = as_mlp(p);
    if (!n || !src) return lh_report_error(LH_EINVAL, "null argument");
    if (count != NPARAM) return lh_report_error(LH_EINVAL, "parameter count mismatch");
    MCHK(hipSetDevice(n->device));
    MCHK(hipMemcpy(n->p, src, NPARAM * 4, hipMemcpyHostToDevice));
    MCHK(hipMemset(n->m, 0, NPARAM * 4));
    MCHK(hipMemset(n->v, 0, NPARAM * 4));
    n->step = 0;
    hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, n->st, n->p, n->p16, NPARAM);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(n->st));
    return LH_OK;
}

lh_status lm_get_params(void *p, float *dst, size_t count) {
    Mlp *n = as_mlp(p);
    if (!n || !dst) return lh_report_error(LH_EINVAL, "null argument");
    if (count != NPARAM) return lh_report_error(LH_EINVAL, "parameter count mismatch");
    MCHK(hipSetDevice(n->device));
    MCHK(hipStreamSynchronize(n->st));
    MCHK(hipMemcpy(dst, n->p, NPARAM * 4, hipMemcpyDeviceToHost));
    return LH_OK;
}

lh_status lm_set_data(void *p, const float *x, const float *y, int cnt) {
    Mlp *n = as_mlp(p);
    if (!n || !x || !y || cnt < 1) return lh_report_error(LH_EINVAL, "bad argument");
    MCHK(hipSetDevice(n->device));
    hipFree(n->dx);
    hipFree(n->dy);
    n->dx = n->dy = nullptr;
    MCHK(hipMalloc(&n->dx, (size_t)cnt * IN * 4));
    MCHK(hipMalloc(&n->dy, (size_t)cnt * 3 * 4));
    MCHK(hipMemcpy(n->dx, x, (size_t)cnt * IN * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(n->dy, y, (size_t)cnt * 3 * 4, hipMemcpyHostToDevice));
    n->ndata = cnt;
    return LH_OK;
}

lh_status lm_train(void *p, const int *order, int n_batches, int batch, float lr, float *mean_loss) {
    Mlp *n = as_mlp(p);
    if (!n || !order || n_batches < 1 || batch < 1) return lh_report_error(LH_EINVAL, "bad argument");
    if (!n->ndata) return lh_report_error(LH_ESTATE, "no training data (lm_set_data)");
    if (batch > n->cap) return lh_report_error(LH_EINVAL, "batch exceeds max_batch");
    for (long long i = 0; i < (long long)n_batches * batch; i++)
        if (order[i] < 0 || order[i] >= n->ndata) return lh_report_error(LH_EINVAL, "order index out of range");
    MCHK(hipSetDevice(n->device));
    const int M = round64(batch);
    double total = 0.0;
    for (int b = 0; b < n_batches; b++) {
        MCHK(hipMemcpyAsync(n->idx, order + (size_t)b * batch, batch * 4, hipMemcpyHostToDevice, n->st));
        hipLaunchKernelGGL(k_gather, dim3(M), dim3(INP), 0, n->st, n->dx, n->dy, n->idx, batch, M, n->X16, n->Y);
        forward(*n, M, true);
        backward(*n, M, batch);
        n->step++;
        const float bc1 = 1.0f - powf(0.9f, (float)n->step), bc2 = 1.0f - powf(0.999f, (float)n->step);
        hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, n->st, n->p, n->g, n->m, n->v, n->p16,
                           NPARAM, lr, bc1, bc2);
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, n->st, n->rl, M, n->lsum);
        MCHK(hipGetLastError());
        float l = 0.0f;
        MCHK(hipMemcpyAsync(&l, n->lsum, 4, hipMemcpyDeviceToHost, n->st));
        MCHK(hipStreamSynchronize(n->st));
        total += l / batch;
    }
    if (mean_loss) *mean_loss = (float)(total / n_batches);
    return LH_OK;
}

lh_status lm_gradients(void *p, const float *x, const float *y, int cnt, float *grad, float *loss) {
    Mlp *n = as_mlp(p);
    if (!n || !x || !y || !grad || cnt < 1) return lh_report_error(LH_EINVAL, "bad argument");
    if (cnt > n->cap) return lh_report_error(LH_EINVAL, "n exceeds max_batch");
    MCHK(hipSetDevice(n->device));
    const int M = round64(cnt);
    float *hx = nullptr, *hy = nullptr;
    MCHK(hipMalloc(&hx, (size_t)cnt * IN * 4));
    MCHK(hipMalloc(&hy, (size_t)cnt * 3 * 4));
    MCHK(hipMemcpy(hx, x, (size_t)cnt * IN * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(hy, y, (size_t)cnt * 3 * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_gather, dim3(M), dim3(INP), 0, n->st, hx, hy, (const int *)nullptr, cnt, M, n->X16, n->Y);
    forward(*n, M, true);
    backward(*n, M, cnt);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, n->st, n->rl, M, n->lsum);
    MCHK(hipGetLastError());
    float l = 0.0f;
    MCHK(hipMemcpyAsync(&l, n->lsum, 4, hipMemcpyDeviceToHost, n->st));
    MCHK(hipMemcpyAsync(grad, n->g, NPARAM * 4, hipMemcpyDeviceToHost, n->st));
    MCHK(hipStreamSynchronize(n->st));
    hipFree(hx);
    hipFree(hy);
    if (loss) *loss = l / cnt;
    return LH_OK;
}

lh_status lm_forward(void *p, const float *x, int cnt, float *out) {
    Mlp *n = as_mlp(p);
    if (!n || !x || !out || cnt < 1) return lh_report_error(LH_EINVAL, "bad argument");
    MCHK(hipSetDevice(n->device));
    float *hx = nullptr;
    const int chunk = n->cap;
    MCHK(hipMalloc(&hx, (size_t)chunk * IN * 4));
    for (int lo = 0; lo < cnt; lo += chunk) {
        const int m = std::min(chunk, cnt - lo), M = round64(m);
        MCHK(hipMemcpyAsync(hx, x + (size_t)lo * IN, (size_t)m * IN * 4, hipMemcpyHostToDevice, n->st));
        hipLaunchKernelGGL(k_gather, dim3(M), dim3(INP), 0, n->st, hx, (const float *)nullptr, (const int *)nullptr, m,
                           M, n->X16, (float *)nullptr);
        forward(*n, M, false);
        hipLaunchKernelGGL(k_take_out, dim3((m + 255) / 256), dim3(256), 0, n->st, n->O, m, n->tmp);
        MCHK(hipGetLastError());
        MCHK(hipMemcpyAsync(out + (size_t)lo * 3, n->tmp, (size_t)m * 3 * 4, hipMemcpyDeviceToHost, n->st));
        MCHK(hipStreamSynchronize(n->st));
    }
    hipFree(hx);
    return LH_OK;
}

lh_status lm_score_sampled(void *p, uint64_t seed, long long cnt, float *logits) {
    Mlp *n = as_mlp(p);
    if (!n || !logits || cnt < 1) return lh_report_error(LH_EINVAL, "bad argument");
    MCHK(hipSetDevice(n->device));
    float *d = nullptr;
    MCHK(hipMalloc(&d, (size_t)cnt * 4));
    for (long long lo = 0; lo < cnt; lo += n->cap) {
        const int m = (int)std::min<long long>(n->cap, cnt - lo), M = round64(m);
        hipLaunchKernelGGL(k_sample, dim3((M + 127) / 128), dim3(128), 0, n->st, seed, lo, m, M, n->X16,
                           (float *)nullptr);
        forward(*n, M, false);
        hipLaunchKernelGGL(k_take_col0, dim3((m + 255) / 256), dim3(256), 0, n->st, n->O, m, d + lo);
    }
    MCHK(hipGetLastError());
    MCHK(hipMemcpyAsync(logits, d, (size_t)cnt * 4, hipMemcpyDeviceToHost, n->st));
    MCHK(hipStreamSynchronize(n->st));
    hipFree(d);
    return LH_OK;
}

lh_status lm_sample_features(void *p, uint64_t seed, long long first, int cnt, float *x) {
    Mlp *n = as_mlp(p);
    if (!n || !x || cnt < 1) return lh_report_error(LH_EINVAL, "bad argument");
    if (cnt > n->cap) return lh_report_error(LH_EINVAL, "n exceeds max_batch");
    MCHK(hipSetDevice(n->device));
    const int M = round64(cnt);
    float *d = nullptr;
    MCHK(hipMalloc(&d, (size_t)cnt * IN * 4));
    hipLaunchKernelGGL(k_sample, dim3((M + 127) / 128), dim3(128), 0, n->st, seed, first, cnt, M, n->X16, d);
    MCHK(hipGetLastError());
    MCHK(hipMemcpyAsync(x, d, (size_t)cnt * IN * 4, hipMemcpyDeviceToHost, n->st));
    MCHK(hipStreamSynchronize(n->st));
    hipFree(d);
    return LH_OK;
}

}  // extern "C"
