// lh_sample.hip -- batched random worlds on the GPU, after the reference's
// generate_random (lenia.cpp:370-429) and seed_grid_gaussian (lenia.cpp:305-334).
//
// The reference draws from the C library's rand(); here every draw is a pure function
// of (seed, set index, slot) through splitmix64, so a set sampled among millions on the
// GPU is reproduced exactly on the host (or by another run) from its index alone:
//   u32(seed, set, slot) = mix64(seed * G + set * 256 + slot + G) >> 32,  G = 0x9E3779B97F4A7C15
// and the reference's `rand() % n` becomes u32 % n.  Slots: 0 neighbourhood, 1 kernel
// count, 2 + 12 i + {0 ring count, 1..2 ring peaks, 3 m, 4 s, 5 c0, 6 c1} for kernel i,
// 128 + 3 b + axis for the centre of blob b.
//
// Ranges and arithmetic are generate_random's: neighbourhood 1 + u%8 (every kernel's
// radius), kernel count u%10 + 1, rings u%2 + 1 with peaks (float)(u%100)/100,
// m = (float)(100 + u%500)/1000, s = m/10 + (float)(u%200)/1000, h = 1, r_a = 1,
// c0 = c1 = 0 for kernel 0 or a single channel, else u%channels each.  Seeding: five
// blobs with centres u%W, u%H, u%D; every channel of every cell gets
// sum_b exp(-d_b^2 / (2 r^2)) (plain distance, no wrap) with r the neighbourhood; the
// exponential is the canonical one of the growth bell (canon_exp), so the host
// restatement matches bit for bit.  GA, S and V are cleared as clear_buffers does.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lenia_hip.h"
#include "lh_device.hpp"

lh_status lh_report_error(lh_status s, const std::string &msg);

namespace {

__host__ __device__ inline uint32_t gr_u32(uint64_t seed, uint64_t set, int slot) {
    const uint64_t G = 0x9E3779B97F4A7C15ull;
    uint64_t z = seed * G + set * 256ull + (uint64_t)slot + G;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (uint32_t)(z >> 32);
}

constexpr int GR_MAX_KERNELS = 10;  // kernel_count = rand() % 10 + 1
constexpr int GR_BLOBS = 5;         // seed_grid_gaussian's loop
static_assert(GR_MAX_KERNELS <= LH_MAX_K, "a sampled set fits one world");

// One thread per set: kernels[set][LH_MAX_K] (unused entries zeroed), counts[set].
__global__ void k_sample_random(uint64_t seed, long long first, int n, int channels, lh_kernel *kernels, int *counts) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint64_t set = (uint64_t)(first + r);
    const int nb = 1 + (int)(gr_u32(seed, set, 0) % 8u);
    const int kc = (int)(gr_u32(seed, set, 1) % 10u) + 1;
    lh_kernel *out = kernels + (size_t)r * LH_MAX_K;
    for (int i = 0; i < LH_MAX_K; i++) {
        lh_kernel k;
        k.c0 = k.c1 = k.radius = k.beta_size = 0;
        for (int e = 0; e < LH_MAX_BETA; e++) k.b[e] = 0.0f;
        k.r_a = k.m = k.s = k.h = 0.0f;
        if (i < kc) {
            const int base = 2 + 12 * i;
            k.beta_size = (int)(gr_u32(seed, set, base) % 2u) + 1;
            k.radius = nb;
            for (int e = 0; e < k.beta_size; e++) k.b[e] = (float)(gr_u32(seed, set, base + 1 + e) % 100u) / 100.0f;
            k.m = (float)(100u + gr_u32(seed, set, base + 3) % 500u) / 1000.0f;
            k.s = k.m / 10.0f + (float)(gr_u32(seed, set, base + 4) % 200u) / 1000.0f;
            if (channels == 1 || i == 0) {
                k.c0 = 0;
                k.c1 = 0;
            } else {
                k.c0 = (int)(gr_u32(seed, set, base + 5) % (uint32_t)channels);
                k.c1 = (int)(gr_u32(seed, set, base + 6) % (uint32_t)channels);
            }
            k.h = 1.0f;
            k.r_a = 1.0f;
        }
        out[i] = k;
    }
    counts[r] = kc;
}

// Seeds world j of the job list from set first_set + j: one thread per cell.
__global__ void k_seed_random(float *pool, const SeedJob *jobs, int n, int chunks, uint64_t seed, long long first_set) {
    const int j = blockIdx.x / chunks;
    if (j >= n) return;
    const SeedJob jb = jobs[j];
    const int N = jb.W * jb.H * jb.D;
    const int i = (blockIdx.x % chunks) * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint64_t set = (uint64_t)(first_set + j);
    const int r = 1 + (int)(gr_u32(seed, set, 0) % 8u);
    const int x = i % jb.W, y = (i / jb.W) % jb.H, z = i / (jb.W * jb.H);
    const float den = (float)(2 * r * r);
    float a = 0.0f;
    for (int b = 0; b < GR_BLOBS; b++) {
        const int cx = (int)(gr_u32(seed, set, 128 + 3 * b) % (uint32_t)jb.W);
        const int cy = (int)(gr_u32(seed, set, 129 + 3 * b) % (uint32_t)jb.H);
        const int cz = (int)(gr_u32(seed, set, 130 + 3 * b) % (uint32_t)jb.D);
        const int dx = x - cx, dy = y - cy, dz = z - cz;
        const float dist = (float)(dx * dx + dy * dy + dz * dz);
        const float q = -dist / den;
        a = a + (float)canon_exp((double)q);
    }
    for (int c = 0; c < jb.C; c++) {
        pool[jb.A + (long long)c * N + i] = a;
        pool[jb.GA + (long long)c * N + i] = 0.0f;
        for (int d = 0; d < 3; d++) pool[jb.V + (long long)(3 * c + d) * N + i] = 0.0f;
    }
    pool[jb.S + i] = 0.0f;
}

}  // namespace

// Enqueues the seeding of n worlds (jobs on the device) on stream st.
void lh_launch_seed_random(float *pool, const SeedJob *d_jobs, int n, int max_cells, uint64_t seed,
                           long long first_set, hipStream_t st) {
    const int chunks = (max_cells + 255) / 256;
    hipLaunchKernelGGL(k_seed_random, dim3(n * chunks), dim3(256), 0, st, pool, d_jobs, n, chunks, seed, first_set);
}

extern "C" lh_status lh_sample_random(int device, uint64_t seed, int64_t first, int n, int channels,
                                      lh_kernel *kernels, int *kernel_count) {
    if (n < 0 || !kernels || !kernel_count) return lh_report_error(LH_EINVAL, "null argument or n < 0");
    if (channels < 1 || channels > LH_MAX_C) return lh_report_error(LH_EINVAL, "channels must be 1..LH_MAX_C");
    if (n == 0) return LH_OK;
    if (hipSetDevice(device) != hipSuccess) return lh_report_error(LH_EHIP, "hipSetDevice failed");
    lh_kernel *dk = nullptr;
    int *dc = nullptr;
    const size_t kb = (size_t)n * LH_MAX_K * sizeof(lh_kernel), cb = (size_t)n * sizeof(int);
    if (hipMalloc(&dk, kb) != hipSuccess) return lh_report_error(LH_ENOMEM, "hipMalloc (sampled kernels)");
    if (hipMalloc(&dc, cb) != hipSuccess) {
        (void)hipFree(dk);
        return lh_report_error(LH_ENOMEM, "hipMalloc (kernel counts)");
    }
    hipLaunchKernelGGL(k_sample_random, dim3((n + 127) / 128), dim3(128), 0, nullptr, seed, (long long)first, n,
                       channels, dk, dc);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(kernels, dk, kb, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(kernel_count, dc, cb, hipMemcpyDeviceToHost);
    (void)hipFree(dk);
    (void)hipFree(dc);
    if (e != hipSuccess) return lh_report_error(LH_EHIP, std::string("lh_sample_random: ") + hipGetErrorString(e));
    return LH_OK;
}
