// lh_kernels.hip -- the shapes built into liblenia_hip.so (the phase kernels of
// lh_kernels.hpp / lh_kernels_big.hpp instantiated per shape), the lookup of a shape's
// launchers (built in, or a shape module built on demand: lh_shape_module.hip), and the
// shape-independent kernels (FP64 mass, moments, marching-cubes case index, copy).
#include <dlfcn.h>
#include <unistd.h>

#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lh_kernels.hpp"

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

const ShapeOps *lh_shape_ops_big(int W, int H, int D);

// Shape modules: liblenia_shape_<W>x<H>x<D>.so (lh_shape_module.hip built for one shape by
// lenia_amd/build.py), looked up in $LENIA_SHAPE_DIR and then in shapes/ beside this
// library, loaded on first use and kept for the life of the process.
static std::string module_dir_of_library() {
    Dl_info info;
    if (!dladdr(reinterpret_cast<const void *>(&module_dir_of_library), &info) || !info.dli_fname) return ".";
    std::string p = info.dli_fname;
    const size_t slash = p.rfind('/');
    return slash == std::string::npos ? std::string(".") : p.substr(0, slash);
}

static std::vector<std::string> module_paths(int W, int H, int D) {
    const std::string name = "/liblenia_shape_" + std::to_string(W) + "x" + std::to_string(H) + "x" + std::to_string(D) + ".so";
    std::vector<std::string> out;
    if (const char *env = std::getenv("LENIA_SHAPE_DIR"))
        if (*env) out.push_back(std::string(env) + name);
    out.push_back(module_dir_of_library() + "/shapes" + name);
    return out;
}

bool lh_shape_module_exists(int W, int H, int D) {
    for (const std::string &p : module_paths(W, H, D))
        if (access(p.c_str(), R_OK) == 0) return true;
    return false;
}

static const ShapeOps *lh_shape_module_ops(int W, int H, int D) {
    static std::mutex mu;
    static std::map<std::string, const ShapeOps *> loaded;  // by path; modules are never unloaded
    std::lock_guard<std::mutex> lock(mu);
    for (const std::string &p : module_paths(W, H, D)) {
        auto it = loaded.find(p);
        if (it != loaded.end()) return it->second;
        if (access(p.c_str(), R_OK) != 0) continue;
        void *h = dlopen(p.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        using fn_t = const ShapeOps *(*)();
        fn_t fn = reinterpret_cast<fn_t>(dlsym(h, "lh_module_shape_ops"));
        const ShapeOps *o = fn ? fn() : nullptr;
        if (!o || o->W != W || o->H != H || o->D != D) continue;  // not this shape's module
        loaded[p] = o;
        return o;
    }
    return nullptr;
}

bool lh_shape_builtin(int W, int H, int D) {
    for (const ShapeOps &o : g_ops)
        if (o.W == W && o.H == H && o.D == D) return true;
    return lh_shape_ops_big(W, H, D) != nullptr;
}

const ShapeOps *lh_shape_ops(int W, int H, int D) {
    for (const ShapeOps &o : g_ops)
        if (o.W == W && o.H == H && o.D == D) return &o;
    if (const ShapeOps *o = lh_shape_ops_big(W, H, D)) return o;
    return lh_shape_module_ops(W, H, D);
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

// Mass and first circular moments per world in FP64: out[7w..7w+6] = sum A, then for
// x, y, z: sum A cos(theta), sum A sin(theta), theta = 2 pi (i + 1/2) / extent, A summed
// over channels.  The angle tables are built once per block in LDS.
__global__ __launch_bounds__(256) void k_moments(PhaseArgs a, double *out) {
    const DevWorld &w = a.worlds[blockIdx.x];
    const int W = w.W, H = w.H, D = w.D, C = w.C;
    const long long N = (long long)W * H * D;
    const float *A = a.pool + (a.parity ? w.A2 : w.A);
    __shared__ double tab[2 * (128 + 128 + 128)];
    double *cx = tab, *sx = tab + W, *cy = tab + 2 * W, *sy = cy + H, *cz = cy + 2 * H, *sz = cz + D;
    for (int i = threadIdx.x; i < W + H + D; i += 256) {
        const int ext = i < W ? W : (i < W + H ? H : D);
        const int k = i < W ? i : (i < W + H ? i - W : i - W - H);
        double sv, cv;
        sincospi((2.0 * k + 1.0) / ext, &sv, &cv);
        if (i < W) { cx[k] = cv; sx[k] = sv; }
        else if (i < W + H) { cy[k] = cv; sy[k] = sv; }
        else { cz[k] = cv; sz[k] = sv; }
    }
    __syncthreads();
    double m[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long long i = threadIdx.x; i < N; i += 256) {
        double v = 0.0;
        for (int c = 0; c < C; c++) v += (double)A[c * N + i];
        const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)(i / ((long long)W * H));
        m[0] += v;
        m[1] += v * cx[x]; m[2] += v * sx[x];
        m[3] += v * cy[y]; m[4] += v * sy[y];
        m[5] += v * cz[z]; m[6] += v * sz[z];
    }
    __shared__ double red[7][256];
    for (int k = 0; k < 7; k++) red[k][threadIdx.x] = m[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 7; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 7) out[7 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

void lh_launch_moments(const PhaseArgs &a, int nworlds, double *out, void *st) {
    hipLaunchKernelGGL(k_moments, dim3(nworlds), dim3(256), 0, (hipStream_t)st, a, out);
}

void lh_launch_mass(const PhaseArgs &a, int nworlds, double *out, void *st) {
    hipLaunchKernelGGL(k_mass, dim3(nworlds), dim3(256), 0, (hipStream_t)st, a, out);
}

void lh_launch_mc(const float *f, int W, int H, int D, float iso, int face_only, unsigned char *out, void *st) {
    const long long n = (long long)W * H * D;
    hipLaunchKernelGGL(k_mc_case, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)st, f, W, H, D,
                       iso, face_only, out);
}
