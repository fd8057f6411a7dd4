// lh_mesh.hip -- marching-cubes meshes of the renderer's three views on the GPU, after
// generate_mesh_v1 / v2 / v3 (mesh.cpp:34-430, 434-1113, 1115-2036).
//
// The reference meshes on CPU threads, each walking a z-section of cubes in order and
// deduplicating vertices through a per-field edge map (V1/V2/V3_EDGES, ETIDX in
// render.h:517).  Its result, for one section spanning the grid, is fixed by the serial
// order of "units" u = ((channel * cubes + cube) * F + field): a vertex belongs to the
// unit and triangle slot that first names its edge, vertices are numbered in that order,
// the index buffer runs in unit order, and a vertex's normal is the sum of the face
// normals of the triangles naming it, in unit order.  Here every rule is data-parallel:
//   count   one thread per channel and cube walks that cube's F units (neighbouring
//           threads read neighbouring cells): case index, triangle count, an atomicMin
//           of the 64-bit key u * 16 + slot on each named edge (the first owner wins),
//           and the unit appended to a list of units with triangles (any order; one
//           global atomic per workgroup);
//   own     per listed unit, the slots whose key won;
//   scan    exclusive prefix sums of index and vertex counts over the units;
//   emit    per listed unit, owned vertices at their ranks (interpolated in the owning cube's
//           corner order, as the reference stores it), edge -> vertex map, and the unit
//           of each triangle;
//   index   one thread per triangle: its indices and face normal at its serial position;
//   normal  per vertex: a gather over the (channel, adjacent cube, triangle) sequence
//           naming it, summed in the reference's order, so the fp32 sums match.
// All interpolation and cross products are written as the reference's expressions
// (no contraction: the build passes -ffp-contract=off).
//
// The triangulation table is an input: the reference passes its own TRIANGLE_TABLE
// (marching_cubes.h:98-355) through lh_mesh; lh_mc_derive builds this library's default
// from the cube's topology (each face's crossings joined, ambiguous faces separating the
// corners above iso, loops oriented with their normal toward the corners at or below
// iso, fan-triangulated from their lowest edge).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lh_common.hpp"

namespace {

// Cube corners (mesh.cpp:19-32) and the corner pairs of the 12 edges
// (marching_cubes.h:19-31).
constexpr int MC_DX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
constexpr int MC_DY[8] = {0, 0, 0, 0, 1, 1, 1, 1};
constexpr int MC_DZ[8] = {0, 0, 1, 1, 0, 0, 1, 1};
constexpr int MC_EA[12] = {0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 2, 3};
constexpr int MC_EB[12] = {1, 2, 3, 3, 5, 6, 7, 7, 4, 5, 6, 7};

// Per-edge facts packed into 64-bit words so that a table-driven (run-time) edge index
// needs no memory lookup: 5 bits per edge = low-corner offset x, y, z and axis (2 bits).
constexpr uint64_t pack_edges() {
    uint64_t p = 0;
    for (int e = 0; e < 12; e++) {
        const int a = MC_EA[e], b = MC_EB[e];
        const int ox = MC_DX[a] < MC_DX[b] ? MC_DX[a] : MC_DX[b];
        const int oy = MC_DY[a] < MC_DY[b] ? MC_DY[a] : MC_DY[b];
        const int oz = MC_DZ[a] < MC_DZ[b] ? MC_DZ[a] : MC_DZ[b];
        const int axis = MC_DX[a] != MC_DX[b] ? 0 : (MC_DY[a] != MC_DY[b] ? 1 : 2);
        p |= (uint64_t)(ox | oy << 1 | oz << 2 | axis << 3) << (5 * e);
    }
    return p;
}
// 3 bits per edge: one of its corners.
constexpr uint64_t pack_corners(const int (&corner)[12]) {
    uint64_t p = 0;
    for (int e = 0; e < 12; e++) p |= (uint64_t)corner[e] << (3 * e);
    return p;
}
constexpr uint64_t EDGE_BITS = pack_edges();
constexpr uint64_t CORNER_A_BITS = pack_corners(MC_EA), CORNER_B_BITS = pack_corners(MC_EB);
constexpr uint32_t DXB = 0b01100110, DYB = 0b11110000, DZB = 0b11001100;  // corner k -> bit k

struct MeshDev {
    MeshJob j;
    const int8_t *tri;     // [256][16]
    const uint8_t *ntri;   // [256] triangles per case
    const uint32_t *colors;
    unsigned long long *first;  // [F * E3] lowest key naming each edge
    uint32_t *vmap;        // [F * E3] edge -> vertex
    uint8_t *ucase;        // [U]
    uint32_t *cnt_i, *cnt_v;  // [U] exclusive offsets of indices and vertices (cnt_v: counts first)
    uint32_t *active;      // [U] units with triangles, in no particular order
    unsigned int *nact;    // their number
    long long U, Nc, E, E3;
};

struct MeshOut {
    float *verts;     // [nv][7] (x, y, z, nx, ny, nz, color bits)
    uint32_t *idx;    // [ni]
    float *fnorm;     // [ni / 3][3] face normal of each triangle
    uint32_t *vedge;  // [nv] edge slot (field * E3 + edge) of each vertex
    uint32_t *tunit;  // [ni / 3] unit of each triangle
};

__device__ __forceinline__ float corner_value(const float *f, const MeshJob &j, int cx, int cy, int cz) {
    const int m = (cx >= j.W) + (cy >= j.H) + (cz >= j.D);
    if (m > 1 && j.face_only) return 0.0f;
    cx = cx >= j.W ? cx - j.W : cx;
    cy = cy >= j.H ? cy - j.H : cy;
    cz = cz >= j.D ? cz - j.D : cz;
    return f[cx + (long long)j.W * (cy + (long long)j.H * cz)];
}

__device__ __forceinline__ void load_corners(const MeshDev &m, int c, int f, int x, int y, int z, float v[8]) {
    const float *fc = m.j.field[f] + (long long)c * m.Nc;
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = corner_value(fc, m.j, x + MC_DX[k], y + MC_DY[k], z + MC_DZ[k]);
}

__device__ __forceinline__ int case_of(const float v[8], float iso) {
    int cs = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) cs |= (v[k] > iso ? 1 : 0) << k;
    return cs;
}

// v[k] for a run-time k without dynamic register indexing.
__device__ __forceinline__ float pick8(const float v[8], int k) {
    float r = v[0];
#pragma unroll
    for (int i = 1; i < 8; i++) r = k == i ? v[i] : r;
    return r;
}

// Edge-map slot of local edge e of cube (x, y, z): axis * E + BIDX of its low corner.
__device__ __forceinline__ long long edge_slot(const MeshDev &m, int x, int y, int z, int e) {
    const uint32_t b = (uint32_t)(EDGE_BITS >> (5 * e)) & 31u;
    const long long px = x + (b & 1), py = y + ((b >> 1) & 1), pz = z + ((b >> 2) & 1);
    return (long long)(b >> 3) * m.E + px + py * (m.j.W + 1) + pz * (long long)(m.j.W + 1) * (m.j.H + 1);
}

// Position of the crossing on local edge e, as mesh.cpp:185-201 computes it.
__device__ __forceinline__ void edge_point(const MeshDev &m, const float v[8], int x, int y, int z, int e,
                                           float &px, float &py, float &pz) {
    const int a = (int)(CORNER_A_BITS >> (3 * e)) & 7, b = (int)(CORNER_B_BITS >> (3 * e)) & 7;
    const float va = pick8(v, a), vb = pick8(v, b);
    const float cell = m.j.cell;
    const float ax = ((float)x + (float)((DXB >> a) & 1)) * cell, bx = ((float)x + (float)((DXB >> b) & 1)) * cell;
    const float ay = ((float)y + (float)((DYB >> a) & 1)) * cell, by = ((float)y + (float)((DYB >> b) & 1)) * cell;
    const float az = ((float)z + (float)((DZB >> a) & 1)) * cell, bz = ((float)z + (float)((DZB >> b) & 1)) * cell;
    const float T = (m.j.iso - va) / (vb - va);
    px = ax + (bx - ax) * T;
    py = ay + (by - ay) * T;
    pz = az + (bz - az) * T;
}

// Thread t = channel * cubes + cube handles the F units t * F + f: consecutive threads
// read consecutive cells of one field.
__device__ __forceinline__ bool cube_coords(const MeshDev &m, long long &t, int &c, int &x, int &y, int &z) {
    t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= m.U / m.j.F) return false;
    const long long cube = t % m.Nc;
    c = (int)(t / m.Nc);
    x = (int)(cube % m.j.W);
    y = (int)((cube / m.j.W) % m.j.H);
    z = (int)(cube / ((long long)m.j.W * m.j.H));
    return true;
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long *p, long long n, unsigned long long v) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = v;
}

__device__ __forceinline__ void unit_cube(const MeshDev &m, long long u, int &c, int &f, int &x, int &y, int &z) {
    f = (int)(u % m.j.F);
    const long long t = u / m.j.F;
    const long long cube = t % m.Nc;
    c = (int)(t / m.Nc);
    x = (int)(cube % m.j.W);
    y = (int)((cube / m.j.W) % m.j.H);
    z = (int)(cube / ((long long)m.j.W * m.j.H));
}

__global__ __launch_bounds__(256) void k_mesh_count(MeshDev m) {
    __shared__ unsigned int wg_n, wg_base;
    if (threadIdx.x == 0) wg_n = 0u;
    __syncthreads();
    long long t;
    int c, x, y, z;
    const bool valid = cube_coords(m, t, c, x, y, z);
    int act = 0;  // bit f: unit t * F + f has triangles
    if (valid) {
        for (int f = 0; f < m.j.F; f++) {
            const long long u = t * m.j.F + f;
            float v[8];
            load_corners(m, c, f, x, y, z, v);
            const int cs = case_of(v, m.j.iso);
            m.ucase[u] = (uint8_t)cs;  // the scan reads index counts from it
            const int ns = 3 * m.ntri[cs];
            const int8_t *row = m.tri + 16 * cs;
            for (int s = 0; s < ns; s++) {
                const long long g = f * m.E3 + edge_slot(m, x, y, z, row[s]);
                atomicMin(m.first + g, (unsigned long long)u * 16ull + (unsigned long long)s);
            }
            if (ns) act |= 1 << f;
        }
    }
    // append the units with triangles to the active list: slots within the workgroup
    // from an LDS counter, one global atomic per workgroup
    const unsigned int na = (unsigned int)__popc(act);
    unsigned int off = 0;
    if (na) off = atomicAdd(&wg_n, na);
    __syncthreads();
    if (threadIdx.x == 0) wg_base = wg_n ? atomicAdd(m.nact, wg_n) : 0u;
    __syncthreads();
    if (na) {
        unsigned int p = wg_base + off;
        for (int f = 0; f < m.j.F; f++)
            if ((act >> f) & 1) m.active[p++] = (uint32_t)(t * m.j.F + f);
    }
}

// Owned slots of each active unit (cnt_v of the others stays 0).
__global__ __launch_bounds__(256) void k_mesh_own(MeshDev m) {
    const unsigned int n = *m.nact;
    for (unsigned int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const long long u = m.active[i];
        int c, f, x, y, z;
        unit_cube(m, u, c, f, x, y, z);
        const int cs = m.ucase[u];
        const int ns = 3 * m.ntri[cs];
        const int8_t *row = m.tri + 16 * cs;
        uint32_t own = 0;
        for (int s = 0; s < ns; s++) {
            const long long g = f * m.E3 + edge_slot(m, x, y, z, row[s]);
            own += m.first[g] == (unsigned long long)u * 16ull + (unsigned long long)s;
        }
        m.cnt_v[u] = own;
    }
}

// ---- exclusive prefix sums of the index and vertex counts at once ------------------
// Index counts are read as 3 * ntri[case] from the case bytes; the offsets go to a.
constexpr int SCAN_T = 256, SCAN_PER = 8, SCAN_TILE = SCAN_T * SCAN_PER;
static_assert(SCAN_T == 256, "one thread per case fills the index-count table");

template <int NT = SCAN_T>
__device__ __forceinline__ uint2 block_exclusive(uint2 x, uint2 *lds, uint2 &total) {
    const int t = threadIdx.x;
    lds[t] = x;
    __syncthreads();
    for (int o = 1; o < NT; o <<= 1) {
        const uint2 y = t >= o ? lds[t - o] : make_uint2(0u, 0u);
        __syncthreads();
        lds[t] = make_uint2(lds[t].x + y.x, lds[t].y + y.y);
        __syncthreads();
    }
    total = lds[NT - 1];
    const uint2 incl = lds[t];
    __syncthreads();
    return make_uint2(incl.x - x.x, incl.y - x.y);
}

// Tile sums (order-free, so the loads run across the workgroup: item k * SCAN_T + t).
__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(const uint8_t *cs, const uint8_t *ntri, const uint32_t *b,
                                                       long long n, uint32_t *ta, uint32_t *tb) {
    __shared__ uint2 lds[SCAN_T];
    __shared__ uint32_t nidx[256];
    nidx[threadIdx.x] = 3u * ntri[threadIdx.x];
    __syncthreads();
    const long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x;
    uint2 s = make_uint2(0u, 0u);
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        const long long i = base + (long long)k * SCAN_T;
        if (i < n) {
            s.x += nidx[cs[i]];
            s.y += b[i];
        }
    }
    uint2 total;
    (void)block_exclusive(s, lds, total);
    if (threadIdx.x == 0) {
        ta[blockIdx.x] = total.x;
        tb[blockIdx.x] = total.y;
    }
}

// One workgroup: the tile sums become exclusive tile offsets; tot[0..1] the totals.
constexpr int SCAN_TOP_T = 1024;
__global__ __launch_bounds__(SCAN_TOP_T) void k_scan_top(uint32_t *ta, uint32_t *tb, long long ntiles,
                                                         uint32_t *tot) {
    __shared__ uint2 lds[SCAN_TOP_T];
    uint2 run = make_uint2(0u, 0u);
    for (long long base = 0; base < ntiles; base += SCAN_TOP_T) {
        const long long i = base + threadIdx.x;
        const uint2 x = i < ntiles ? make_uint2(ta[i], tb[i]) : make_uint2(0u, 0u);
        uint2 total;
        const uint2 ex = block_exclusive<SCAN_TOP_T>(x, lds, total);
        if (i < ntiles) {
            ta[i] = run.x + ex.x;
            tb[i] = run.y + ex.y;
        }
        run = make_uint2(run.x + total.x, run.y + total.y);
    }
    if (threadIdx.x == 0) {
        tot[0] = run.x;
        tot[1] = run.y;
    }
}

// In-place exclusive offsets of one tile: staged through LDS so that global loads and
// stores run across the workgroup while each thread scans SCAN_PER consecutive items.
__global__ __launch_bounds__(SCAN_T) void k_scan_apply(const uint8_t *cs, const uint8_t *ntri, uint32_t *a,
                                                       uint32_t *b, long long n, const uint32_t *ta,
                                                       const uint32_t *tb) {
    __shared__ uint2 lds[SCAN_T];
    __shared__ uint32_t sa[SCAN_TILE], sb[SCAN_TILE];
    __shared__ uint32_t nidx[256];
    nidx[threadIdx.x] = 3u * ntri[threadIdx.x];
    __syncthreads();
    const long long tile = (long long)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        const int j = k * SCAN_T + threadIdx.x;
        sa[j] = tile + j < n ? nidx[cs[tile + j]] : 0u;
        sb[j] = tile + j < n ? b[tile + j] : 0u;
    }
    __syncthreads();
    uint32_t va[SCAN_PER], vb[SCAN_PER];
    uint2 s = make_uint2(0u, 0u);
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        va[k] = sa[threadIdx.x * SCAN_PER + k];
        vb[k] = sb[threadIdx.x * SCAN_PER + k];
        s.x += va[k];
        s.y += vb[k];
    }
    uint2 total;
    uint2 run = block_exclusive(s, lds, total);  // ends with a barrier: sa, sb reads done
    run.x += ta[blockIdx.x];
    run.y += tb[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        sa[threadIdx.x * SCAN_PER + k] = run.x;
        sb[threadIdx.x * SCAN_PER + k] = run.y;
        run.x += va[k];
        run.y += vb[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        const int j = k * SCAN_T + threadIdx.x;
        if (tile + j < n) {
            a[tile + j] = sa[j];
            b[tile + j] = sb[j];
        }
    }
}

__global__ __launch_bounds__(256) void k_mesh_emit(MeshDev m, MeshOut o, unsigned int nact) {
    const unsigned int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nact) return;
    const long long u = m.active[i];
    int c, f, x, y, z;
    unit_cube(m, u, c, f, x, y, z);
    const int cs = m.ucase[u];
    const int ns = 3 * m.ntri[cs];
    float v[8];
    load_corners(m, c, f, x, y, z, v);
    const int8_t *row = m.tri + 16 * cs;
    uint32_t *tu = o.tunit + m.cnt_i[u] / 3;
    for (int tri = 0; tri < ns / 3; tri++) tu[tri] = (uint32_t)u;
    uint32_t vid = m.cnt_v[u];
    for (int s = 0; s < ns; s++) {
        const int e = row[s];
        const long long g = f * m.E3 + edge_slot(m, x, y, z, e);
        if (m.first[g] != (unsigned long long)u * 16ull + (unsigned long long)s) continue;
        float px, py, pz;
        edge_point(m, v, x, y, z, e, px, py, pz);
        float *out = o.verts + 7ll * vid;
        out[0] = px;
        out[1] = py;
        out[2] = pz;
        out[6] = __uint_as_float(m.colors[c]);
        m.vmap[g] = vid;
        o.vedge[vid] = (uint32_t)g;
        vid++;
    }
}

// The index buffer, and the face normal of every triangle (mesh.cpp:229-300) from its
// cube's own interpolated corners: one thread per triangle, at its serial position.
__global__ __launch_bounds__(256) void k_mesh_index(MeshDev m, MeshOut o, long long ntri) {
    const long long tau = (long long)blockIdx.x * 256 + threadIdx.x;
    if (tau >= ntri) return;
    const long long u = o.tunit[tau];
    const uint32_t base = m.cnt_i[u];
    const int tri = (int)(tau - base / 3);
    int c, f, x, y, z;
    unit_cube(m, u, c, f, x, y, z);
    const int8_t *row = m.tri + 16 * m.ucase[u];
    const int e0 = row[3 * tri], e1 = row[3 * tri + 1], e2 = row[3 * tri + 2];
    uint32_t *ix = o.idx + 3 * tau;
    ix[0] = m.vmap[f * m.E3 + edge_slot(m, x, y, z, e0)];
    ix[1] = m.vmap[f * m.E3 + edge_slot(m, x, y, z, e1)];
    ix[2] = m.vmap[f * m.E3 + edge_slot(m, x, y, z, e2)];
    float v[8];
    load_corners(m, c, f, x, y, z, v);
    float ax, ay, az, bx, by, bz, qx, qy, qz;
    edge_point(m, v, x, y, z, e0, ax, ay, az);
    edge_point(m, v, x, y, z, e1, bx, by, bz);
    edge_point(m, v, x, y, z, e2, qx, qy, qz);
    const float ux = bx - ax, uy = by - ay, uz = bz - az;
    const float vx = qx - ax, vy = qy - ay, vz = qz - az;
    float *fn = o.fnorm + 3 * tau;
    fn[0] = uy * vz - uz * vy;
    fn[1] = uz * vx - ux * vz;
    fn[2] = ux * vy - uy * vx;
}

// Normal of vertex i: the face normals of every triangle naming it, summed over
// channels, then the (up to four) cubes around its edge in sweep order, then each cube's
// triangles in table order -- the reference's order, so the fp32 sums agree.
__global__ __launch_bounds__(256) void k_mesh_normal(MeshDev m, MeshOut o, long long nv) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= nv) return;
    const long long g = o.vedge[i];
    const int f = (int)(g / m.E3);
    const long long r = g % m.E3;
    const int axis = (int)(r / m.E);
    const long long b = r % m.E;
    const int W1 = m.j.W + 1, H1 = m.j.H + 1;
    const int bx = (int)(b % W1), by = (int)((b / W1) % H1), bz = (int)(b / ((long long)W1 * H1));
    // the two axes across the edge, lower first: cube = base - (s on p) - (t on q)
    const int p = axis == 0 ? 1 : 0, q = axis == 2 ? 1 : 2;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    for (int c = 0; c < m.j.C; c++) {
        for (int k = 0; k < 4; k++) {
            const int s = (k & 1) ? 0 : 1, t = (k & 2) ? 0 : 1;  // (p-1,q-1), (p,q-1), (p-1,q), (p,q)
            int cx = bx, cy = by, cz = bz;
            if (p == 0) cx -= s; else cy -= s;
            if (q == 1) cy -= t; else cz -= t;
            if (cx < 0 || cy < 0 || cz < 0 || cx >= m.j.W || cy >= m.j.H || cz >= m.j.D) continue;
            const long long cube = cx + (long long)m.j.W * (cy + (long long)m.j.H * cz);
            const long long u = ((long long)c * m.Nc + cube) * m.j.F + f;
            const int nt = m.ntri[m.ucase[u]];
            const uint32_t base = m.cnt_i[u];
            for (int tri = 0; tri < nt; tri++) {
                const uint32_t *ix = o.idx + base + 3 * tri;
                const int hits = (ix[0] == (uint32_t)i) + (ix[1] == (uint32_t)i) + (ix[2] == (uint32_t)i);
                if (!hits) continue;
                const float *fn = o.fnorm + base + 3 * tri;
                const float fx = fn[0], fy = fn[1], fz = fn[2];
                for (int h = 0; h < hits; h++) {
                    nx = nx + fx;
                    ny = ny + fy;
                    nz = nz + fz;
                }
            }
        }
    }
    float *out = o.verts + 7ll * i;
    out[3] = nx;
    out[4] = ny;
    out[5] = nz;
}

inline size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

hipError_t grow(void **p, size_t *cap, size_t need) {
    if (*cap >= need) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(p, need);
    if (e == hipSuccess) *cap = need;
    return e;
}

unsigned grid_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ---- the built-in triangulation table ---------------------------------------------

void lh_mc_derive(int *tri) {
    // corner cycles of the six faces, each consecutive pair one cube edge
    int faces[6][4];
    int nf = 0;
    for (int axis = 0; axis < 3; axis++)
        for (int side = 0; side < 2; side++) {
            int cs[4], n = 0;
            for (int k = 0; k < 8; k++) {
                const int co[3] = {MC_DX[k], MC_DY[k], MC_DZ[k]};
                if (co[axis] == side) cs[n++] = k;
            }
            int *cyc = faces[nf++];
            bool used[4] = {true, false, false, false};
            cyc[0] = cs[0];
            for (int i = 1; i < 4; i++)
                for (int j = 1; j < 4; j++) {
                    const int a = cyc[i - 1], b = cs[j];
                    const int d = (MC_DX[a] != MC_DX[b]) + (MC_DY[a] != MC_DY[b]) + (MC_DZ[a] != MC_DZ[b]);
                    if (!used[j] && d == 1) {
                        cyc[i] = b;
                        used[j] = true;
                        break;
                    }
                }
        }
    auto edge_of = [](int a, int b) {
        for (int e = 0; e < 12; e++)
            if ((MC_EA[e] == a && MC_EB[e] == b) || (MC_EA[e] == b && MC_EB[e] == a)) return e;
        return -1;
    };
    auto mid = [](int e, double m[3]) {
        const int a = MC_EA[e], b = MC_EB[e];
        m[0] = 0.5 * (MC_DX[a] + MC_DX[b]);
        m[1] = 0.5 * (MC_DY[a] + MC_DY[b]);
        m[2] = 0.5 * (MC_DZ[a] + MC_DZ[b]);
    };
    for (int cs = 0; cs < 256; cs++) {
        int *row = tri + 16 * cs;
        for (int i = 0; i < 16; i++) row[i] = -1;
        auto in = [cs](int k) { return (cs >> k) & 1; };
        int adj[12][2], deg[12] = {0};
        for (int fi = 0; fi < 6; fi++) {
            const int *q = faces[fi];
            int fe[4], cr[4], nc = 0;
            for (int i = 0; i < 4; i++) {
                fe[i] = edge_of(q[i], q[(i + 1) & 3]);
                if (in(q[i]) != in(q[(i + 1) & 3])) cr[nc++] = i;
            }
            int pairs[2][2], np = 0;
            if (nc == 2) {
                pairs[np][0] = fe[cr[0]];
                pairs[np++][1] = fe[cr[1]];
            } else if (nc == 4) {  // separate the corners above iso
                const int o = in(q[0]) ? 3 : 0;
                pairs[0][0] = fe[o];
                pairs[0][1] = fe[(o + 1) & 3];
                pairs[1][0] = fe[(o + 2) & 3];
                pairs[1][1] = fe[(o + 3) & 3];
                np = 2;
            }
            for (int i = 0; i < np; i++) {
                const int a = pairs[i][0], b = pairs[i][1];
                adj[a][deg[a]++] = b;
                adj[b][deg[b]++] = a;
            }
        }
        bool seen[12] = {false};
        int out = 0;
        for (int s = 0; s < 12; s++) {
            if (!deg[s] || seen[s]) continue;
            int loop[12], n = 0, prev = -1, cur = s;
            while (true) {
                loop[n++] = cur;
                seen[cur] = true;
                const int nx = adj[cur][0] != prev && !seen[adj[cur][0]] ? adj[cur][0]
                             : (!seen[adj[cur][1]] ? adj[cur][1] : -1);
                if (nx < 0) break;
                prev = cur;
                cur = nx;
            }
            // orientation: fan area vector against the sum of in -> out edge directions
            double m0[3], mi[3], mj[3], nrm[3] = {0, 0, 0}, dir[3] = {0, 0, 0};
            mid(loop[0], m0);
            for (int i = 1; i + 1 < n; i++) {
                mid(loop[i], mi);
                mid(loop[i + 1], mj);
                const double u[3] = {mi[0] - m0[0], mi[1] - m0[1], mi[2] - m0[2]};
                const double w[3] = {mj[0] - m0[0], mj[1] - m0[1], mj[2] - m0[2]};
                nrm[0] += u[1] * w[2] - u[2] * w[1];
                nrm[1] += u[2] * w[0] - u[0] * w[2];
                nrm[2] += u[0] * w[1] - u[1] * w[0];
            }
            for (int i = 0; i < n; i++) {
                int a = MC_EA[loop[i]], b = MC_EB[loop[i]];
                if (!in(a)) {
                    const int t = a;
                    a = b;
                    b = t;
                }
                dir[0] += MC_DX[b] - MC_DX[a];
                dir[1] += MC_DY[b] - MC_DY[a];
                dir[2] += MC_DZ[b] - MC_DZ[a];
            }
            if (nrm[0] * dir[0] + nrm[1] * dir[1] + nrm[2] * dir[2] < 0)
                for (int i = 1, j = n - 1; i < j; i++, j--) {
                    const int t = loop[i];
                    loop[i] = loop[j];
                    loop[j] = t;
                }
            for (int i = 1; i + 1 < n && out + 3 <= 15; i++) {
                row[out++] = loop[0];
                row[out++] = loop[i];
                row[out++] = loop[i + 1];
            }
        }
    }
}

const char *lh_mc_invalid(const int *tri) {
    for (int cs = 0; cs < 256; cs++) {
        const int *row = tri + 16 * cs;
        int n = 0;
        while (n < 16 && row[n] >= 0) n++;
        if (n == 16 || n % 3 != 0 || n > 15) return "a row holds more than 5 triangles or no -1 terminator";
        for (int i = 0; i < n; i++) {
            const int e = row[i];
            if (e > 11) return "an edge index is outside 0..11";
            if (((cs >> MC_EA[e]) & 1) == ((cs >> MC_EB[e]) & 1)) return "a triangle names an edge its case does not cross";
        }
    }
    return nullptr;
}

// ---- host driver ------------------------------------------------------------------

hipError_t lh_mesh_run(const MeshJob &job, const int *tri, const uint32_t *colors, MeshBufs &b, hipStream_t st,
                       void *verts, int64_t max_v, uint32_t *idx, int64_t max_i, int64_t *nv, int64_t *ni) {
    MeshDev m;
    m.j = job;
    m.Nc = (long long)job.W * job.H * job.D;
    m.U = m.Nc * job.C * job.F;
    m.E = (long long)(job.W + 1) * (job.H + 1) * (job.D + 1);
    m.E3 = 3 * m.E;
    const long long ntiles = (m.U + SCAN_TILE - 1) / SCAN_TILE;
    // work layout: table | ntri | colors | first | vmap | ucase | cnt_i | cnt_v | tiles | totals
    size_t off = 0;
    const size_t o_tri = off;
    off = up256(off + 256 * 16);
    const size_t o_ntri = off;
    off = up256(off + 256);
    const size_t o_col = off;
    off = up256(off + LHD_MAX_C * sizeof(uint32_t));
    const size_t o_first = off;
    off = up256(off + (size_t)(job.F * m.E3) * sizeof(unsigned long long));
    const size_t o_vmap = off;
    off = up256(off + (size_t)(job.F * m.E3) * sizeof(uint32_t));
    const size_t o_case = off;
    off = up256(off + (size_t)m.U);
    const size_t o_ci = off;
    off = up256(off + (size_t)m.U * sizeof(uint32_t));
    const size_t o_cv = off;
    off = up256(off + (size_t)m.U * sizeof(uint32_t));
    const size_t o_act = off;
    off = up256(off + (size_t)m.U * sizeof(uint32_t));
    const size_t o_ta = off;
    off = up256(off + (size_t)ntiles * sizeof(uint32_t));
    const size_t o_tb = off;
    off = up256(off + (size_t)ntiles * sizeof(uint32_t));
    const size_t o_tot = off;
    off = up256(off + 3 * sizeof(uint32_t));  // index and vertex totals, active units
    hipError_t e = grow(&b.work, &b.work_cap, off);
    if (e != hipSuccess) return e;
    char *w = (char *)b.work;
    int8_t htri[256 * 16];
    uint8_t hn[256];
    // A row ends at its first negative entry (lh_mc_invalid checks exactly that prefix);
    // whatever a caller's row holds after it is replaced by -1 and never read.
    for (int cs = 0; cs < 256; cs++) {
        int n = 0;
        while (n < 16 && tri[16 * cs + n] >= 0) n++;
        for (int i = 0; i < 16; i++) htri[16 * cs + i] = (int8_t)(i < n ? tri[16 * cs + i] : -1);
        hn[cs] = (uint8_t)(n / 3);
    }
    uint32_t hcol[LHD_MAX_C];
    for (int c = 0; c < LHD_MAX_C; c++) hcol[c] = colors && c < job.C ? colors[c] : 0xFFFFFFFFu;
    m.tri = (const int8_t *)(w + o_tri);
    m.ntri = (const uint8_t *)(w + o_ntri);
    m.colors = (const uint32_t *)(w + o_col);
    m.first = (unsigned long long *)(w + o_first);
    m.vmap = (uint32_t *)(w + o_vmap);
    m.ucase = (uint8_t *)(w + o_case);
    m.cnt_i = (uint32_t *)(w + o_ci);
    m.cnt_v = (uint32_t *)(w + o_cv);
    m.active = (uint32_t *)(w + o_act);
    uint32_t *ta = (uint32_t *)(w + o_ta), *tb = (uint32_t *)(w + o_tb), *tot = (uint32_t *)(w + o_tot);
    m.nact = tot + 2;
    if ((e = hipMemcpyAsync(w + o_tri, htri, sizeof(htri), hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(w + o_ntri, hn, sizeof(hn), hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(w + o_col, hcol, sizeof(hcol), hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    const long long nfirst = job.F * m.E3;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)std::min<long long>(grid_for(nfirst), 4096)), dim3(256), 0, st,
                       m.first, nfirst, ~0ull);
    if ((e = hipMemsetAsync(m.cnt_v, 0, (size_t)m.U * sizeof(uint32_t), st)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(m.nact, 0, sizeof(uint32_t), st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_mesh_count, dim3(grid_for(m.U / job.F)), dim3(256), 0, st, m);
    hipLaunchKernelGGL(k_mesh_own, dim3((unsigned)std::min<long long>(grid_for(m.U), 2048)), dim3(256), 0, st, m);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(SCAN_T), 0, st, m.ucase, m.ntri, m.cnt_v, m.U, ta,
                       tb);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_TOP_T), 0, st, ta, tb, ntiles, tot);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)ntiles), dim3(SCAN_T), 0, st, m.ucase, m.ntri, m.cnt_i, m.cnt_v,
                       m.U, ta, tb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t htot[3];
    if ((e = hipMemcpyAsync(htot, tot, sizeof(htot), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    *ni = htot[0];
    *nv = htot[1];
    if (!verts || !idx || *nv > max_v || *ni > max_i || *nv == 0) return hipSuccess;
    const long long ntri = *ni / 3;
    const size_t vb = up256((size_t)*nv * 28), ib = up256((size_t)*ni * 4), eb = up256((size_t)*nv * 4);
    const size_t ub = up256((size_t)ntri * 4);
    if ((e = grow(&b.out, &b.out_cap, vb + 2 * ib + eb + ub)) != hipSuccess) return e;
    MeshOut o;
    o.verts = (float *)b.out;
    o.idx = (uint32_t *)((char *)b.out + vb);
    o.fnorm = (float *)((char *)b.out + vb + ib);
    o.vedge = (uint32_t *)((char *)b.out + vb + 2 * ib);
    o.tunit = (uint32_t *)((char *)b.out + vb + 2 * ib + eb);
    hipLaunchKernelGGL(k_mesh_emit, dim3(grid_for(htot[2])), dim3(256), 0, st, m, o, htot[2]);
    hipLaunchKernelGGL(k_mesh_index, dim3(grid_for(ntri)), dim3(256), 0, st, m, o, ntri);
    hipLaunchKernelGGL(k_mesh_normal, dim3(grid_for(*nv)), dim3(256), 0, st, m, o, (long long)*nv);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(verts, o.verts, (size_t)*nv * 28, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(idx, o.idx, (size_t)*ni * 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    return hipStreamSynchronize(st);
}
