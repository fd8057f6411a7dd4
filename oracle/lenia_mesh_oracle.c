/* Marching-cubes meshes of the renderer's views -- CPU restatement, TEST INFRASTRUCTURE.
 *
 * Follows generate_mesh_v1 / v2 / v3 (mesh.cpp:34-430, 434-1113, 1115-2036) with one
 * mesh section spanning the whole grid (as if MESH_THREADS, render.h:15, were 1, so the border
 * stitching of main.cpp:105-170 has nothing to join):
 *   for channel c, for cube z, y, x, for view field f (1: A; 2: A, U_SUM; 3: A, U_SUM, H):
 *     case index from the 8 corner values > iso (mesh.cpp:173-180);
 *     each crossing edge interpolated in the cube's own corner order (mesh.cpp:185-201):
 *       T = (iso - v_a) / (v_b - v_a);  p = p_a + (p_b - p_a) * T,  p_k = (g + delta_k) * cell;
 *     triangles from the row of the triangulation table (mesh.cpp:203-227), face normal
 *       n = (B - A) x (C - A) unnormalised (mesh.cpp:229-300);
 *     each triangle corner looks up the vertex of its edge in field f's edge map; a miss
 *       appends Vertex(p, color of c) with zero normal (render.h:469-474); the face normal
 *       is added to the vertex and its index appended (mesh.cpp:304-415).
 * The edge maps are per field but shared by the channels (ETIDX, render.h:517, carries
 * no channel), so a later channel's triangle reuses the vertex an earlier channel made on
 * the same edge -- kept as the reference does it.  Corner values come from the torus face
 * pads of transpose_mesh_borders; edge and corner pads read 0 when face_only
 * (transposition.cpp:881-995), as in NAME(mc_case).
 *
 * Pinned by tests/test_mesh_oracle.py: closed, consistently oriented surfaces with the
 * Euler characteristic of a sphere (2) and a torus (0), outward normals, vertices on the
 * iso-surface of a linear field, and hand-counted small cases. */
#include <stdint.h>
#include <stdlib.h>

#include "lenia_oracle.h"

static const int MO_DX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int MO_DY[8] = {0, 0, 0, 0, 1, 1, 1, 1};
static const int MO_DZ[8] = {0, 0, 1, 1, 0, 0, 1, 1};
/* corner pairs of the 12 cube edges (marching_cubes.h:19-31) */
static const int MO_EA[12] = {0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 2, 3};
static const int MO_EB[12] = {1, 2, 3, 3, 5, 6, 7, 7, 4, 5, 6, 7};

static float mo_value(const float *f, int W, int H, int D, int cx, int cy, int cz, int face_only) {
    int m = (cx >= W) + (cy >= H) + (cz >= D);
    if (m > 1 && face_only) return 0.0f;
    return f[(size_t)(cx % W) + (size_t)(cy % H) * W + (size_t)(cz % D) * W * H];
}

/* Edge-map slot of local edge e of cube (x, y, z): axis * E + BIDX of its low corner. */
static int64_t mo_edge(int W, int H, int D, int x, int y, int z, int e) {
    const int a = MO_EA[e], b = MO_EB[e];
    int ox = MO_DX[a] < MO_DX[b] ? MO_DX[a] : MO_DX[b];
    int oy = MO_DY[a] < MO_DY[b] ? MO_DY[a] : MO_DY[b];
    int oz = MO_DZ[a] < MO_DZ[b] ? MO_DZ[a] : MO_DZ[b];
    int axis = MO_DX[a] != MO_DX[b] ? 0 : (MO_DY[a] != MO_DY[b] ? 1 : 2);
    const int64_t E = (int64_t)(W + 1) * (H + 1) * (D + 1);
    return axis * E + (x + ox) + (int64_t)(y + oy) * (W + 1) + (int64_t)(z + oz) * (W + 1) * (H + 1);
}

int lo_mesh(int W, int H, int D, int C, int F, const float *f0, const float *f1, const float *f2, float iso,
            float cell, int face_only, const int *tri, const uint32_t *colors, lo_vertex *verts, int64_t max_verts,
            uint32_t *indices, int64_t max_indices, int64_t *n_verts, int64_t *n_indices) {
    const float *fields[3] = {f0, f1, f2};
    const int64_t N = (int64_t)W * H * D;
    const int64_t E3 = 3 * (int64_t)(W + 1) * (H + 1) * (D + 1);
    int64_t *map = (int64_t *)malloc((size_t)(F * E3) * sizeof(int64_t));
    if (!map) return -2;
    for (int64_t i = 0; i < F * E3; i++) map[i] = -1;
    int64_t nv = 0, ni = 0;
    for (int c = 0; c < C; c++)
        for (int z = 0; z < D; z++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++)
                    for (int f = 0; f < F; f++) {
                        const float *fc = fields[f] + (size_t)c * N;
                        float val[8], px[8], py[8], pz[8];
                        int cs = 0;
                        for (int k = 0; k < 8; k++) {
                            val[k] = mo_value(fc, W, H, D, x + MO_DX[k], y + MO_DY[k], z + MO_DZ[k], face_only);
                            px[k] = ((float)x + (float)MO_DX[k]) * cell;
                            py[k] = ((float)y + (float)MO_DY[k]) * cell;
                            pz[k] = ((float)z + (float)MO_DZ[k]) * cell;
                            if (val[k] > iso) cs |= 1 << k;
                        }
                        const int *row = tri + 16 * cs;
                        float ex[12], ey[12], ez[12];
                        for (int e = 0; e < 12; e++) {
                            const int a = MO_EA[e], b = MO_EB[e];
                            if (((cs >> a) & 1) == ((cs >> b) & 1)) continue;
                            float T = (iso - val[a]) / (val[b] - val[a]);
                            ex[e] = px[a] + (px[b] - px[a]) * T;
                            ey[e] = py[a] + (py[b] - py[a]) * T;
                            ez[e] = pz[a] + (pz[b] - pz[a]) * T;
                        }
                        for (int t = 0; t < 5 && row[3 * t] >= 0; t++) {
                            const int ea = row[3 * t], eb = row[3 * t + 1], ec = row[3 * t + 2];
                            float ux = ex[eb] - ex[ea], uy = ey[eb] - ey[ea], uz = ez[eb] - ez[ea];
                            float vx = ex[ec] - ex[ea], vy = ey[ec] - ey[ea], vz = ez[ec] - ez[ea];
                            float nx = uy * vz - uz * vy;
                            float ny = uz * vx - ux * vz;
                            float nz = ux * vy - uy * vx;
                            for (int k = 0; k < 3; k++) {
                                const int e = row[3 * t + k];
                                int64_t *slot = map + f * E3 + mo_edge(W, H, D, x, y, z, e);
                                if (*slot < 0) {
                                    if (nv < max_verts) {
                                        lo_vertex v = {ex[e], ey[e], ez[e], 0.0f, 0.0f, 0.0f,
                                                       colors ? colors[c] : 0xFFFFFFFFu};
                                        verts[nv] = v;
                                    }
                                    *slot = nv++;
                                }
                                if (*slot < max_verts) {
                                    verts[*slot].nx += nx;
                                    verts[*slot].ny += ny;
                                    verts[*slot].nz += nz;
                                }
                                if (ni < max_indices) indices[ni] = (uint32_t)*slot;
                                ni++;
                            }
                        }
                    }
    free(map);
    *n_verts = nv;
    *n_indices = ni;
    return (nv <= max_verts && ni <= max_indices) ? 0 : -1;
}
