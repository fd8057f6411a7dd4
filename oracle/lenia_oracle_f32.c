/* Single-precision instantiation of the CPU oracle (test infrastructure only). */
#define REAL float
#define PFX(n) lo_f32_##n
#include "lenia_oracle_impl.h"

int lo_nthreads = 1;

/* Threads of the checker's parallel loops (1 = serial); results are identical for any
 * count.  Returns the count in effect. */
int lo_set_threads(int n) {
    lo_nthreads = n < 1 ? 1 : n;
    return lo_nthreads;
}

/* Edge mask of a marching-cubes case: edge e is cut iff its two corners lie on
 * different sides of the iso level.  Corner pairs follow the cube numbering of
 * marching_cubes.h:17-22 (a 12-entry edge list that must match the reference). */
int lo_edge_mask(int v) {
    static const int E[12][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {4, 5}, {5, 6},
                                 {6, 7}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    int m = 0;
    for (int e = 0; e < 12; e++)
        if (((v >> E[e][0]) & 1) != ((v >> E[e][1]) & 1)) m |= 1 << e;
    return m;
}
