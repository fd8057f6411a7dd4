// lh_kernels_big.hip -- the large shapes built into liblenia_hip.so (lh_kernels_big.hpp
// instantiated per shape).
#include "lh_kernels_big.hpp"

static const ShapeOps g_ops_big[] = {LaunchBig<64, 64, 64>::ops(), LaunchBig<96, 96, 96>::ops(),
                                     LaunchBig<128, 128, 128>::ops()};

const ShapeOps *lh_shape_ops_big(int W, int H, int D) {
    for (const ShapeOps &o : g_ops_big)
        if (o.W == W && o.H == H && o.D == D) return &o;
    return nullptr;
}
