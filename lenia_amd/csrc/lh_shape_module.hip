// lh_shape_module.hip -- the phase kernels of one world shape as a loadable module
// (liblenia_shape_<W>x<H>x<D>.so), for shapes not built into liblenia_hip.so.  The
// reference is compiled for one GRID_WIDTH x GRID_HEIGHT x GRID_WIDTH with sides
// divisible by 4 (render.h:27-29, README.md:30); here any W x H x D with sides divisible
// by 4 from 4 to 128 is built on demand by lenia_amd/build.py (-DLH_MODULE_W/H/D) and
// loaded by the engine when a world of that shape is added (lh_shape_ops).  Sides up to
// 48 take the register-resident kernels of lh_kernels.hpp, larger ones the LDS-staged
// spectral kernels of lh_kernels_big.hpp; both give results bit-identical to the CPU
// restatement for any shape.
#include "lh_kernels.hpp"
#include "lh_kernels_big.hpp"

#if !defined(LH_MODULE_W) || !defined(LH_MODULE_H) || !defined(LH_MODULE_D)
#error "build with -DLH_MODULE_W=<W> -DLH_MODULE_H=<H> -DLH_MODULE_D=<D>"
#endif

constexpr int MW = LH_MODULE_W, MH = LH_MODULE_H, MD = LH_MODULE_D;
static_assert(MW % 4 == 0 && MH % 4 == 0 && MD % 4 == 0, "sides must be divisible by 4");
static_assert(MW >= 4 && MH >= 4 && MD >= 4 && MW <= 128 && MH <= 128 && MD <= 128, "sides must be 4..128");

template <int W, int H, int D>
constexpr ShapeOps module_ops() {
    if constexpr (W > 48 || H > 48 || D > 48) return LaunchBig<W, H, D>::ops();
    else return Launch<W, H, D>::ops();
}

static const ShapeOps g_module_ops = module_ops<MW, MH, MD>();

extern "C" __attribute__((visibility("default"))) const ShapeOps *lh_module_shape_ops(void) { return &g_module_ops; }
