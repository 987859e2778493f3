// Host build of the reference's Julia module (src/main/cuda/fractals/julia.cu)
// for pinning the oracle.  TEST INFRASTRUCTURE ONLY.  See ref_mandelbrot.cpp.
#include "cuda_host_shim.hpp"
#define lb lt
#define rt rb
/* both modules define these; distinct names keep the two template
   instantiations from merging at link time */
#define debugFractal ref_debug_julia
#define computeFractal ref_compute_julia
#include "fractals/julia.cu"
#undef lb
#undef rt
#undef debugFractal

extern "C" float ref_julia_d(unsigned max_iter, double zx, double zy, double jcx, double jcy) {
    julia_c[0] = jcx;
    julia_c[1] = jcy;
    return ref_compute_julia<double>(max_iter, Point<double>(zx, zy));
}
extern "C" float ref_julia_f(unsigned max_iter, float zx, float zy, double jcx, double jcy) {
    julia_c[0] = jcx;
    julia_c[1] = jcy;
    return ref_compute_julia<float>(max_iter, Point<float>(zx, zy));
}
