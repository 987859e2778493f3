// Host build of the reference's Mandelbrot module (src/main/cuda/fractals/mandelbrot.cu)
// for pinning the oracle.  TEST INFRASTRUCTURE ONLY.
//
// helpers.cuh:83 names its constructor parameters lt/rb but initialises from
// lb/rt; nvcc accepts that, g++ does not.  The two macros below rename the
// initialisers for the duration of the include.
#include "cuda_host_shim.hpp"
#define lb lt
#define rt rb
/* both modules define these; distinct names keep the two template
   instantiations from merging at link time */
#define debugFractal ref_debug_mandelbrot
#define computeFractal ref_compute_mandelbrot
#include "fractals/mandelbrot.cu"
#undef lb
#undef rt
#undef debugFractal

extern "C" float ref_mandelbrot_d(unsigned max_iter, double cx, double cy) {
    return ref_compute_mandelbrot<double>(max_iter, Point<double>(cx, cy));
}
extern "C" float ref_mandelbrot_f(unsigned max_iter, float cx, float cy) {
    return ref_compute_mandelbrot<float>(max_iter, Point<float>(cx, cy));
}
