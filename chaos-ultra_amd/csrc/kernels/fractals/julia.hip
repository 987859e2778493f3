// julia.hip — the Julia fractal module (z0 = the plane point, c = julia_c).
//
// Counterpart of src/main/cuda/fractals/julia.cu: c is the module global
// julia_c cast to the kernel's precision (:7), the escape count is returned as
// is, so an orbit that used every iteration returns max_iter (:16).  The host
// writes julia_c through the module-global path ("re;im" custom parameters).
#include "../escape_core.hpp"

__constant__ double julia_c[2];

struct Fractal {
    static constexpr int kKind = CU_FRACTAL_JULIA;

    template <class Real>
    __device__ static void start(Real X, Real Y, Real &zx, Real &zy, Real &cx, Real &cy) {
        zx = X;
        zy = Y;
        cx = (Real)julia_c[0];
        cy = (Real)julia_c[1];
    }

    __device__ static float value(uint n, uint) { return (float)n; }

    // julia_c (in the kernel's precision) zero or at least `threshold` in magnitude.
    template <class Real>
    __device__ static bool c_fused_valid(double threshold) {
        const double re = (double)(Real)julia_c[0], im = (double)(Real)julia_c[1];
        return (re == 0 || fabs(re) >= threshold) && (im == 0 || fabs(im) >= threshold);
    }

    __device__ static bool blocks_valid(double bailout_sq) {
        return cu::blocks_valid(bailout_sq, true, julia_c[0], julia_c[1]);
    }
};

__device__ void debugFractal() { printf("hello from julia\n"); }
