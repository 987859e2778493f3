// mandelbrot.hip — the Mandelbrot fractal module (z0 = 0, c = the plane point).
//
// Counterpart of src/main/cuda/fractals/mandelbrot.cu: computeFractal (:11-25)
// returns the escape count, 0 for an orbit that used every iteration; colorize
// (:27-35) and debugFractal (:37-39) keep their behaviour.  Compiled into
// mandelbrot.hsaco together with render_generic.hip.
#include "../escape_core.hpp"

struct Fractal {
    static constexpr int kKind = CU_FRACTAL_MANDELBROT;

    template <class Real>
    __device__ static void start(Real X, Real Y, Real &zx, Real &zy, Real &cx, Real &cy) {
        zx = 0;
        zy = 0;
        cx = X;
        cy = Y;
    }

    // computeFractal's return value for an orbit of n of max_iter iterations.
    __device__ static float value(uint n, uint max_iter) { return n == max_iter ? 0.f : (float)n; }

    // c is the plane point: the fused-step rule has nothing to add.
    template <class Real>
    __device__ static bool c_fused_valid(double) { return true; }

    __device__ static bool blocks_valid(double bailout_sq) {
        return cu::blocks_valid(bailout_sq, false, 0, 0);
    }
};

__device__ void debugFractal() { printf("hello from mandelbrot\n"); }
