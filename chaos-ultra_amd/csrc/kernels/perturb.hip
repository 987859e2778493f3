// perturb.hip — entry points of the deep-zoom render form (part of
// escape_family.hsaco); the device code is in perturb_core.hpp.
//
//   referenceOrbit{Float,Double}   one lane: double-double orbit of C -> entries, M
//   perturb{Float,Double}          lane per pixel (8x8 pixels per wave, 16x16 / 32x16 per block,
//                                  orbit staged in LDS)
//   deepDirectDouble               direct FP64 at the same pixel -> c mapping
#include "perturb_core.hpp"

extern "C" __global__ __launch_bounds__(64) void referenceOrbitFloat(
    double cx_hi, double cx_lo, double cy_hi, double cy_lo, uint max_iter, double bailout_sq,
    OrbitEntry<float> *orbit, uint *length) {
    reference_orbit<float>(cx_hi, cx_lo, cy_hi, cy_lo, max_iter, bailout_sq, orbit, length);
}
extern "C" __global__ __launch_bounds__(64) void referenceOrbitDouble(
    double cx_hi, double cx_lo, double cy_hi, double cy_lo, uint max_iter, double bailout_sq,
    OrbitEntry<double> *orbit, uint *length) {
    reference_orbit<double>(cx_hi, cx_lo, cy_hi, cy_lo, max_iter, bailout_sq, orbit, length);
}
// `orbit` repeats a.orbit as its own read-only argument, so uniform loads from
// it can go through the scalar cache.  The second bound (8 waves per SIMD)
// keeps the kernels within 64 VGPRs, so that a CU's 32 wave slots can all be
// filled; without it the compiler keeps a whole block's orbit reads in flight.
// Blocks of 32 steps: against 16, 7-9 % faster at 1e-10 and 1e-20 and the same at
// 1e-6, where a rebase reruns more of a longer block.  Block shapes: perturb_frame.
extern "C" __global__ __launch_bounds__(256, 8) void perturbFloat(
    DeepArgs a, const OrbitEntry<float> *__restrict__ orbit) {
    perturb_frame<float, 32, 2, 2, 20u * 1024u>(a, orbit);
}
extern "C" __global__ __launch_bounds__(512, 8) void perturbDouble(
    DeepArgs a, const OrbitEntry<double> *__restrict__ orbit) {
    perturb_frame<double, 32, 4, 2, 40u * 1024u>(a, orbit);
}
extern "C" __global__ __launch_bounds__(256) void deepDirectDoubleExact(DeepArgs a) {
    deep_direct<16, false>(a);
}
extern "C" __global__ __launch_bounds__(256) void deepDirectDoubleFused(DeepArgs a) {
    deep_direct<16, true>(a);
}
