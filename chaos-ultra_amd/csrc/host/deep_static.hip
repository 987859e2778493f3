// deep_static.hip — the deep-zoom render for callers that cannot allocate
// (cu_deep_render_static).
//
// cu_deep_render needs a cu_deep frame, whose creation allocates the reference
// orbit, its length word, the pixel planes and the iteration counters.  This
// entry point renders the same rebasing perturbation frame into planes the
// caller owns and keeps its scratch — the reference orbit and its length M —
// in `scratch`, a file-scope device array of fixed size: each GPU the library
// runs on holds one copy until the process ends.  Nothing here calls hipMalloc,
// hipFree or any allocating routine; the kernels are built into the library
// itself (no module load), so the call only enqueues two kernels on `stream`.
//
// The array's name is a host-side stand-in, not a device address: its address
// on each device comes from hipGetSymbolAddress, taken once per device with that
// device current (at cu_init, or at the first call on a device) and kept in
// g_scratch; kernels and copies get that address, never the name.
//
// Every call on a GPU shares that GPU's array, and the array keeps the last
// call's contents.  Nothing in it must start at zero: the orbit kernel writes
// every entry 0..M and then M before the pixel kernel (ordered after it on the
// same stream) reads them.  Calls that overlap on one GPU (other streams, other
// threads) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../kernels/perturb_core.hpp"
#include "common.hpp"
#include "engine_helpers.hpp"
#include "handles.hpp"

namespace {

// The orbit at its widest (FP64 entries, 32 B), then M in its own 256-byte line.
constexpr size_t kOrbitBytes = ((size_t)CU_DEEP_STATIC_MAX_ITER + 1) * sizeof(OrbitEntry<double>);
constexpr size_t kScratchBytes = kOrbitBytes + 256;
static_assert(kScratchBytes == CU_DEEP_STATIC_BYTES, "scratch size in the header");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

// The address of `scratch` on `device`, which must be current.
void *scratch_on(int device, bool lookup) {
    if (device < 0 || device >= kMaxDevices) cu::fail(CU_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device] && lookup) {
        void *p = nullptr;
        cu::check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch[device] = p;
    }
    return g_scratch[device];
}

template <class Real>
__global__ __launch_bounds__(64) void staticReferenceOrbit(double cx_hi, double cx_lo,
                                                           double cy_hi, double cy_lo,
                                                           uint max_iter, double bailout_sq,
                                                           OrbitEntry<Real> *orbit, uint *length) {
    reference_orbit<Real>(cx_hi, cx_lo, cy_hi, cy_lo, max_iter, bailout_sq, orbit, length);
}

// The block shapes and LDS stages of perturbFloat / perturbDouble (perturb.hip).
__global__ __launch_bounds__(256, 8) void staticPerturbFloat(DeepArgs a,
                                                             const OrbitEntry<float> *__restrict__ orbit) {
    perturb_frame<float, 32, 2, 2, 20u * 1024u>(a, orbit);
}
__global__ __launch_bounds__(512, 8) void staticPerturbDouble(DeepArgs a,
                                                              const OrbitEntry<double> *__restrict__ orbit) {
    perturb_frame<double, 32, 4, 2, 40u * 1024u>(a, orbit);
}

template <class Real>
void render(const cu_deep_desc &c, unsigned char *base, uint32_t *counts, double *smooth,
            const uint32_t *palette, uint32_t palette_length, uint32_t *rgba,
            unsigned long long *iters, hipStream_t st) {
    OrbitEntry<Real> *orbit = (OrbitEntry<Real> *)base;
    uint *len = (uint *)(base + kOrbitBytes);
    hipLaunchKernelGGL(staticReferenceOrbit<Real>, dim3(1), dim3(64), 0, st, c.centre_x,
                       c.centre_x_lo, c.centre_y, c.centre_y_lo, c.max_iter, c.bailout_sq, orbit,
                       len);
    cu::check(hipGetLastError(), "launch staticReferenceOrbit");
    DeepArgs a{};
    a.orbit = orbit;
    a.orbit_len = len;
    a.counts = counts;
    a.smooth = smooth;
    a.glitch = nullptr;
    a.rgba = rgba;
    a.palette = palette;
    a.iters = iters;
    a.centre_x = c.centre_x;
    a.centre_y = c.centre_y;
    a.pixel_w = c.pixel_w;
    a.pixel_h = c.pixel_h;
    a.bailout_sq = c.bailout_sq;
    a.smooth_c = cu::smooth_constant(c.bailout_sq);
    a.width = c.width;
    a.height = c.height;
    a.max_iter = c.max_iter;
    a.flags = c.flags;
    a.palette_length = palette ? palette_length : 0u;
    a.glitch_eps = c.glitch_eps;
    const uint32_t bw = sizeof(Real) == 8 ? 32 : 16, bh = 16;
    const dim3 grid((c.width + bw - 1) / bw, (c.height + bh - 1) / bh, 1);
    if constexpr (sizeof(Real) == 8)
        hipLaunchKernelGGL(staticPerturbDouble, grid, dim3(bw * bh), 0, st, a, orbit);
    else
        hipLaunchKernelGGL(staticPerturbFloat, grid, dim3(bw * bh), 0, st, a, orbit);
    cu::check(hipGetLastError(), "launch staticPerturb");
}

// cu_init looks the address up on the device it selects (with that device current).
struct RegisterSetup {
    RegisterSetup() {
        cu::on_device_init([](int device) { (void)scratch_on(device, true); });
    }
} g_register_setup;

}  // namespace

using namespace cu;

extern "C" int cu_deep_render_static(int method, const cu_deep_desc *desc, uint32_t *d_counts,
                                     double *d_smooth, const uint32_t *d_palette,
                                     uint32_t palette_length, uint32_t *d_rgba,
                                     unsigned long long *d_iters, cu_stream stream) {
    return guarded([&] {
        if (!desc || !d_counts) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        if (method != CU_DEEP_PERTURB_F32 && method != CU_DEEP_PERTURB_F64)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "method must be CU_DEEP_PERTURB_F32 or _F64");
        validate_deep(*desc, method);
        if (!(desc->flags & CU_DEEP_REBASE))
            fail(CU_ERR_ILLEGAL_ARGUMENT,
                 "cu_deep_render_static needs CU_DEEP_REBASE (glitch rounds read flags back)");
        if (desc->max_iter > CU_DEEP_STATIC_MAX_ITER)
            fail(CU_ERR_ILLEGAL_ARGUMENT,
                 "max_iter " + std::to_string(desc->max_iter) + " above CU_DEEP_STATIC_MAX_ITER (" +
                     std::to_string(CU_DEEP_STATIC_MAX_ITER) +
                     "): the static scratch holds that many orbit entries plus one");
        if ((d_rgba == nullptr) != (d_palette == nullptr))
            fail(CU_ERR_ILLEGAL_ARGUMENT, "palette and rgba go together");
        if (d_palette && palette_length == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "empty palette");
        ensure_device();
        unsigned char *base = (unsigned char *)scratch_on(current_device(), true);
        hipStream_t st = stream ? stream->stream : nullptr;
        if (method == CU_DEEP_PERTURB_F32)
            render<float>(*desc, base, d_counts, d_smooth, d_palette, palette_length, d_rgba,
                          d_iters, st);
        else
            render<double>(*desc, base, d_counts, d_smooth, d_palette, palette_length, d_rgba,
                           d_iters, st);
    });
}
