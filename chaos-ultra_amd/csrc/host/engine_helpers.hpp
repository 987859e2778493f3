// engine_helpers.hpp — helpers shared by the escape-family launchers (engine.cpp, frame.cpp).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <string>

#include "../../../include/chaosultra.h"

namespace cu {

constexpr uint32_t kMaxPersistentTiles = 1024;  // tile tables staged in LDS

cu_function function(cu_engine e, const std::string &name);
void launch(cu_engine e, cu_function f, dim3 grid, dim3 block, void **params, cu_stream s,
            bool timed);
const char *real_name(int precision);
bool fused_safe(const cu_tile_desc &t, int precision);
void validate_tile(const cu_tile_desc &t);
void validate_tune(const cu_tune &tn);
// the checks cu_deep_create runs on a deep-frame description (CU_ERR_ILLEGAL_ARGUMENT)
void validate_deep(const cu_deep_desc &d, int method);
bool pick_fused(int variant, const cu_tile_desc &t, int precision);
void *device_alloc(size_t bytes);  // hipMalloc, nullptr for 0 bytes; throws CuError
// fill `words` u32 at p with `value` by the fillU32 kernel (safe under graph capture)
void fill_u32(cu_engine e, cu_stream s, void *p, uint32_t value, uint64_t words);

// C = log2(log2(R^2)), the per-tile term of the smooth value
// n + 1 - (log2(log2(|z|^2)) - C) computed by the kernels.
inline double smooth_constant(double bailout_sq) { return std::log2(std::log2(bailout_sq)); }

}  // namespace cu
