// cu_types.hpp — value types shared by the gfx950 kernels and their host launcher.
//
// The reference's device types (src/main/cuda/helpers.cuh:22-130) restated as
// plain aggregates so a kernel argument has the same bytes as the reference's:
// Point<uint> is 8 B, Rectangle<double> is 32 B (left_bottom, right_top), and
// pixel_info_t is the 16-byte record {value, weight, isReused, weightOfNewSamples}.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/chaosultra.h"

typedef unsigned int uint;

template <class T>
struct Point {
    T x, y;
};

template <class T>
struct Rectangle {
    Point<T> left_bottom;
    Point<T> right_top;
};

struct pixel_info_t {
    float value;               // the fractal value (mean escape count)
    float weight;              // samples behind the value (decays under reuse)
    bool isReused;             // value (partly) taken from the previous frame
    float weightOfNewSamples;  // fresh samples in a reused pixel
};
static_assert(sizeof(pixel_info_t) == 16, "pixel_info_t must stay 16 bytes");
static_assert(sizeof(Point<uint>) == 8, "Point<uint> must stay 8 bytes");
static_assert(sizeof(Rectangle<double>) == 32, "Rectangle<double> must stay 32 bytes");
static_assert(sizeof(cu_tile_desc) == 128, "cu_tile_desc must stay 128 bytes");

constexpr uint kMaxSuperSampling = 64;   // fractalRendererGeneric.cu:12
constexpr uint kAdaptiveThreshold = 10;  // fractalRendererGeneric.cu:93
constexpr uint kVoteGroupW = 8;          // a 32-lane vote group covers 8x4 pixels
constexpr uint kVoteGroupH = 4;
