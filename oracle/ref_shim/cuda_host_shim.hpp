// cuda_host_shim.hpp — lets the reference's per-fractal CUDA source compile as
// host C++ (g++), so its own computeFractal can pin the oracle.
//
// TEST INFRASTRUCTURE ONLY.  This header is force-included before the
// reference's file; it supplies the few CUDA names the device code uses and
// nothing else.  The reference sources are read from /root/reference by
// oracle/Makefile and never copied.
#pragma once
#include <cassert>
#include <cmath>
#include <cstdio>

#define CUDART_VERSION 10000
#define __device__
#define __host__
#define __constant__
#define __forceinline__ inline

typedef unsigned long long cudaSurfaceObject_t;
template <class T>
inline void surf2Dread(T *out, cudaSurfaceObject_t, int, int) {
    *out = T();
}

using std::round;
using std::sqrt;
