// driver.cpp — layer 1 of the C ABI: the driver calls the reference's JCuda
// binding makes on the render path, implemented on the HIP runtime.
//
// Each entry point names the JCuda call it replaces and the reference site that
// makes it.  Argument validation that the reference does on the host (missing
// module file, unknown names, undersized globals) happens before any device
// call, so it is testable without a GPU.
#include <sys/stat.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "handles.hpp"

namespace cu {

namespace {
thread_local std::string t_last_error;
std::mutex g_init_mutex;
std::atomic<int> g_device{-1};
int g_compute_units = 0;
}  // namespace

void set_last_error(const std::string &message) { t_last_error = message; }

int current_device() { return g_device.load(); }

std::vector<void (*)(int)> &device_setups() {
    static std::vector<void (*)(int)> steps;
    return steps;
}
void on_device_init(void (*setup)(int device)) { device_setups().push_back(setup); }
int compute_units() { return g_compute_units; }

void ensure_device() {
    int d = g_device.load();
    if (d < 0) {
        if (cu_init(0) != CU_OK) fail(CU_ERR_INITIALIZATION, t_last_error);
        d = g_device.load();
    }
    check(hipSetDevice(d), "hipSetDevice", CU_ERR_INITIALIZATION);
}

}  // namespace cu

using namespace cu;

extern "C" {

const char *cu_last_error(void) { return t_last_error.c_str(); }

// CudaHelpers.cudaInit (cudarenderer/CudaHelpers.java:27-45)
int cu_init(int device) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_init_mutex);
        if (g_device.load() >= 0) {
            if (g_device.load() != device)
                fail(CU_ERR_ILLEGAL_STATE, "already initialised on device " +
                                               std::to_string(g_device.load()));
            return;
        }
        check(hipInit(0), "hipInit (is a ROCm device present?)", CU_ERR_INITIALIZATION);
        int n = 0;
        check(hipGetDeviceCount(&n), "hipGetDeviceCount", CU_ERR_INITIALIZATION);
        if (device < 0 || device >= n)
            fail(CU_ERR_INITIALIZATION,
                 "no device " + std::to_string(device) + " (" + std::to_string(n) + " visible)");
        check(hipSetDevice(device), "hipSetDevice", CU_ERR_INITIALIZATION);
        int cus = 0;
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device),
              "hipDeviceGetAttribute", CU_ERR_INITIALIZATION);
        g_compute_units = cus;
        g_device.store(device);
        for (void (*setup)(int) : device_setups()) {
            try {
                setup(device);
            } catch (const CuError &) {
            }
        }
    });
}

int cu_device_count(int *count) {
    return guarded([&] {
        if (!count) fail(CU_ERR_ILLEGAL_ARGUMENT, "count is NULL");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) {
            (void)hipGetLastError();
            n = 0;
        }
        *count = n;
    });
}

// cuModuleLoad (cudarenderer/FractalRenderingModule.java:73-88)
int cu_module_load(const char *path, cu_module *out) {
    return guarded([&] {
        if (!path || !out) fail(CU_ERR_ILLEGAL_ARGUMENT, "path/out is NULL");
        struct stat st;
        if (stat(path, &st) != 0 || !S_ISREG(st.st_mode))
            fail(CU_ERR_ILLEGAL_ARGUMENT, std::string("Invalid module file name: ") + path +
                                              "\nHave you set the kernels directory properly?");
        ensure_device();
        hipModule_t m = nullptr;
        hipError_t rc = hipModuleLoad(&m, path);
        if (rc == hipErrorInvalidContext) check(rc, "Invalid HIP context", CU_ERR_INITIALIZATION);
        check(rc, (std::string("hipModuleLoad ") + path).c_str(), CU_ERR_INITIALIZATION);
        *out = new cu_module_s{m, path, {}};
    });
}

// cuModuleUnload (cudarenderer/FractalRenderingModule.java:140-147)
int cu_module_unload(cu_module module) {
    return guarded([&] {
        if (!module) fail(CU_ERR_ILLEGAL_ARGUMENT, "module is NULL");
        ensure_device();
        for (auto &kv : module->functions) delete kv.second;
        hipError_t rc = hipModuleUnload(module->module);
        delete module;
        check(rc, "hipModuleUnload");
    });
}

// cuModuleGetFunction (cudarenderer/CudaKernel.java:28-43)
int cu_module_get_function(cu_module module, const char *name, cu_function *out) {
    return guarded([&] {
        if (!module || !name || !out) fail(CU_ERR_ILLEGAL_ARGUMENT, "module/name/out is NULL");
        auto it = module->functions.find(name);
        if (it != module->functions.end()) {
            *out = it->second;
            return;
        }
        ensure_device();
        hipFunction_t f = nullptr;
        hipError_t rc = hipModuleGetFunction(&f, module->module, name);
        if (rc != hipSuccess) {
            (void)hipGetLastError();
            fail(CU_ERR_ILLEGAL_ARGUMENT, std::string("Function with this name not found: ") + name);
        }
        auto *fn = new cu_function_s{f, name};
        module->functions[name] = fn;
        *out = fn;
    });
}

// cuModuleGetGlobal (cudarenderer/FractalRenderingModule.java:163-169)
int cu_module_get_global(cu_module module, const char *name, void **dptr, size_t *bytes) {
    return guarded([&] {
        if (!module || !name || !dptr || !bytes)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "module/name/dptr/bytes is NULL");
        ensure_device();
        hipDeviceptr_t p = nullptr;
        size_t n = 0;
        hipError_t rc = hipModuleGetGlobal(&p, &n, module->module, name);
        if (rc != hipSuccess) {
            (void)hipGetLastError();
            fail(CU_ERR_ILLEGAL_ARGUMENT, std::string("No module global named ") + name);
        }
        *dptr = (void *)p;
        *bytes = n;
    });
}

// writeToConstantMemory (cudarenderer/FractalRenderingModule.java:171-227)
int cu_module_write_global(cu_module module, const char *name, const void *value, size_t bytes) {
    return guarded([&] {
        if (!value && bytes) fail(CU_ERR_ILLEGAL_ARGUMENT, "value is NULL");
        void *dptr = nullptr;
        size_t size = 0;
        if (cu_module_get_global(module, name, &dptr, &size) != CU_OK)
            fail(CU_ERR_ILLEGAL_ARGUMENT, t_last_error);
        if (size < bytes)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "Attempt to write " + std::to_string(bytes) +
                                              " bytes to device memory allocated to size " +
                                              std::to_string(size));
        std::string staged(size, '\0');
        if (bytes) std::memcpy(&staged[0], value, bytes);
        check(hipMemcpyHtoD((hipDeviceptr_t)dptr, &staged[0], size), "hipMemcpyHtoD");
    });
}

int cu_mem_alloc(void **dptr, size_t bytes) {
    return guarded([&] {
        if (!dptr) fail(CU_ERR_ILLEGAL_ARGUMENT, "dptr is NULL");
        *dptr = nullptr;
        if (bytes == 0) return;
        ensure_device();
        check(hipMalloc(dptr, bytes), "hipMalloc");
    });
}

// cuMemAllocPitch (cudarenderer/DeviceMemoryDoubleBuffer2D.java:116-146)
int cu_mem_alloc_pitch(void **dptr, size_t *pitch, size_t width_bytes, size_t height,
                       unsigned element_bytes) {
    return guarded([&] {
        if (!dptr || !pitch) fail(CU_ERR_ILLEGAL_ARGUMENT, "dptr/pitch is NULL");
        if (element_bytes == 0 || width_bytes % element_bytes)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "width_bytes must be a multiple of element_bytes");
        *dptr = nullptr;
        *pitch = 0;
        if (width_bytes == 0 || height == 0) return;  // the reference allocates nothing
        ensure_device();
        check(hipMallocPitch(dptr, pitch, width_bytes, height), "hipMallocPitch");
        if (*pitch == 0) fail(CU_ERR_DEVICE, "hipMallocPitch returned pitch 0");
    });
}

int cu_mem_free(void *dptr) {
    return guarded([&] {
        if (!dptr) return;
        ensure_device();
        check(hipFree(dptr), "hipFree");
    });
}

int cu_host_alloc(void **ptr, size_t bytes) {
    return guarded([&] {
        if (!ptr) fail(CU_ERR_ILLEGAL_ARGUMENT, "ptr is NULL");
        *ptr = nullptr;
        if (!bytes) fail(CU_ERR_ILLEGAL_ARGUMENT, "bytes must be > 0");
        ensure_device();
        check(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
    });
}

int cu_host_free(void *ptr) {
    return guarded([&] {
        if (!ptr) return;
        ensure_device();
        check(hipHostFree(ptr), "hipHostFree");
    });
}

int cu_memcpy_htod(void *dst, const void *src, size_t bytes) {
    return guarded([&] {
        if (!bytes) return;
        ensure_device();
        check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    });
}

int cu_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    return guarded([&] {
        if (!bytes) return;
        ensure_device();
        check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    });
}

int cu_memcpy_dtod(void *dst, const void *src, size_t bytes) {
    return guarded([&] {
        if (!bytes) return;
        ensure_device();
        check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
    });
}

// cuMemcpy2D device -> host (cudarenderer/CudaFractalRenderer.java:119-134)
int cu_memcpy_2d_dtoh(void *dst, size_t dst_pitch, const void *src, size_t src_pitch,
                      size_t width_bytes, size_t height) {
    return guarded([&] {
        if (!width_bytes || !height) return;
        ensure_device();
        check(hipMemcpy2D(dst, dst_pitch, src, src_pitch, width_bytes, height,
                          hipMemcpyDeviceToHost),
              "hipMemcpy2D D2H");
    });
}

int cu_memset_u32(void *dptr, uint32_t value, size_t count) {
    return guarded([&] {
        if (!count) return;
        ensure_device();
        check(hipMemsetD32((hipDeviceptr_t)dptr, (int)value, count), "hipMemsetD32");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

// cuLaunchKernel on the null stream (cudarenderer/CudaFractalRenderer.java:371-378)
int cu_launch_kernel(cu_function f, unsigned gx, unsigned gy, unsigned gz, unsigned bx,
                     unsigned by, unsigned bz, unsigned shared_bytes, void **params) {
    return guarded([&] {
        if (!f) fail(CU_ERR_ILLEGAL_ARGUMENT, "function is NULL");
        if (!gx || !gy || !gz || !bx || !by || !bz)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "grid and block dimensions must be positive");
        if ((uint64_t)bx * by * bz > 1024)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "block larger than 1024 threads");
        ensure_device();
        check(hipModuleLaunchKernel(f->function, gx, gy, gz, bx, by, bz, shared_bytes, nullptr,
                                    params, nullptr),
              (std::string("hipModuleLaunchKernel ") + f->name).c_str());
    });
}

// cuCtxSynchronize
int cu_ctx_synchronize(void) {
    return guarded([&] {
        ensure_device();
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    });
}

int cu_stream_create(cu_stream *out) {
    return guarded([&] {
        if (!out) fail(CU_ERR_ILLEGAL_ARGUMENT, "out is NULL");
        ensure_device();
        hipStream_t s = nullptr;
        check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        *out = new cu_stream_s{s};
    });
}

int cu_stream_destroy(cu_stream s) {
    return guarded([&] {
        if (!s) return;
        ensure_device();
        hipError_t rc = hipStreamDestroy(s->stream);
        delete s;
        check(rc, "hipStreamDestroy");
    });
}

int cu_stream_synchronize(cu_stream s) {
    return guarded([&] {
        ensure_device();
        check(hipStreamSynchronize(s ? s->stream : nullptr), "hipStreamSynchronize");
    });
}

int cu_stream_query(cu_stream s, int *done) {
    return guarded([&] {
        if (!done) fail(CU_ERR_ILLEGAL_ARGUMENT, "done is NULL");
        ensure_device();
        hipError_t rc = hipStreamQuery(s ? s->stream : nullptr);
        if (rc == hipErrorNotReady) {
            (void)hipGetLastError();
            *done = 0;
            return;
        }
        check(rc, "hipStreamQuery");
        *done = 1;
    });
}

}  // extern "C"
