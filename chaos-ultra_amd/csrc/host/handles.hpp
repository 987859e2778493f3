// handles.hpp — the structs behind the opaque driver-layer handles.
#pragma once
#include <hip/hip_runtime_api.h>

#include <map>
#include <string>

struct cu_function_s {
    hipFunction_t function;
    std::string name;
};

struct cu_module_s {
    hipModule_t module;
    std::string path;
    std::map<std::string, cu_function_s *> functions;  // owned; freed on unload
};

struct cu_stream_s {
    hipStream_t stream;
};

#include <mutex>

#include "../../../include/chaosultra.h"

struct cu_engine_s {
    cu_module module = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    bool timed = false;
    cu_tune tune{16, 32, 256, 16};  // defaults of every launch (cu_engine_configure)
    std::mutex mutex;
};

// A captured launch sequence (cu_escape_tiles_persistent_capture) and the events
// around its last replay.
struct cu_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    bool launched = false;
};
