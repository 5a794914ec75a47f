/* Implementation of the bhhip C-ABI over the HIP runtime and hiprtc. */
#include "../../../include/bhhip.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;
std::mutex g_mu;
bool g_init = false;
int g_device = 0;
hipStream_t g_stream = nullptr;
std::string g_arch;
std::string g_info;
uint64_t g_compiles = 0, g_mem_hits = 0, g_disk_hits = 0;

// One loaded code object per (source, options); its kernels are looked up by name.
struct Loaded {
    hipModule_t module = nullptr;
    std::unordered_map<std::string, hipFunction_t> functions;
};
std::unordered_map<uint64_t, Loaded> g_modules;

int fail(hipError_t e, const char *what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return static_cast<int>(e);
}
#define HIPCHK(expr, what)                       \
    do {                                         \
        hipError_t e_ = (expr);                  \
        if (e_ != hipSuccess) return fail(e_, what); \
    } while (0)

// FNV-1a 64-bit.
uint64_t fnv1a(const std::string &s, uint64_t h = 1469598103934665603ull) {
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

bool read_file(const std::string &path, std::vector<char> &out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    out.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    return !out.empty();
}

void write_file_atomic(const std::string &dir, const std::string &path, const std::vector<char> &data) {
    mkdir(dir.c_str(), 0755);
    const std::string tmp = path + ".tmp" + std::to_string(getpid());
    {
        std::ofstream out(tmp, std::ios::binary);
        if (!out) return;
        out.write(data.data(), static_cast<std::streamsize>(data.size()));
    }
    std::rename(tmp.c_str(), path.c_str());
}

const char *kOptions[] = {"-O3", "-ffp-contract=fast", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                          "-std=c++17"};

} // namespace

extern "C" {

int bhhip_init(int device) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_init) return 0;
    HIPCHK(hipSetDevice(device), "hipSetDevice");
    g_device = device;
    HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking), "hipStreamCreate");
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    g_arch = prop.gcnArchName;
    const size_t colon = g_arch.find(':');
    if (colon != std::string::npos) g_arch = g_arch.substr(0, colon);
    std::ostringstream ss;
    ss << prop.name << " (" << g_arch << "), " << prop.multiProcessorCount << " CUs, "
       << prop.totalGlobalMem / (1024 * 1024) << " MiB, clock " << prop.clockRate / 1000 << " MHz";
    g_info = ss.str();
    g_init = true;
    return 0;
}

const char *bhhip_last_error(void) { return g_err.c_str(); }

const char *bhhip_device_info(void) { return g_info.c_str(); }

int bhhip_device_props(int *num_cus, uint64_t *total_mem) {
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, g_device), "hipGetDeviceProperties");
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (total_mem) *total_mem = prop.totalGlobalMem;
    return 0;
}

int bhhip_malloc(void **dptr, size_t bytes) {
    HIPCHK(hipMalloc(dptr, bytes == 0 ? 1 : bytes), "hipMalloc");
    return 0;
}

int bhhip_free(void *dptr) {
    HIPCHK(hipFree(dptr), "hipFree");
    return 0;
}

int bhhip_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return 0;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream), "hipMemcpyAsync(H2D)");
    // The host buffer may be reused by the caller right away.
    HIPCHK(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
    return 0;
}

int bhhip_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return 0;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream), "hipMemcpyAsync(D2H)");
    HIPCHK(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
    return 0;
}

int bhhip_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return 0;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream), "hipMemcpyAsync(D2D)");
    return 0;
}

int bhhip_memcpy_h2d_async(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return 0;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream), "hipMemcpyAsync(H2D)");
    return 0;
}

int bhhip_memzero(void *dptr, size_t bytes) {
    if (bytes == 0) return 0;
    HIPCHK(hipMemsetAsync(dptr, 0, bytes, g_stream), "hipMemsetAsync");
    HIPCHK(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
    return 0;
}

int bhhip_host_alloc(void **ptr, size_t bytes) {
    HIPCHK(hipHostMalloc(ptr, bytes == 0 ? 1 : bytes, hipHostMallocDefault), "hipHostMalloc");
    return 0;
}

int bhhip_host_free(void *ptr) {
    HIPCHK(hipHostFree(ptr), "hipHostFree");
    return 0;
}

int bhhip_set_arch(const char *arch) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_init) {
        g_err = "bhhip_set_arch: the device is already initialised";
        return 1;
    }
    g_arch = arch;
    return 0;
}

int bhhip_compile(const char *prelude, const char *source, const char *name, const char *cache_dir,
                  void **function) {
    if (g_arch.empty()) {
        g_err = "bhhip_compile: no target (call bhhip_init or bhhip_set_arch)";
        return 1;
    }
    std::string full = std::string(prelude) + "\n" + source;
    std::string opts_key = g_arch;
    for (const char *o : kOptions) opts_key += std::string(" ") + o;
    const uint64_t key = fnv1a(full, fnv1a(opts_key));
    if (function != nullptr) {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_modules.find(key);
        if (it != g_modules.end()) {
            ++g_mem_hits;
            auto f = it->second.functions.find(name);
            if (f == it->second.functions.end()) {
                hipFunction_t fn;
                HIPCHK(hipModuleGetFunction(&fn, it->second.module, name), "hipModuleGetFunction");
                f = it->second.functions.emplace(name, fn).first;
            }
            *function = reinterpret_cast<void *>(f->second);
            return 0;
        }
    }
    std::vector<char> code;
    std::string path;
    const std::string dir = cache_dir == nullptr ? "" : cache_dir;
    if (!dir.empty()) {
        char hex[32];
        std::snprintf(hex, sizeof(hex), "%016llx", static_cast<unsigned long long>(key));
        path = dir + "/" + hex + ".hsaco";
        if (read_file(path, code)) ++g_disk_hits;
    }
    if (code.empty()) {
        hiprtcProgram prog;
        hiprtcResult r = hiprtcCreateProgram(&prog, full.c_str(), "bh_kernel.hip", 0, nullptr, nullptr);
        if (r != HIPRTC_SUCCESS) {
            g_err = std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r);
            return 1000 + static_cast<int>(r);
        }
        std::vector<const char *> opts;
        const std::string arch = "--offload-arch=" + g_arch;
        opts.push_back(arch.c_str());
        for (const char *o : kOptions) opts.push_back(o);
        r = hiprtcCompileProgram(prog, static_cast<int>(opts.size()), opts.data());
        if (r != HIPRTC_SUCCESS) {
            size_t n = 0;
            hiprtcGetProgramLogSize(prog, &n);
            std::string log(n, '\0');
            if (n) hiprtcGetProgramLog(prog, &log[0]);
            g_err = std::string("hiprtc compile failed: ") + log + "\n---- source ----\n" + source;
            hiprtcDestroyProgram(&prog);
            return 1000 + static_cast<int>(r);
        }
        size_t cs = 0;
        hiprtcGetCodeSize(prog, &cs);
        code.resize(cs);
        hiprtcGetCode(prog, code.data());
        hiprtcDestroyProgram(&prog);
        ++g_compiles;
        if (!path.empty()) write_file_atomic(dir, path, code);
    }
    if (function == nullptr) return 0;  // compile only
    Loaded l;
    hipFunction_t fn;
    HIPCHK(hipModuleLoadData(&l.module, code.data()), "hipModuleLoadData");
    HIPCHK(hipModuleGetFunction(&fn, l.module, name), "hipModuleGetFunction");
    l.functions[name] = fn;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        g_modules[key] = l;
    }
    *function = reinterpret_cast<void *>(fn);
    return 0;
}

void bhhip_compile_stats(uint64_t *compiles, uint64_t *memory_hits, uint64_t *disk_hits) {
    if (compiles) *compiles = g_compiles;
    if (memory_hits) *memory_hits = g_mem_hits;
    if (disk_hits) *disk_hits = g_disk_hits;
}

int bhhip_launch(void *function, uint64_t grid, uint32_t block, const void *args, size_t args_bytes) {
    size_t size = args_bytes;
    void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void *>(args), HIP_LAUNCH_PARAM_BUFFER_SIZE, &size,
                      HIP_LAUNCH_PARAM_END};
    if (grid == 0) return 0;
    if (grid > 0x7fffffffull) {
        g_err = "bhhip_launch: grid too large";
        return 1;
    }
    HIPCHK(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(function), static_cast<unsigned>(grid), 1, 1, block,
                                 1, 1, 0, g_stream, nullptr, config),
           "hipModuleLaunchKernel");
    return 0;
}

int bhhip_launch_3d(void *function, const uint32_t grid[3], const uint32_t block[3], const void *args,
                    size_t args_bytes) {
    size_t size = args_bytes;
    void *config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void *>(args), HIP_LAUNCH_PARAM_BUFFER_SIZE, &size,
                      HIP_LAUNCH_PARAM_END};
    if (grid[0] == 0 || grid[1] == 0 || grid[2] == 0) return 0;
    HIPCHK(hipModuleLaunchKernel(reinterpret_cast<hipFunction_t>(function), grid[0], grid[1], grid[2], block[0],
                                 block[1], block[2], 0, g_stream, nullptr, config),
           "hipModuleLaunchKernel");
    return 0;
}

int bhhip_launch_cooperative(void *function, uint32_t grid, uint32_t block, void **params) {
    if (grid == 0) return 0;
    HIPCHK(hipModuleLaunchCooperativeKernel(reinterpret_cast<hipFunction_t>(function), grid, 1, 1, block, 1, 1, 0,
                                            g_stream, params),
           "hipModuleLaunchCooperativeKernel");
    return 0;
}

void *bhhip_stream(void) { return g_stream; }

int bhhip_occupancy(void *function, uint32_t block, int *blocks_per_cu) {
    int n = 0;
    HIPCHK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<hipFunction_t>(function),
                                                              static_cast<int>(block), 0),
           "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
    *blocks_per_cu = n;
    return 0;
}

int bhhip_synchronize(void) {
    HIPCHK(hipStreamSynchronize(g_stream), "hipStreamSynchronize");
    return 0;
}

int bhhip_event_create(void **event) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e), "hipEventCreate");
    *event = e;
    return 0;
}

int bhhip_event_record(void *event) {
    HIPCHK(hipEventRecord(static_cast<hipEvent_t>(event), g_stream), "hipEventRecord");
    return 0;
}

int bhhip_event_elapsed_ms(void *start, void *stop, float *ms) {
    HIPCHK(hipEventSynchronize(static_cast<hipEvent_t>(stop)), "hipEventSynchronize");
    HIPCHK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)),
           "hipEventElapsedTime");
    return 0;
}

int bhhip_event_done(void *event, int *done) {
    const hipError_t e = hipEventQuery(static_cast<hipEvent_t>(event));
    if (e == hipErrorNotReady) {
        *done = 0;
        return 0;
    }
    HIPCHK(e, "hipEventQuery");
    *done = 1;
    return 0;
}

int bhhip_event_synchronize(void *event) {
    HIPCHK(hipEventSynchronize(static_cast<hipEvent_t>(event)), "hipEventSynchronize");
    return 0;
}

int bhhip_event_destroy(void *event) {
    HIPCHK(hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
    return 0;
}

int bhhip_capture_begin(void) {
    HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    return 0;
}

int bhhip_capture_end(void **graph_exec) {
    hipGraph_t graph;
    HIPCHK(hipStreamEndCapture(g_stream, &graph), "hipStreamEndCapture");
    hipGraphExec_t exec;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(e, "hipGraphInstantiate");
    *graph_exec = exec;
    return 0;
}

int bhhip_graph_launch(void *graph_exec) {
    HIPCHK(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), g_stream), "hipGraphLaunch");
    return 0;
}

int bhhip_graph_destroy(void *graph_exec) {
    HIPCHK(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)), "hipGraphExecDestroy");
    return 0;
}

} // extern "C"
