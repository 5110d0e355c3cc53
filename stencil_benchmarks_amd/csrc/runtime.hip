// Library lifecycle, errors, device memory and cache flushing.
#include <cstring>
#include <mutex>
#include <vector>

#include "sb_internal.h"

namespace sb {

thread_local ExtEvents *t_ext_events = nullptr;

namespace {
thread_local std::string g_error;
std::mutex g_ctx_mutex;
std::vector<Context> g_contexts;

// Read every 16-byte chunk of the flush buffer: the lines displace the
// caller's data from L2 and the Infinity Cache, and, unlike a
// read-modify-write sweep, leave no dirty lines whose write-back would compete
// with the next (timed) kernel for HBM.  The buffer holds zeros, so the
// conditional store never happens; it keeps the loads alive.
__global__ __launch_bounds__(256) void flush_kernel(uint4 *buf, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
    for (; i < n; i += stride) {
        const uint4 v = buf[i];  // plain load: the line must allocate in the caches
        acc.x ^= v.x;
        acc.y ^= v.y;
        acc.z ^= v.z;
        acc.w ^= v.w;
    }
    if ((acc.x & acc.y & acc.z & acc.w) == 0xFFFFFFFFu) buf[0] = acc;
}

int init_context(Context &c, int device) {
    c.device = device;
    SB_HIP(hipSetDevice(device));
    SB_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    SB_HIP(hipEventCreate(&c.start));
    SB_HIP(hipEventCreate(&c.stop));
    hipDeviceProp_t p;
    SB_HIP(hipGetDeviceProperties(&p, device));
    c.compute_units = p.multiProcessorCount;
    c.lds_per_block = p.sharedMemPerBlock;
    // CDNA3/CDNA4 parts split their CUs over 8 accelerator complex dies.
    c.xcds = (std::strncmp(p.gcnArchName, "gfx94", 5) == 0 ||
              std::strncmp(p.gcnArchName, "gfx95", 5) == 0)
                 ? 8
                 : 1;
    vadv_noalloc_setup(c);
    return 0;
}
}  // namespace

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

Context *context() {
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) {
        (void)hipGetLastError();
        set_error("no HIP device available");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    if (int(g_contexts.size()) <= device) g_contexts.resize(device + 1);
    Context &c = g_contexts[device];
    if (c.device < 0) {
        if (init_context(c, device)) {
            c = Context{};
            return nullptr;
        }
    }
    return &c;
}

int flush_caches(Context *ctx) {
    // 1 GiB: four times the Infinity Cache plus all L2s.
    const size_t bytes = size_t(1) << 30;
    if (!ctx->flush_buffer) {
        SB_HIP(hipMalloc(&ctx->flush_buffer, bytes));
        SB_HIP(hipMemsetAsync(ctx->flush_buffer, 0, bytes, ctx->stream));
        ctx->flush_bytes = bytes;
    }
    const size_t n = ctx->flush_bytes / sizeof(uint4);
    hipLaunchKernelGGL(flush_kernel, dim3(ctx->compute_units * 8), dim3(256), 0,
                       ctx->stream, static_cast<uint4 *>(ctx->flush_buffer), n);
    SB_HIP(hipGetLastError());
    return 0;
}

int check_geom(const sb_geom *g, int64_t need_hi, int64_t need_hj,
               int64_t need_hk) {
    SB_REQUIRE(g, "geometry pointer is NULL");
    SB_REQUIRE(g->nx >= 0 && g->ny >= 0 && g->nz >= 0,
               "negative domain size (%lld, %lld, %lld)", (long long)g->nx,
               (long long)g->ny, (long long)g->nz);
    SB_REQUIRE(g->hi >= need_hi && g->hj >= need_hj && g->hk >= need_hk,
               "halo (%lld, %lld, %lld) smaller than the stencil needs "
               "(%lld, %lld, %lld)",
               (long long)g->hi, (long long)g->hj, (long long)g->hk,
               (long long)need_hi, (long long)need_hj, (long long)need_hk);
    SB_REQUIRE(g->si != 0 && g->sj != 0 && g->sk != 0, "zero stride");
    return 0;
}

int vector_width(size_t elem_bytes, const void *const *ptrs, int nptrs,
                 const sb_geom *g) {
    if (g->si != 1) return 1;
    int v = int(16 / elem_bytes);
    for (; v > 1; v /= 2) {
        const size_t vb = size_t(v) * elem_bytes;
        bool ok = (g->sj % v == 0) && (g->sk % v == 0);
        for (int p = 0; ok && p < nptrs; ++p)
            ok = (reinterpret_cast<uintptr_t>(ptrs[p]) % vb) == 0;
        if (ok) break;
    }
    return v;
}

}  // namespace sb

using namespace sb;

extern "C" {

const char *sb_last_error(void) { return g_error.c_str(); }

int sb_abi_version(void) { return SB_ABI_VERSION; }

int sb_device_count(int *count) {
    SB_REQUIRE(count, "count pointer is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return 0;
}

int sb_init(int device) {
    int n = 0;
    SB_HIP(hipGetDeviceCount(&n));
    SB_REQUIRE(device >= 0 && device < n, "device %d not present (%d devices)",
               device, n);
    SB_HIP(hipSetDevice(device));
    return context() ? 0 : 1;
}

int sb_get_device_info(int device, sb_device_info *info) {
    SB_REQUIRE(info, "info pointer is NULL");
    hipDeviceProp_t p;
    SB_HIP(hipGetDeviceProperties(&p, device));
    std::memset(info, 0, sizeof(*info));
    std::strncpy(info->name, p.name, sizeof(info->name) - 1);
    std::strncpy(info->arch, p.gcnArchName, sizeof(info->arch) - 1);
    info->compute_units = p.multiProcessorCount;
    info->wave_size = p.warpSize;
    info->max_threads_per_block = p.maxThreadsPerBlock;
    info->xcds = (std::strncmp(p.gcnArchName, "gfx94", 5) == 0 ||
                  std::strncmp(p.gcnArchName, "gfx95", 5) == 0)
                     ? 8
                     : 1;
    info->lds_per_block = int64_t(p.sharedMemPerBlock);
    info->lds_per_cu = int64_t(p.maxSharedMemoryPerMultiProcessor);
    info->total_mem = int64_t(p.totalGlobalMem);
    info->l2_bytes = int64_t(p.l2CacheSize);
    info->clock_khz = p.clockRate;
    info->mem_clock_khz = p.memoryClockRate;
    info->mem_bus_width = p.memoryBusWidth;
    return 0;
}

int sb_mem_info(size_t *free_bytes, size_t *total_bytes) {
    SB_REQUIRE(free_bytes && total_bytes, "NULL output pointer");
    if (!context()) return 1;
    SB_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return 0;
}

int sb_malloc(void **ptr, size_t bytes) {
    SB_REQUIRE(ptr, "NULL output pointer");
    if (!context()) return 1;
    *ptr = nullptr;
    SB_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}

int sb_free(void *ptr) {
    if (!ptr) return 0;
    SB_HIP(hipFree(ptr));
    return 0;
}

int sb_host_malloc(void **ptr, size_t bytes) {
    SB_REQUIRE(ptr, "NULL output pointer");
    if (!context()) return 1;
    *ptr = nullptr;
    SB_HIP(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return 0;
}

int sb_host_free(void *ptr) {
    if (!ptr) return 0;
    SB_HIP(hipHostFree(ptr));
    return 0;
}

static int copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    Context *ctx = context();
    if (!ctx) return 1;
    if (!bytes) return 0;
    SB_REQUIRE(dst && src, "NULL pointer in copy");
    SB_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    SB_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int sb_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    return copy_sync(dst, src, bytes, hipMemcpyHostToDevice);
}

int sb_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    return copy_sync(dst, src, bytes, hipMemcpyDeviceToHost);
}

int sb_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    return copy_sync(dst, src, bytes, hipMemcpyDeviceToDevice);
}

int sb_memset(void *dst, int value, size_t bytes) {
    Context *ctx = context();
    if (!ctx) return 1;
    if (!bytes) return 0;
    SB_REQUIRE(dst, "NULL pointer in memset");
    SB_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
    SB_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int sb_synchronize(void) {
    Context *ctx = context();
    if (!ctx) return 1;
    SB_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int sb_event_create(void **event) {
    SB_REQUIRE(event, "NULL event pointer");
    Context *ctx = context();
    if (!ctx) return 1;
    hipEvent_t e = nullptr;
    SB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    *event = e;
    return 0;
}

int sb_event_record(void *event) {
    SB_REQUIRE(event, "NULL event");
    Context *ctx = context();
    if (!ctx) return 1;
    SB_HIP(hipEventRecord(static_cast<hipEvent_t>(event), ctx->stream));
    return 0;
}

int sb_event_synchronize(void *event) {
    SB_REQUIRE(event, "NULL event");
    SB_HIP(hipEventSynchronize(static_cast<hipEvent_t>(event)));
    return 0;
}

int sb_event_destroy(void *event) {
    if (event) SB_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return 0;
}

int sb_flush_caches(void) {
    Context *ctx = context();
    if (!ctx) return 1;
    if (flush_caches(ctx)) return 1;
    SB_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"
