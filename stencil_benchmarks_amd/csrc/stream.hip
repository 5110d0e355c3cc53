// STREAM copy / scale / add / triad on HBM-resident arrays.
//
// Each thread moves U vectors of V elements per pass (V elements of 4 or 8
// bytes: one global_load/store_dword{,x2,x4} per vector); the U vectors of a
// thread are one block width apart so every load instruction of a wave covers
// 64 V element-widths contiguously.  Passes stride over the arrays by the
// whole grid, so the same kernel runs a grid that covers the arrays once or
// a persistent grid of a few workgroups per CU.  Optional non-temporal hints
// keep the streamed lines from displacing L2.  The thread and block index
// come from one grid axis (x, y or z), as the reference's `axis` option
// chooses.
#include <type_traits>

#include "sb_internal.h"

namespace sb {
namespace {

template <typename T, int V>
using vec_t = std::conditional_t<V == 1, T, T __attribute__((ext_vector_type(V)))>;

template <bool NT, typename VT>
__device__ __forceinline__ VT load(const VT *p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <bool NT, typename VT>
__device__ __forceinline__ void store(VT *p, VT v) {
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

struct GridAxis {
    uint32_t thread, block_size;
    size_t block, blocks;
};

__device__ __forceinline__ GridAxis grid_axis(int axis) {
    if (axis == 1) return {threadIdx.y, blockDim.y, blockIdx.y, gridDim.y};
    if (axis == 2) return {threadIdx.z, blockDim.z, blockIdx.z, gridDim.z};
    return {threadIdx.x, blockDim.x, blockIdx.x, gridDim.x};
}

// c = a; b = s c; c = a + b; a = b + s c  (stream/cuda_hip.j2:147-173)
template <typename T, int OP, int V, int U, bool NTL, bool NTS>
__global__ __launch_bounds__(1024) void stream_kernel(T *__restrict__ a, T *__restrict__ b,
                                                      T *__restrict__ c, T s, size_t nvec,
                                                      int axis) {
    using VT = vec_t<T, V>;
    VT *A = reinterpret_cast<VT *>(a);
    VT *B = reinterpret_cast<VT *>(b);
    VT *Cv = reinterpret_cast<VT *>(c);
    const GridAxis g = grid_axis(axis);
    const size_t pass = size_t(U) * g.block_size;
    for (size_t base = g.block * pass + g.thread; base < nvec; base += g.blocks * pass) {
        VT x[U], y[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t idx = base + size_t(u) * g.block_size;
            if (idx < nvec) {
                if constexpr (OP == SB_STREAM_COPY) {
                    x[u] = load<NTL>(A + idx);
                } else if constexpr (OP == SB_STREAM_SCALE) {
                    x[u] = load<NTL>(Cv + idx);
                } else if constexpr (OP == SB_STREAM_ADD) {
                    x[u] = load<NTL>(A + idx);
                    y[u] = load<NTL>(B + idx);
                } else {
                    x[u] = load<NTL>(B + idx);
                    y[u] = load<NTL>(Cv + idx);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t idx = base + size_t(u) * g.block_size;
            if (idx < nvec) {
                if constexpr (OP == SB_STREAM_COPY) {
                    store<NTS>(Cv + idx, x[u]);
                } else if constexpr (OP == SB_STREAM_SCALE) {
                    store<NTS>(B + idx, s * x[u]);
                } else if constexpr (OP == SB_STREAM_ADD) {
                    store<NTS>(Cv + idx, x[u] + y[u]);
                } else {
                    store<NTS>(A + idx, x[u] + s * y[u]);
                }
            }
        }
    }
}

// Elements past the last full vector.
template <typename T, int OP>
__global__ void stream_tail(T *a, T *b, T *c, T s, size_t first, size_t n) {
    const size_t i = first + threadIdx.x;
    if (i >= n) return;
    if constexpr (OP == SB_STREAM_COPY)
        c[i] = a[i];
    else if constexpr (OP == SB_STREAM_SCALE)
        b[i] = s * c[i];
    else if constexpr (OP == SB_STREAM_ADD)
        c[i] = a[i] + b[i];
    else
        a[i] = b[i] + s * c[i];
}

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(T *x, size_t n, T value) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x)
        x[i] = value;
}

// splitmix64 finaliser of (seed, index): uniform [0,1) with 53 (double) or
// 24 (float) random mantissa bits.
__device__ __host__ inline uint64_t mix64(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
__global__ __launch_bounds__(256) void fill_random_kernel(T *x, size_t n, uint64_t seed) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        const uint64_t z = mix64(seed, i);
        if constexpr (sizeof(T) == 8)
            x[i] = T(double(z >> 11) * 0x1.0p-53);
        else
            x[i] = T(float(uint32_t(z >> 40)) * 0x1.0p-24f);
    }
}

// Several arrays of n elements in one launch: blockIdx.y is the array.
constexpr int kMaxFillFields = 16;
template <typename T>
struct FillFields {
    T *p[kMaxFillFields];
    uint64_t seed[kMaxFillFields];
};

template <typename T>
__global__ __launch_bounds__(256) void fill_fields_kernel(FillFields<T> f, size_t n) {
    T *x = f.p[blockIdx.y];
    const uint64_t seed = f.seed[blockIdx.y];
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        const uint64_t z = mix64(seed, i);
        if constexpr (sizeof(T) == 8)
            x[i] = T(double(z >> 11) * 0x1.0p-53);
        else
            x[i] = T(float(uint32_t(z >> 40)) * 0x1.0p-24f);
    }
}

int grid_for(size_t n, Context *ctx);

template <typename T>
int fill_fields(void *const *fields, int count, size_t n, const uint64_t *seeds, Context *ctx) {
    FillFields<T> f{};
    for (int q = 0; q < count; ++q) {
        SB_REQUIRE(fields[q], "NULL array %d", q);
        f.p[q] = static_cast<T *>(fields[q]);
        f.seed[q] = seeds[q];
    }
    const unsigned gx = unsigned(std::max(1, grid_for(n, ctx) / std::max(1, count)));
    hipLaunchKernelGGL(fill_fields_kernel<T>, dim3(gx, unsigned(count)), dim3(256), 0, ctx->stream,
                       f, n);
    SB_HIP(hipGetLastError());
    return 0;
}

// Max relative deviation from `expected`, as the bit pattern of a
// non-negative double (monotonic under unsigned comparison).
template <typename T>
__global__ __launch_bounds__(256) void check_kernel(const T *x, size_t n, double expected,
                                                    unsigned long long *result) {
    double worst = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        const double rel = fabs((double(x[i]) - expected) / expected);
        worst = rel > worst ? rel : worst;
    }
    for (int off = 32; off > 0; off /= 2) {
        const double other = __shfl_down(worst, off);
        worst = other > worst ? other : worst;
    }
    if ((threadIdx.x & 63) == 0)
        atomicMax(result, (unsigned long long)__double_as_longlong(worst));
}

// Largest grid extent along y and z (HIP limit); larger x extents are fine.
constexpr size_t kMaxGridYZ = 65535;

template <typename T, int OP, int V, int U>
int launch_stream(T *a, T *b, T *c, size_t n, double scalar, const sb_launch *l,
                  double *seconds, int block, int per_cu, int axis) {
    const size_t nvec = n / V;
    const size_t pass = size_t(U) * block;
    const int nt = l ? (l->variant & 3) : 0;
    Context *ctx = context();
    if (!ctx) return 1;
    size_t blocks = (nvec + pass - 1) / pass;
    if (per_cu > 0) blocks = std::min(blocks, size_t(per_cu) * size_t(ctx->compute_units));
    if (axis != 0) blocks = std::min(blocks, kMaxGridYZ);
    const unsigned gb = unsigned(blocks), tb = unsigned(block);
    const dim3 grid = axis == 0 ? dim3(gb) : axis == 1 ? dim3(1, gb) : dim3(1, 1, gb);
    const dim3 threads = axis == 0 ? dim3(tb) : axis == 1 ? dim3(1, tb) : dim3(1, 1, tb);
    const T s = T(scalar);
    return timed_launch(l, seconds, [&]() -> hipError_t {
        if (blocks) {
            switch (nt) {
                case 0:
                    enqueue_kernel((stream_kernel<T, OP, V, U, false, false>), grid, threads, 0,
                                   ctx->stream, a, b, c, s, nvec, axis);
                    break;
                case 1:
                    enqueue_kernel((stream_kernel<T, OP, V, U, true, false>), grid, threads, 0,
                                   ctx->stream, a, b, c, s, nvec, axis);
                    break;
                case 2:
                    enqueue_kernel((stream_kernel<T, OP, V, U, false, true>), grid, threads, 0,
                                   ctx->stream, a, b, c, s, nvec, axis);
                    break;
                default:
                    enqueue_kernel((stream_kernel<T, OP, V, U, true, true>), grid, threads, 0,
                                   ctx->stream, a, b, c, s, nvec, axis);
                    break;
            }
        }
        if (nvec * V < n)
            enqueue_kernel((stream_tail<T, OP>), dim3(1), dim3(V), 0, ctx->stream, a, b, c, s,
                           nvec * V, n);
        return hipGetLastError();
    });
}

template <typename T, int OP, int V>
int dispatch_unroll(T *a, T *b, T *c, size_t n, double scalar, const sb_launch *l,
                    double *seconds, int block, int per_cu, int axis) {
    const int unroll = (l && l->unroll) ? l->unroll : 4;
    switch (unroll) {
        case 1: return launch_stream<T, OP, V, 1>(a, b, c, n, scalar, l, seconds, block, per_cu, axis);
        case 2: return launch_stream<T, OP, V, 2>(a, b, c, n, scalar, l, seconds, block, per_cu, axis);
        case 4: return launch_stream<T, OP, V, 4>(a, b, c, n, scalar, l, seconds, block, per_cu, axis);
        case 8: return launch_stream<T, OP, V, 8>(a, b, c, n, scalar, l, seconds, block, per_cu, axis);
        default: set_error("STREAM unroll %d not in {1, 2, 4, 8}", unroll); return 2;
    }
}

// launch->by: elements per vector (0: 16 bytes), ->bz: workgroups per CU of
// a persistent grid (0: one pass over the arrays), ->variant bits 2-3: the
// grid axis (0 x, 1 y, 2 z).
template <typename T, int OP>
int dispatch_vector(T *a, T *b, T *c, size_t n, double scalar, const sb_launch *l,
                    double *seconds) {
    const int block = (l && l->bx) ? l->bx : 256;
    SB_REQUIRE(block >= 64 && block <= 1024 && block % 64 == 0,
               "STREAM block size %d must be a multiple of 64 in [64, 1024]", block);
    constexpr int kWide = int(16 / sizeof(T));
    const int vec = (l && l->by) ? l->by : kWide;
    const int per_cu = l ? l->bz : 0;
    SB_REQUIRE(per_cu >= 0 && per_cu <= 64, "STREAM workgroups per CU %d not in 0..64", per_cu);
    const int axis = l ? (l->variant >> 2) & 3 : 0;
    SB_REQUIRE(axis <= 2, "STREAM grid axis %d not in 0..2", axis);
    for (T *p : {a, b, c})
        SB_REQUIRE(reinterpret_cast<uintptr_t>(p) % (size_t(vec) * sizeof(T)) == 0,
                   "STREAM arrays must be aligned to the vector size");
    switch (vec) {
        case 1: return dispatch_unroll<T, OP, 1>(a, b, c, n, scalar, l, seconds, block, per_cu, axis);
        case 2: return dispatch_unroll<T, OP, 2>(a, b, c, n, scalar, l, seconds, block, per_cu, axis);
        case 4:
            if constexpr (sizeof(T) == 4)
                return dispatch_unroll<T, OP, 4>(a, b, c, n, scalar, l, seconds, block, per_cu, axis);
            [[fallthrough]];
        default:
            set_error("STREAM vector size %d not in {1, %d} elements of %zu bytes", vec, kWide,
                      sizeof(T));
            return 2;
    }
}

template <typename T>
int dispatch_op(int op, T *a, T *b, T *c, size_t n, double scalar, const sb_launch *l,
                double *seconds) {
    switch (op) {
        case SB_STREAM_COPY: return dispatch_vector<T, SB_STREAM_COPY>(a, b, c, n, scalar, l, seconds);
        case SB_STREAM_SCALE: return dispatch_vector<T, SB_STREAM_SCALE>(a, b, c, n, scalar, l, seconds);
        case SB_STREAM_ADD: return dispatch_vector<T, SB_STREAM_ADD>(a, b, c, n, scalar, l, seconds);
        case SB_STREAM_TRIAD: return dispatch_vector<T, SB_STREAM_TRIAD>(a, b, c, n, scalar, l, seconds);
        default: set_error("invalid STREAM kernel %d", op); return 2;
    }
}

int grid_for(size_t n, Context *ctx) {
    const size_t want = (n + 255) / 256;
    const size_t cap = size_t(ctx->compute_units) * 32;
    return int(want < cap ? (want ? want : 1) : cap);
}

}  // namespace
}  // namespace sb

using namespace sb;

extern "C" {

int sb_stream_fill(int prec, void *x, size_t n, double value) {
    SB_REQUIRE(prec == SB_F64 || prec == SB_F32, "STREAM precision must be SB_F64 or SB_F32");
    SB_REQUIRE(x || !n, "NULL array");
    Context *ctx = context();
    if (!ctx || !n) return ctx ? 0 : 1;
    if (prec == SB_F64)
        hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(n, ctx)), dim3(256), 0, ctx->stream,
                           static_cast<double *>(x), n, value);
    else
        hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n, ctx)), dim3(256), 0, ctx->stream,
                           static_cast<float *>(x), n, float(value));
    SB_HIP(hipGetLastError());
    SB_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int sb_stream_fill_random(int prec, void *x, size_t n, uint64_t seed) {
    SB_REQUIRE(prec == SB_F64 || prec == SB_F32, "STREAM precision must be SB_F64 or SB_F32");
    SB_REQUIRE(x || !n, "NULL array");
    Context *ctx = context();
    if (!ctx || !n) return ctx ? 0 : 1;
    if (prec == SB_F64)
        hipLaunchKernelGGL(fill_random_kernel<double>, dim3(grid_for(n, ctx)), dim3(256), 0,
                           ctx->stream, static_cast<double *>(x), n, seed);
    else
        hipLaunchKernelGGL(fill_random_kernel<float>, dim3(grid_for(n, ctx)), dim3(256), 0,
                           ctx->stream, static_cast<float *>(x), n, seed);
    SB_HIP(hipGetLastError());
    SB_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int sb_fill_random_fields(int prec, void *const *fields, int count, size_t n,
                          const uint64_t *seeds) {
    SB_REQUIRE(prec == SB_F64 || prec == SB_F32, "fill precision must be SB_F64 or SB_F32");
    SB_REQUIRE(count >= 0 && count <= kMaxFillFields, "field count %d not in 0..%d", count,
               kMaxFillFields);
    SB_REQUIRE((fields && seeds) || !count, "NULL field or seed array");
    Context *ctx = context();
    if (!ctx) return 1;
    if (!n || !count) return 0;
    return prec == SB_F64 ? fill_fields<double>(fields, count, n, seeds, ctx)
                          : fill_fields<float>(fields, count, n, seeds, ctx);
}

int sb_stream_kernel(int prec, int op, void *a, void *b, void *c, size_t n, double scalar,
                     const sb_launch *l, double *seconds) {
    SB_REQUIRE(a && b && c, "NULL STREAM array");
    if (prec == SB_F64)
        return dispatch_op<double>(op, static_cast<double *>(a), static_cast<double *>(b),
                                   static_cast<double *>(c), n, scalar, l, seconds);
    if (prec == SB_F32)
        return dispatch_op<float>(op, static_cast<float *>(a), static_cast<float *>(b),
                                  static_cast<float *>(c), n, scalar, l, seconds);
    set_error("STREAM precision must be SB_F64 or SB_F32");
    return 2;
}

int sb_stream_check(int prec, const void *x, size_t n, double expected, double *max_rel_err) {
    SB_REQUIRE(prec == SB_F64 || prec == SB_F32, "STREAM precision must be SB_F64 or SB_F32");
    SB_REQUIRE(max_rel_err, "NULL result pointer");
    SB_REQUIRE(expected != 0.0, "expected value must be non-zero");
    Context *ctx = context();
    if (!ctx) return 1;
    unsigned long long *d_res = nullptr;
    SB_HIP(hipMallocAsync(reinterpret_cast<void **>(&d_res), sizeof(*d_res), ctx->stream));
    SB_HIP(hipMemsetAsync(d_res, 0, sizeof(*d_res), ctx->stream));
    if (n) {
        if (prec == SB_F64)
            hipLaunchKernelGGL(check_kernel<double>, dim3(grid_for(n, ctx)), dim3(256), 0,
                               ctx->stream, static_cast<const double *>(x), n, expected, d_res);
        else
            hipLaunchKernelGGL(check_kernel<float>, dim3(grid_for(n, ctx)), dim3(256), 0,
                               ctx->stream, static_cast<const float *>(x), n, expected, d_res);
        SB_HIP(hipGetLastError());
    }
    unsigned long long bits = 0;
    SB_HIP(hipMemcpyAsync(&bits, d_res, sizeof(bits), hipMemcpyDeviceToHost, ctx->stream));
    SB_HIP(hipFreeAsync(d_res, ctx->stream));
    SB_HIP(hipStreamSynchronize(ctx->stream));
    double v;
    static_assert(sizeof(v) == sizeof(bits), "size");
    __builtin_memcpy(&v, &bits, sizeof(v));
    *max_rel_err = v;
    return 0;
}

}  // extern "C"
