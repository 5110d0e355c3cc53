// Internal helpers shared by the libsbmi355x.so translation units: error
// state, per-device context, launch timing and small device utilities.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/sbmi355x.h"

namespace sb {

// ---- error state --------------------------------------------------------
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

#define SB_HIP(call)                                                          \
    do {                                                                      \
        hipError_t sb_status_ = (call);                                       \
        if (sb_status_ != hipSuccess) {                                       \
            ::sb::set_error("%s failed: %s (%s:%d)", #call,                   \
                            hipGetErrorString(sb_status_), __FILE__,          \
                            __LINE__);                                        \
            (void)hipGetLastError();                                          \
            return 1;                                                         \
        }                                                                     \
    } while (0)

#define SB_REQUIRE(cond, ...)                                                 \
    do {                                                                      \
        if (!(cond)) {                                                        \
            ::sb::set_error(__VA_ARGS__);                                     \
            return 2;                                                         \
        }                                                                     \
    } while (0)

// ---- per-device context -------------------------------------------------
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    void *flush_buffer = nullptr;
    size_t flush_bytes = 0;
    int compute_units = 0;
    int xcds = 1;
    size_t lds_per_block = 0;
    // device address of the no-allocation entry point's scratch array on this
    // device (nullptr if the lookup failed, with the reason kept)
    void *noalloc_scratch = nullptr;
    size_t noalloc_scratch_bytes = 0;
    char noalloc_setup_error[160] = {};
};

// Look up this device's copy of the no-allocation scratch array (vadv.hip);
// called once per device at set-up, with the device current.
void vadv_noalloc_setup(Context &c);

// Context of the current device; initialises it on first use.  Returns
// nullptr (with the error set) on failure.
Context *context();

// Evict L2 and the Infinity Cache by streaming a large scratch buffer.
int flush_caches(Context *ctx);

// Events that the kernel launches of this thread record themselves (the
// SB_TIMER_HIP_EXT timer): the first launch records `start` as it begins,
// every launch records `stop` as it ends, so the pair brackets the kernels
// with no stream operation in between.
struct ExtEvents {
    hipEvent_t start, stop;
    int launches;
};
extern thread_local ExtEvents *t_ext_events;

// Enqueue `kernel` on `stream`; under the hip-ext timer through
// hipExtLaunchKernelGGL with the thread's events.
template <typename F, typename... Args>
inline void enqueue_kernel(F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t stream,
                          Args... args) {
    if (ExtEvents *ev = t_ext_events) {
        hipExtLaunchKernelGGL(kernel, grid, block, uint32_t(lds), stream,
                              ev->launches ? nullptr : ev->start, ev->stop, 0u, args...);
        ++ev->launches;
    } else {
        hipLaunchKernelGGL(kernel, grid, block, lds, stream, args...);
    }
}

// Run `launch` (a callable returning hipError_t after enqueueing the
// kernels of one launch on ctx->stream, through enqueue_kernel) under the
// timing contract of the C ABI: l->timer SB_TIMER_EVENTS records events on
// the stream around the launch, SB_TIMER_HIP_EXT lets the kernels record
// them (reference timers 'gpu' and 'hip-ext').
template <typename Launch>
int timed_launch(const sb_launch *l, double *seconds, Launch &&launch) {
    Context *ctx = context();
    if (!ctx) return 1;
    const int dry = l ? l->dry_runs : 0;
    if (!seconds) {
        SB_HIP(launch());
        return 0;
    }
    const int timer = l ? l->timer : SB_TIMER_EVENTS;
    SB_REQUIRE(timer == SB_TIMER_EVENTS || timer == SB_TIMER_HIP_EXT, "unknown timer %d",
               timer);
    for (int d = 0; d < dry; ++d) SB_HIP(launch());
    if (l && l->flush) {
        if (flush_caches(ctx)) return 1;
    }
    if (timer == SB_TIMER_HIP_EXT) {
        ExtEvents ev{ctx->start, ctx->stop, 0};
        t_ext_events = &ev;
        const hipError_t e = launch();
        t_ext_events = nullptr;
        SB_HIP(e);
        SB_REQUIRE(ev.launches > 0, "this entry point does not support the hip-ext timer");
    } else {
        SB_HIP(hipEventRecord(ctx->start, ctx->stream));
        SB_HIP(launch());
        SB_HIP(hipEventRecord(ctx->stop, ctx->stream));
    }
    SB_HIP(hipEventSynchronize(ctx->stop));
    float ms = 0.f;
    SB_HIP(hipEventElapsedTime(&ms, ctx->start, ctx->stop));
    *seconds = double(ms) * 1e-3;
    return 0;
}

// Validate pointer / geometry preconditions common to all stencils.
int check_geom(const sb_geom *g, int64_t need_hi, int64_t need_hj,
               int64_t need_hk);

// Largest vector width (elements, power of two, at most 16 bytes) such that
// every pointer in `ptrs` and the strides sj, sk are multiples of it; the
// unit stride si must be 1, else 1 is returned.
int vector_width(size_t elem_bytes, const void *const *ptrs, int nptrs,
                 const sb_geom *g);

// Map a linear workgroup index so that consecutive "logical" tiles land on
// the same XCD: hardware dispatches workgroup b to XCD b % xcds.
__host__ __device__ inline uint32_t xcd_remap(uint32_t b, uint32_t total, uint32_t xcds) {
    if (xcds <= 1 || total < xcds) return b;
    const uint32_t per = total / xcds;      // full rounds
    const uint32_t full = per * xcds;
    if (b >= full) return b;                // ragged tail keeps its order
    return (b % xcds) * per + b / xcds;
}

// Division of a dividend below 2^31 by a divisor fixed at launch: q =
// (n * mul) >> shift with mul = ceil(2^shift / d), shift = 31 + ceil(log2 d).
// The rounding error of mul stays below 1/d of the quotient's unit for every
// such n, so q = floor(n / d) exactly.  Wave-uniform dividends then take a
// few scalar instructions instead of the VALU reciprocal sequence of a
// division by a run-time divisor.
struct UDiv {
    uint64_t mul;
    uint32_t shift, d;
};

inline UDiv make_udiv(uint32_t d) {
    uint32_t l = 0;
    while ((uint64_t(1) << l) < d) ++l;
    const uint32_t shift = 31 + l;
    return UDiv{((uint64_t(1) << shift) + d - 1) / d, shift, d};
}

__host__ __device__ inline uint32_t udiv(uint32_t n, const UDiv &u) {
    return uint32_t((uint64_t(n) * u.mul) >> u.shift);
}

// The grid of a kernel with one workgroup per (tile_i, tile_j, plane), its
// XCD remap (the one of xcd_remap, for a power-of-two XCD count) and tile
// decomposition precomputed on the host, so a workgroup finds its tile with
// scalar instructions.
struct TileGrid {
    uint32_t full;              // workgroups below it are remapped (0: none)
    uint32_t per;               // full rounds: blocks / xcds
    uint32_t xcd_mask, xcd_shift;
    UDiv tiles, tiles_i;        // by tiles_i * tiles_j and by tiles_i
};

inline TileGrid make_tile_grid(uint32_t blocks, uint32_t tiles_i, uint32_t tiles_j, int xcds) {
    TileGrid g{};
    const bool pow2 = xcds > 1 && (xcds & (xcds - 1)) == 0;
    if (pow2 && blocks >= uint32_t(xcds)) {
        g.per = blocks / uint32_t(xcds);
        g.full = g.per * uint32_t(xcds);
        g.xcd_mask = uint32_t(xcds) - 1;
        while ((1u << g.xcd_shift) < uint32_t(xcds)) ++g.xcd_shift;
    }
    g.tiles = make_udiv(tiles_i * tiles_j);
    g.tiles_i = make_udiv(tiles_i);
    return g;
}

// Tile (ti, tj) and plane k of workgroup b of such a grid.
__host__ __device__ inline void tile_of(uint32_t b, const TileGrid &g, uint32_t &ti, uint32_t &tj,
                               uint32_t &k) {
    if (b < g.full) b = (b & g.xcd_mask) * g.per + (b >> g.xcd_shift);
    k = udiv(b, g.tiles);
    const uint32_t t = b - k * g.tiles.d;
    tj = udiv(t, g.tiles_i);
    ti = t - tj * g.tiles_i.d;
}

template <int P>
struct Prec;
template <>
struct Prec<SB_F64> {
    using T = double;
    using C = double;
};
template <>
struct Prec<SB_F32> {
    using T = float;
    using C = float;
};
template <>
struct Prec<SB_F32_F64> {
    using T = float;
    using C = double;
};

inline size_t prec_bytes(int prec) { return prec == SB_F64 ? 8 : 4; }

}  // namespace sb

// Dispatch a templated functor on the runtime precision code.
#define SB_DISPATCH_PREC(prec, FUNCTOR, ...)                                  \
    ((prec) == SB_F64       ? FUNCTOR<SB_F64>(__VA_ARGS__)                    \
     : (prec) == SB_F32     ? FUNCTOR<SB_F32>(__VA_ARGS__)                    \
     : (prec) == SB_F32_F64 ? FUNCTOR<SB_F32_F64>(__VA_ARGS__)                \
                            : (::sb::set_error("invalid precision code %d",  \
                                               (int)(prec)),                  \
                               2))
