// Vertical advection: implicit tridiagonal (Thomas) solve per (i, j) column.
//
// Arithmetic follows stencils/base.py:410-473 operation by operation (the
// FP64 path is bit-identical to it when compiled with -ffp-contract=off):
//   gav = -0.25 (wcon(i+is,j+js,k) + wcon(k)),  gcv = 0.25 (wcon(i+is,j+js,k+1) + wcon(k+1))
//   as = gav bet_m,  cs = gcv bet_m,  a = gav bet_p,  c = gcv bet_p
//   b  = dtr - a - c   (k = 0: dtr - c; k = nz-1: dtr - a)
//   d  = ((dtr pos + tens) + tensstage) + corr,
//        corr = (-as)(stage(k-1) - stage(k)) - cs (stage(k+1) - stage(k))
//   k = 0:       c' = c / b,  d' = d / b
//   0 < k < n-1: r = 1 / (b - c'(k-1) a),  c' = c r,  d' = (d - d'(k-1) a) r
//   k = n-1:     d' = (d - d'(k-1) a) / (b - c'(k-1) a)
//   backward:    x(n-1) = d'(n-1),  x(k) = d'(k) - c'(k) x(k+1)
//                tensstage(k) = dtr (x(k) - pos(k))
// with dtr = 3/20 and bet_m = bet_p = 0.5.
//
// Work decomposition.  One lane owns one column; the 64 lanes of the
// compute wave own 64 consecutive i at fixed j (variants 1 and 2), or 64
// consecutive columns in i-fastest order that may run on into the next row
// (variants 3 and 4, vadv_column.h), so a level's field access is one or
// two contiguous row segments.  The compute wave keeps
// stage / wcon of the previous level in registers, loads the next PF levels
// while it computes the current ones, stores (c', d') of levels 0..nz-2 and
// runs the backward substitution from that storage in the same kernel.
//
// Storage of (c', d'): variant 1 keeps it in LDS ((nz-1) x 64 x 16 bytes
// for FP64 arithmetic, so nz <= 160 fits one workgroup); variant 2 keeps it
// in a library-owned FP64 global scratch laid out [k][column].
//
// Memory-level parallelism.  With LDS holding the coefficients, a CU runs a
// single 64-column compute wave at nz = 160 and one wave can only track 63
// outstanding loads, far less than HBM latency needs.  Variant 1 therefore
// pairs every compute wave with a helper wave in the same workgroup that
// walks LEAD levels ahead and touches each cache line of the coming levels
// (one 4-byte load per 64-byte half line, results folded into a dummy
// sum), so the compute wave's own loads hit L2.  The helper paces itself on
// the compute wave's progress counter in LDS with bounded waits (it never
// blocks the compute wave), and after the forward sweep it does the same
// for the pos lines of the backward substitution.
#include <algorithm>

#include "vadv_column.h"
#include "vadv_regs.h"
#include "vadv_ws.h"

// (c', d') scratch of the no-allocation entry point sb_vadv_noalloc: the
// global-scratch solver (variant 2) on a persistent grid sized to this array.
// Each device the library runs on holds its own copy until the process ends.
// Host code never uses this name as an address: the per-device address comes
// from hipGetSymbolAddress at set-up (vadv_noalloc_setup).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace sb {
namespace {

constexpr int PB = 16;        // backward-sweep prefetch depth (levels)
constexpr int LEAD = 32;      // helper lead over the compute wave (levels)
constexpr int MAX_SPINS = 4096;

// Progress word shared by the compute and helper waves of a workgroup:
// forward level k published as k, backward level k as 2 nz - 1 - k.
__device__ __forceinline__ void publish(int *progress, int value) {
    if (progress) __hip_atomic_store(progress, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int observe(const int *progress) {
    return __hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Wait (bounded) until the compute wave has published at least `target`.
__device__ __forceinline__ void pace(const int *progress, int target) {
    for (int spin = 0; spin < MAX_SPINS && observe(progress) < target; ++spin)
        __builtin_amdgcn_s_sleep(4);
}

template <typename T, typename C, int PF, bool NT>
__device__ __forceinline__ void solve_column(const Component &cp, const T *__restrict__ wcon,
                                             int64_t base, int64_t sk, int nz, CD<C> *scratch,
                                             int64_t sstride, bool active, int *progress) {
    const T *__restrict__ stage = static_cast<const T *>(cp.stage) + base;
    const T *__restrict__ pos = static_cast<const T *>(cp.pos) + base;
    const T *__restrict__ tens = static_cast<const T *>(cp.tens) + base;
    T *__restrict__ tstage = static_cast<T *>(cp.tensstage) + base;
    const T *__restrict__ w0 = wcon + base;
    const T *__restrict__ wn0 = wcon + base + cp.wshift;

    const C dtr = C(3.0 / 20.0);
    const C bet_m = C(0.5), bet_p = C(0.5);
    const int kmid_last = nz - 2;  // last interior level with a k+1 neighbour

    // Level record of k; callers clamp k into [0, nz-2] so every load is in
    // bounds (records past the end are re-reads of level nz-2, discarded).
    auto load = [&](int k, Level<T> &r) {
        const int64_t p = int64_t(k) * sk;
        r.pos = pos[p];
        r.tens = stream_load<NT>(tens + p);
        r.ts = stream_load<NT>(tstage + p);
        r.st = stream_load<NT>(stage + p + sk);
        r.w = stream_load<NT>(w0 + p + sk);
        r.wn = stream_load<NT>(wn0 + p + sk);
    };

    // ---- k = 0 ----------------------------------------------------------
    Level<T> first;
    load(0, first);
    Level<T> cur[PF], nxt[PF];
#pragma unroll
    for (int p = 0; p < PF; ++p) load(min(1 + p, kmid_last), cur[p]);
    const C st0 = C(stage[0]);

    C st_m, st_c, w_c, wn_c, c_prev, d_prev;
    {
        const C gcv = C(0.25) * (C(first.wn) + C(first.w));
        const C cs = gcv * bet_m;
        C cc = gcv * bet_p;
        const C bcol = dtr - cc;
        const C corr = -cs * (C(first.st) - st0);
        C dd = ((dtr * C(first.pos) + C(first.tens)) + C(first.ts)) + corr;
        cc = cc / bcol;
        dd = dd / bcol;
        scratch[0] = CD<C>{cc, dd};
        c_prev = cc;
        d_prev = dd;
        st_m = st0;
        st_c = C(first.st);
        w_c = C(first.w);
        wn_c = C(first.wn);
    }

    auto middle = [&](int k, const Level<T> &r) {
        const C gav = C(-0.25) * (wn_c + w_c);
        const C w_n = C(r.w), wn_n = C(r.wn), st_n = C(r.st);
        const C gcv = C(0.25) * (wn_n + w_n);
        const C as = gav * bet_m;
        const C cs = gcv * bet_m;
        const C acol = gav * bet_p;
        C cc = gcv * bet_p;
        const C bcol = (dtr - acol) - cc;
        const C corr = -as * (st_m - st_c) - cs * (st_n - st_c);
        C dd = ((dtr * C(r.pos) + C(r.tens)) + C(r.ts)) + corr;
        const C divided = C(1.0) / (bcol - c_prev * acol);
        cc = cc * divided;
        dd = (dd - d_prev * acol) * divided;
        scratch[int64_t(k) * sstride] = CD<C>{cc, dd};
        c_prev = cc;
        d_prev = dd;
        st_m = st_c;
        st_c = st_n;
        w_c = w_n;
        wn_c = wn_n;
    };

    // ---- 0 < k < nz-1: full blocks of PF levels, next block in flight -----
    int kb = 1;
    for (; kb + PF - 1 <= kmid_last; kb += PF) {
#pragma unroll
        for (int p = 0; p < PF; ++p) load(min(kb + PF + p, kmid_last), nxt[p]);
#pragma unroll
        for (int p = 0; p < PF; ++p) middle(kb + p, cur[p]);
#pragma unroll
        for (int p = 0; p < PF; ++p) cur[p] = nxt[p];
        publish(progress, kb + PF);
    }
    // last level's own inputs, issued before the partial block
    const int kl = nz - 1;
    const T last_pos = pos[int64_t(kl) * sk], last_tens = tens[int64_t(kl) * sk],
            last_ts = tstage[int64_t(kl) * sk];
#pragma unroll
    for (int p = 0; p < PF; ++p)
        if (kb + p <= kmid_last) middle(kb + p, cur[p]);

    // ---- k = nz-1 (kept in registers, never stored) -----------------------
    C x;
    {
        const C gav = C(-0.25) * (wn_c + w_c);
        const C as = gav * bet_m;
        const C acol = gav * bet_p;
        const C bcol = dtr - acol;
        const C corr = -as * (st_m - st_c);
        C dd = ((dtr * C(last_pos) + C(last_tens)) + C(last_ts)) + corr;
        x = (dd - d_prev * acol) / (bcol - c_prev * acol);
    }
    publish(progress, nz);
    if (active) tstage[int64_t(kl) * sk] = T(dtr * (x - C(last_pos)));

    // ---- backward substitution over k = nz-2 .. 0, pos PB levels ahead ----
    T pcur[PB], pnxt[PB];
#pragma unroll
    for (int p = 0; p < PB; ++p) pcur[p] = pos[int64_t(max(kl - 1 - p, 0)) * sk];
    int kt = kl - 1;
    for (; kt - PB + 1 >= 0; kt -= PB) {
#pragma unroll
        for (int p = 0; p < PB; ++p) pnxt[p] = pos[int64_t(max(kt - PB - p, 0)) * sk];
#pragma unroll
        for (int p = 0; p < PB; ++p) {
            const int k = kt - p;
            const CD<C> cd = scratch[int64_t(k) * sstride];
            x = cd.d - cd.c * x;
            if (active) tstage[int64_t(k) * sk] = T(dtr * (x - C(pcur[p])));
        }
#pragma unroll
        for (int p = 0; p < PB; ++p) pcur[p] = pnxt[p];
        publish(progress, 2 * nz - 1 - (kt - PB + 1));
    }
#pragma unroll
    for (int p = 0; p < PB; ++p) {
        const int k = kt - p;
        if (k >= 0) {
            const CD<C> cd = scratch[int64_t(k) * sstride];
            x = cd.d - cd.c * x;
            if (active) tstage[int64_t(k) * sk] = T(dtr * (x - C(pcur[p])));
        }
    }
}

// Helper wave: touches every 64-byte half line of the 64-column row segments
// of stage, pos, tens, tensstage and wcon (one extra element for the wcon
// neighbour / ragged tails), LEAD levels ahead of the compute wave, then the
// pos segments backwards ahead of the backward substitution.  Addresses are
// clamped into [0, nx] along i (the halo is at least 1 wide) and into
// [0, nz-1] along k, so every touch is in bounds.
template <typename T>
__device__ float helper_wave(const Component &cp, const T *wcon, int64_t row_base, int64_t i0,
                             const VadvArgs &a, int *progress) {
    constexpr int EPT = 64 / int(sizeof(T));  // elements per touch
    constexpr int NT = 64 / EPT + 1;          // touches per field row segment
    constexpr int WIN = 32;                   // touch loads in flight
    constexpr int64_t FPE = int64_t(sizeof(T)) / 4;  // floats per element
    const int lane = threadIdx.x & 63;
    const int nz = int(a.nz);
    const int f = lane / NT, t = lane - (lane / NT) * NT;
    const int64_t elem = min(i0 + int64_t(t) * EPT, a.nx);
    const T *field = f == 0   ? static_cast<const T *>(cp.stage)
                     : f == 1 ? static_cast<const T *>(cp.pos)
                     : f == 2 ? static_cast<const T *>(cp.tens)
                     : f == 3 ? static_cast<const T *>(cp.tensstage)
                              : wcon;
    const float *line = reinterpret_cast<const float *>(field + row_base + elem * a.si);
    auto touch = [&](int k) { return line[int64_t(k) * a.sk * FPE]; };

    float acc = 0.f;
    float buf[WIN];
#pragma unroll
    for (int u = 0; u < WIN; ++u) buf[u] = touch(min(u, nz - 1));
    for (int kb = WIN; kb < nz; kb += WIN) {
        pace(progress, kb + WIN - LEAD);
#pragma unroll
        for (int u = 0; u < WIN; ++u) {
            acc += buf[u];
            buf[u] = touch(min(kb + u, nz - 1));
        }
    }
#pragma unroll
    for (int u = 0; u < WIN; ++u) acc += buf[u];

    // backward: pos row segments for k = nz-2 down to 0, NT lanes per level
    constexpr int PER = 64 / NT;  // levels per touch instruction
    const int64_t pelem = min(i0 + int64_t(lane % NT) * EPT, a.nx);
    const float *pline = reinterpret_cast<const float *>(static_cast<const T *>(cp.pos) + row_base +
                                                         pelem * a.si);
    const int sub = lane / NT < PER ? lane / NT : 0;
    for (int kt = nz - 2; kt >= 0; kt -= PER * WIN) {
        pace(progress, 2 * nz - 1 - kt - LEAD);
#pragma unroll
        for (int u = 0; u < WIN; ++u) {
            const int k = max(kt - u * PER - sub, 0);
            buf[u] = pline[int64_t(k) * a.sk * FPE];
        }
#pragma unroll
        for (int u = 0; u < WIN; ++u) acc += buf[u];
    }
    return acc;
}

// Variants 1 and 2: one 64-column tile per workgroup at a time; workgroups
// stride over the tiles by the grid size (one pass where the grid covers
// them all).  Variant 2 keeps (c', d') in global scratch, one slice of 64
// columns per workgroup, reused for each of its tiles.
template <typename T, typename C, bool LDS, int PF, bool NT, bool HELPER>
__global__ __launch_bounds__(HELPER ? 128 : 64) __attribute__((amdgpu_waves_per_eu(1, LDS ? 2 : 8)))
void vadv_kernel(Components comps, const T *__restrict__ wcon, CD<C> *__restrict__ gscratch,
                 VadvArgs a) {
    extern __shared__ unsigned char lds_raw[];
    __shared__ int s_progress;
    const int lane = threadIdx.x & 63;
    const int wave = HELPER ? int(threadIdx.x >> 6) : 0;
    CD<C> *scratch;
    int64_t sstride;
    if constexpr (LDS) {
        scratch = reinterpret_cast<CD<C> *>(lds_raw) + lane;
        sstride = 64;
    } else {
        scratch = gscratch + int64_t(blockIdx.x) * 64 + lane;
        sstride = a.scratch_cols;
    }
    for (int64_t tile = blockIdx.x; tile < a.tiles; tile += gridDim.x) {
        const int64_t tile_i = tile % a.tiles_i;
        const int64_t j = tile / a.tiles_i;
        const int64_t i0 = tile_i * 64;
        const int64_t i = i0 + lane;
        const bool active = i < a.nx;
        const int64_t ic = active ? i : a.nx - 1;  // idle lanes re-read a valid column
        const int64_t base = ic * a.si + j * a.sj;
        for (int c = 0; c < a.ncomp; ++c) {
            if constexpr (HELPER) {
                if (threadIdx.x == 0) s_progress = -(1 << 30);
                __syncthreads();
                if (wave == 0)
                    solve_column<T, C, PF, NT>(comps.c[c], wcon, base, a.sk, int(a.nz), scratch,
                                               sstride, active, &s_progress);
                else {
                    const float sum = helper_wave<T>(comps.c[c], wcon, j * a.sj, i0, a, &s_progress);
                    // an empty asm that reads the sum keeps the touch loads alive
                    __asm__ volatile("" : : "v"(sum));
                }
                __syncthreads();
            } else {
                solve_column<T, C, PF, NT>(comps.c[c], wcon, base, a.sk, int(a.nz), scratch,
                                           sstride, active, nullptr);
            }
        }
    }
}

template <typename T, typename C, int PF, bool NT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
void vadv_pipe_kernel(Components comps, const T *__restrict__ wcon, PipeArgs a) {
    extern __shared__ unsigned char lds_raw[];
    constexpr int AUX = NT ? 2 : 0;  // cache policy: non-temporal for streamed fields
    CD<C> *lds = reinterpret_cast<CD<C> *>(lds_raw) + (threadIdx.x & 63);
    const int64_t first = xcd_remap(blockIdx.x, gridDim.x, a.xcds);
    if (first >= a.ngroups) return;
    const int64_t count = (a.ngroups - first + gridDim.x - 1) / gridDim.x;
    for (int c = 0; c < a.ncomp; ++c) {
        const Component &cp = comps.c[c];
        const Rsrc R{make_rsrc(cp.stage, a.span_bytes), make_rsrc(cp.pos, a.span_bytes),
                     make_rsrc(cp.tens, a.span_bytes), make_rsrc(cp.tensstage, a.span_bytes),
                     make_rsrc(wcon, a.span_bytes)};
        C x = C(0);
        T last_pos = T(0);
        pipe_iteration<T, C, PF, AUX, true, false>(R, cp.wshift, a, first, 0, false, lds, x,
                                                   last_pos);
        for (int64_t n = 1; n < count; ++n)
            pipe_iteration<T, C, PF, AUX, true, true>(R, cp.wshift, a, first + n * gridDim.x,
                                                      first + (n - 1) * gridDim.x, (n & 1) != 0,
                                                      lds, x, last_pos);
        // the backward-only tail has one load per step: prefetch deeper
        pipe_iteration<T, C, 32, AUX, false, true>(R, cp.wshift, a, 0,
                                                   first + (count - 1) * gridDim.x,
                                                   (count & 1) != 0, lds, x, last_pos);
    }
}

// Variant 4: warp-specialised persistent solver (vadv_ws.h); four waves per
// workgroup, group sequence w, w + gridDim, ... for the workgroup's logical
// index w (consecutive w on one XCD); MB levels per barrier step.
template <typename T, typename C, bool NT, bool FAST, int MB>
__global__ __launch_bounds__(64 * kWsWaves) void vadv_ws_kernel(Components comps,
                                                               const T *__restrict__ wcon,
                                                               PipeArgs a) {
    extern __shared__ unsigned char lds_raw[];
    constexpr int AUX = NT ? 2 : 0;
    // adjacent groups of a row share the cache lines at their edges (and the
    // wcon neighbour column): run them on one XCD so its L2 fetches each once
    const int64_t first = xcd_remap(blockIdx.x, gridDim.x, a.xcds);
    if (first >= a.ngroups) return;
    const WsSeq q{first, int64_t(gridDim.x), (a.ngroups - first + gridDim.x - 1) / gridDim.x};
    C *cbuf = reinterpret_cast<C *>(lds_raw) + (threadIdx.x & 63);
    C *dbuf = cbuf + ws_rows(a.nz, MB) * 64;
    for (int c = 0; c < a.ncomp; ++c) {
        const Component &cp = comps.c[c];
        const WsFields f{cp.stage, cp.pos, cp.tens, cp.tensstage, wcon, a.span_bytes};
        WsSolver<T, C, AUX, FAST, MB>::run(f, cp.wshift, a, q, cbuf, dbuf);
    }
}

// Variant 5: one wave per 64-column group, (c', d') of the lowest levels in
// registers and of the rest in LDS (vadv_regs.h); consecutive groups on one
// XCD.
template <typename C>
constexpr int hybrid_nr() { return sizeof(C) == 8 ? 40 : 80; }
// levels of loads in flight: 8 levels of six loads stay below the 63
// outstanding vector-memory instructions a wave can count (20 levels measured
// 8 % slower at 256 x 256 x 80 with float32 storage)
template <typename T>
constexpr int hybrid_pf() { return 8; }

template <typename T, typename C, bool NT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
void vadv_hybrid_kernel(Components comps, const T *__restrict__ wcon, PipeArgs a) {
    extern __shared__ unsigned char lds_raw[];
    constexpr int AUX = NT ? 2 : 0;
    CD<C> *lds = reinterpret_cast<CD<C> *>(lds_raw) + (threadIdx.x & 63);
    // rounds of gridDim.x consecutive groups, consecutive groups of a round on
    // one XCD; a grid of every resident wave walks the rounds
    const int64_t slot = xcd_remap(blockIdx.x, gridDim.x, a.xcds);
    for (int64_t g = slot; g < a.ngroups; g += gridDim.x) {
        for (int c = 0; c < a.ncomp; ++c) {
            const Component &cp = comps.c[c];
            const Rsrc R{make_rsrc(cp.stage, a.span_bytes), make_rsrc(cp.pos, a.span_bytes),
                         make_rsrc(cp.tens, a.span_bytes), make_rsrc(cp.tensstage, a.span_bytes),
                         make_rsrc(wcon, a.span_bytes)};
            HybridColumn<T, C, hybrid_nr<C>(), hybrid_pf<T>(), AUX>::solve(R, cp.wshift, a, g, lds);
        }
    }
}

template <typename T, typename C>
size_t hybrid_lds_bytes(int64_t nz) {
    return size_t(hybrid_lds_rows(nz, hybrid_nr<C>(), hybrid_pf<T>())) * 64 * sizeof(CD<C>);
}

template <typename T, typename C>
hipError_t launch_hybrid(bool nt, uint64_t groups, size_t lds_bytes, const Components &comps,
                         const T *wcon, const PipeArgs &pa, Context *ctx) {
    PipeArgs args = pa;
    // one wave per SIMD (its registers), as many per CU as the LDS allows
    const uint64_t per_cu = lds_bytes ? std::min<uint64_t>(4, ctx->lds_per_block / lds_bytes) : 4;
    const uint64_t grid = std::min<uint64_t>(groups, std::max<uint64_t>(1, per_cu) * ctx->compute_units);
    args.xcds = grid >= uint64_t(ctx->xcds) ? uint32_t(ctx->xcds) : 1u;
    const void *fn = nt ? reinterpret_cast<const void *>(&vadv_hybrid_kernel<T, C, true>)
                        : reinterpret_cast<const void *>(&vadv_hybrid_kernel<T, C, false>);
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes));
    if (nt)
        enqueue_kernel((vadv_hybrid_kernel<T, C, true>), dim3(unsigned(grid)), dim3(64),
                       lds_bytes, ctx->stream, comps, wcon, args);
    else
        enqueue_kernel((vadv_hybrid_kernel<T, C, false>), dim3(unsigned(grid)), dim3(64),
                       lds_bytes, ctx->stream, comps, wcon, args);
    return hipGetLastError();
}

int ensure_scratch(Context *ctx, void **scratch, size_t *have, size_t bytes) {
    if (*have >= bytes) return 0;
    if (*scratch) {
        SB_HIP(hipStreamSynchronize(ctx->stream));
        SB_HIP(hipFree(*scratch));
        *scratch = nullptr;
        *have = 0;
    }
    SB_HIP(hipMalloc(scratch, bytes));
    *have = bytes;
    return 0;
}

void *g_scratch[64] = {};
size_t g_scratch_bytes[64] = {};

template <int PF, bool NT, bool HELPER, typename T, typename C>
hipError_t launch_kernel(bool lds, uint64_t blocks, size_t lds_bytes, const Components &comps,
                         const T *wcon, CD<C> *gscratch, const VadvArgs &a, Context *ctx) {
    if (lds) {
        (void)hipFuncSetAttribute(
            reinterpret_cast<const void *>(&vadv_kernel<T, C, true, PF, NT, HELPER>),
            hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes));
        enqueue_kernel((vadv_kernel<T, C, true, PF, NT, HELPER>), dim3(unsigned(blocks)),
                           dim3(HELPER ? 128 : 64), lds_bytes, ctx->stream, comps, wcon, gscratch, a);
    } else {
        enqueue_kernel((vadv_kernel<T, C, false, PF, NT, false>), dim3(unsigned(blocks)),
                           dim3(64), 0, ctx->stream, comps, wcon, gscratch, a);
    }
    return hipGetLastError();
}

template <int PF, typename T, typename C>
hipError_t launch_depth(int variant, int flags, uint64_t blocks, size_t lds_bytes,
                        const Components &comps, const T *wcon, CD<C> *gscratch, const VadvArgs &a,
                        Context *ctx) {
    const bool lds = variant == 1;
    const bool nt = flags & 1, helper = lds && (flags & 2);
    if (nt && helper)
        return launch_kernel<PF, true, true>(lds, blocks, lds_bytes, comps, wcon, gscratch, a, ctx);
    if (nt)
        return launch_kernel<PF, true, false>(lds, blocks, lds_bytes, comps, wcon, gscratch, a, ctx);
    if (helper)
        return launch_kernel<PF, false, true>(lds, blocks, lds_bytes, comps, wcon, gscratch, a, ctx);
    return launch_kernel<PF, false, false>(lds, blocks, lds_bytes, comps, wcon, gscratch, a, ctx);
}

// Static LDS of the variant-1 kernel (progress word), with slack for
// alignment.
constexpr size_t kStaticLds = 64;

template <typename T, typename C, int PF, bool NT>
hipError_t launch_pipe_k(unsigned grid, size_t lds_bytes, const Components &comps, const T *wcon,
                         const PipeArgs &pa, Context *ctx) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&vadv_pipe_kernel<T, C, PF, NT>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes));
    enqueue_kernel((vadv_pipe_kernel<T, C, PF, NT>), dim3(grid), dim3(64), lds_bytes,
                       ctx->stream, comps, wcon, pa);
    return hipGetLastError();
}

template <typename T, typename C>
hipError_t launch_pipe(int depth, bool nt, unsigned grid, size_t lds_bytes, const Components &comps,
                       const T *wcon, const PipeArgs &pa, Context *ctx) {
    if (depth == 4)
        return nt ? launch_pipe_k<T, C, 4, true>(grid, lds_bytes, comps, wcon, pa, ctx)
                  : launch_pipe_k<T, C, 4, false>(grid, lds_bytes, comps, wcon, pa, ctx);
    if (depth == 16)
        return nt ? launch_pipe_k<T, C, 16, true>(grid, lds_bytes, comps, wcon, pa, ctx)
                  : launch_pipe_k<T, C, 16, false>(grid, lds_bytes, comps, wcon, pa, ctx);
    return nt ? launch_pipe_k<T, C, 8, true>(grid, lds_bytes, comps, wcon, pa, ctx)
              : launch_pipe_k<T, C, 8, false>(grid, lds_bytes, comps, wcon, pa, ctx);
}

// Persistent grid for `groups` column groups and at most `resident`
// workgroups, and the XCD count for the group remap.  Consecutive groups of a
// round share the cache lines at their edges, so the remap runs them on one
// XCD, and then the grid is sized so that every workgroup takes the same
// number of groups (give or take one in total): a short last round would
// otherwise run on the few XCDs that hold its workgroups.  That costs loads
// in flight when it shrinks the grid much, which matters with few rounds:
// with fewer than four rounds and a balanced grid below 7/8 of the resident
// workgroups, every resident workgroup takes part and the groups go in
// dispatch order (measured on MI355X: 256x256x80 FP64 4% faster remapped,
// 200x200x80 4% slower).
struct PersistentGrid {
    unsigned grid, xcds;
};

PersistentGrid persistent_grid(uint64_t groups, uint64_t resident, int xcds) {
    resident = std::max<uint64_t>(resident, 1);
    const uint64_t rounds = (groups + resident - 1) / resident;
    const uint64_t balanced = (groups + rounds - 1) / std::max<uint64_t>(rounds, 1);
    if (rounds < 4 && balanced * 8 < resident * 7) return {unsigned(std::min(groups, resident)), 1u};
    return {unsigned(balanced), unsigned(xcds)};
}

template <typename T, typename C, bool NT, bool FAST, int MB>
hipError_t launch_ws_k(int per_cu_limit, uint64_t blocks, size_t lds_bytes, const Components &comps,
                       const T *wcon, const PipeArgs &pa, Context *ctx) {
    const void *fn = reinterpret_cast<const void *>(&vadv_ws_kernel<T, C, NT, FAST, MB>);
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes));
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * kWsWaves, lds_bytes);
    if (e != hipSuccess) return e;
    per_cu = std::max(per_cu, 1);
    if (per_cu_limit > 0) per_cu = std::min(per_cu, per_cu_limit);
    const PersistentGrid pg = persistent_grid(
        blocks, uint64_t(ctx->compute_units) * uint64_t(per_cu), ctx->xcds);
    PipeArgs args = pa;
    args.xcds = pg.xcds;
    enqueue_kernel((vadv_ws_kernel<T, C, NT, FAST, MB>), dim3(pg.grid), dim3(64 * kWsWaves),
                       lds_bytes, ctx->stream, comps, wcon, args);
    return hipGetLastError();
}

template <typename T, typename C, int MB>
hipError_t launch_ws_m(bool nt, bool fast, int per_cu_limit, uint64_t blocks, size_t lds_bytes,
                       const Components &comps, const T *wcon, const PipeArgs &pa, Context *ctx) {
    if (fast)
        return nt ? launch_ws_k<T, C, true, true, MB>(per_cu_limit, blocks, lds_bytes, comps, wcon, pa, ctx)
                  : launch_ws_k<T, C, false, true, MB>(per_cu_limit, blocks, lds_bytes, comps, wcon, pa, ctx);
    return nt ? launch_ws_k<T, C, true, false, MB>(per_cu_limit, blocks, lds_bytes, comps, wcon, pa, ctx)
              : launch_ws_k<T, C, false, false, MB>(per_cu_limit, blocks, lds_bytes, comps, wcon, pa, ctx);
}

template <typename T, typename C>
hipError_t launch_ws(int block, bool nt, bool fast, int per_cu_limit, uint64_t blocks,
                     size_t lds_bytes, const Components &comps, const T *wcon, const PipeArgs &pa,
                     Context *ctx) {
    if (block == 8)
        return launch_ws_m<T, C, 8>(nt, fast, per_cu_limit, blocks, lds_bytes, comps, wcon, pa, ctx);
    return launch_ws_m<T, C, kWsBlock>(nt, fast, per_cu_limit, blocks, lds_bytes, comps, wcon, pa,
                                       ctx);
}

// The checked fields and geometry of one call (sb_vadv, sb_vadv_noalloc).
template <typename T>
struct VadvCall {
    Components comps{};
    const T *wcon = nullptr;
    VadvArgs a{};
    uint64_t blocks = 0;  // 64-column tiles
    bool empty = false;   // nothing to compute
};

template <typename T>
int prepare_vadv(int all_components, void *const *f, const sb_geom *g, VadvCall<T> &call) {
    SB_REQUIRE(f, "NULL field array");
    const int ncomp = all_components ? 3 : 1;
    for (int q = 0; q < 4 * ncomp + 1; ++q) SB_REQUIRE(f[q], "NULL field pointer (argument %d)", q);
    if (g->nx == 0 || g->ny == 0 || g->nz == 0) {
        call.empty = true;
        return 0;
    }
    SB_REQUIRE(g->nz >= 2, "vertical advection needs nz >= 2 (got %lld)", (long long)g->nz);
    SB_REQUIRE(g->nz <= (1 << 24), "nz too large");
    const int64_t shifts[3] = {g->si, g->sj, 0};
    for (int q = 0; q < ncomp; ++q)
        call.comps.c[q] = Component{f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3], shifts[q]};
    call.wcon = static_cast<const T *>(f[4 * ncomp]);
    VadvArgs &a = call.a;
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.si = g->si;
    a.sj = g->sj;
    a.sk = g->sk;
    a.tiles_i = int32_t((g->nx + 63) / 64);
    a.ncomp = ncomp;
    call.blocks = uint64_t(a.tiles_i) * uint64_t(g->ny);
    a.tiles = int64_t(call.blocks);
    SB_REQUIRE(call.blocks < (uint64_t(1) << 26), "domain too large for one launch");
    return 0;
}

template <int P>
int launch_vadv(int all_components, void *const *f, const sb_geom *g, const sb_launch *l,
                double *seconds) {
    using T = typename Prec<P>::T;
    using C = typename Prec<P>::C;
    VadvCall<T> call;
    if (int status = prepare_vadv<T>(all_components, f, g, call)) return status;
    Context *ctx = context();
    if (!ctx) return 1;
    if (call.empty) {
        if (seconds) *seconds = 0.0;
        return 0;
    }
    const Components &comps = call.comps;
    const T *wcon = call.wcon;
    VadvArgs a = call.a;
    const uint64_t blocks = call.blocks;
    // column groups of the pipelined and warp-specialised solvers (rows packed)
    const int64_t groups = vadv_groups(g->nx, g->ny);

    const size_t lds_bytes = size_t(g->nz - 1) * 64 * sizeof(CD<C>);
    const bool lds_fits = lds_bytes + kStaticLds <= ctx->lds_per_block;
    int variant = l ? l->variant : 0;
    const int64_t last_elem = (g->nx - 1 + g->hi) * g->si + (g->ny - 1 + g->hj) * g->sj +
                              (g->nz - 1 + g->hk) * g->sk;
    const bool buffer_ok = g->si > 0 && g->sj > 0 && g->sk > 0 &&
                           uint64_t(last_elem + 1) * sizeof(T) < (uint64_t(1) << 32);
    const size_t ws_lds = ws_lds_bytes<C>(g->nz);
    const bool ws_fits = ws_lds <= ctx->lds_per_block;
    // default: the warp-specialised solver where every workgroup gets at
    // least two groups and the columns are tall enough: nz >= 96 for FP64
    // storage (the single-wave pipeline is then held to one group per CU by
    // LDS), nz >= 48 for float32 storage (measured on MI355X against the
    // pipeline: 512x512x160 FP64 1.18x, float32 storage 1.9x; 256x256x80
    // FP64 0.96x, float32 storage 1.20x; float32 storage 512x512x64 1.24x,
    // 512x512x56 1.21x, 320x320x56 1.19x, 384x384x48 1.09x, 512x512x48
    // 1.07x, 512x512x40 0.84x)
    const int ws_min_nz = sizeof(T) == 8 ? 96 : 48;
    const bool ws_default = ws_fits && buffer_ok &&
                            uint64_t(last_elem + 1 + g->sk) * sizeof(T) < kWsDropOffset &&
                            g->nz >= ws_min_nz && groups >= 2 * int64_t(ctx->compute_units);
    // variant 5 (one wave per group, no pipeline fill or drain) where every
    // group is in flight at once: at most one wave per SIMD (its registers)
    // and as many per CU as its LDS allows (measured on MI355X, 256x256x80:
    // float32 storage 34.6 against 39.6 us, FP64 53.5 against 54.7 us;
    // with more groups than that, as at 384x384x80, it ran 8-20 % slower)
    const size_t v5_lds = hybrid_lds_bytes<T, C>(g->nz);
    const int64_t v5_per_cu =
        v5_lds ? std::min<int64_t>(4, int64_t(ctx->lds_per_block / v5_lds)) : 4;
    const bool v5_default = buffer_ok && g->nz <= 96 && v5_per_cu > 0 &&
                            uint64_t(last_elem + 1 + g->sk) * sizeof(T) < kRegDropOffset &&
                            groups <= v5_per_cu * int64_t(ctx->compute_units);
    // beyond the LDS of variants 1, 3 and 4 (nz > 160 with FP64 arithmetic)
    // variant 5 still holds the column while its LDS part fits (nz <= 201),
    // ahead of the global scratch (512x512x160 FP64: 765 against 863 us)
    const bool v5_fits = buffer_ok && v5_lds <= ctx->lds_per_block &&
                         uint64_t(last_elem + 1 + g->sk) * sizeof(T) < kRegDropOffset;
    if (variant == 0)
        variant = v5_default ? 5
                  : ws_default ? 4
                  : lds_fits   ? (buffer_ok ? 3 : 1)
                  : v5_fits    ? 5
                               : 2;
    SB_REQUIRE(variant >= 1 && variant <= 5, "vadv variant %d not available", variant);
    if (variant == 5) {
        const size_t hlds = hybrid_lds_bytes<T, C>(g->nz);
        SB_REQUIRE(hlds <= ctx->lds_per_block,
                   "nz = %lld needs %zu bytes of LDS per column group, device allows %zu",
                   (long long)g->nz, hlds, ctx->lds_per_block);
        SB_REQUIRE(buffer_ok && uint64_t(last_elem + 1 + g->sk) * sizeof(T) < kRegDropOffset,
                   "vadv variant 5 needs positive strides and fields below 4 GiB");
        SB_REQUIRE(uint64_t(groups) < (uint64_t(1) << 31), "domain too large for one launch");
        PipeArgs pa{a.nx, a.ny, a.nz, a.si, a.sj, a.sk, a.ncomp, groups,
                    uint32_t(uint64_t(last_elem + 1) * sizeof(T)), uint32_t(ctx->xcds)};
        const bool nt = l && l->bz > 0 && (l->bz & 1);
        return timed_launch(l, seconds, [&]() -> hipError_t {
            return launch_hybrid<T, C>(nt, uint64_t(groups), hlds, comps, wcon, pa, ctx);
        });
    }
    SB_REQUIRE(variant == 2 || variant == 4 || lds_fits,
               "nz = %lld needs %zu bytes of LDS per workgroup, device allows %zu",
               (long long)g->nz, lds_bytes + kStaticLds, ctx->lds_per_block);
    // variant 4 chosen explicitly: launch->unroll = levels per barrier step
    // (4 or 8, 0 = 4); the default choice runs 4
    const bool ws_explicit = l && l->variant == 4;
    SB_REQUIRE(!ws_explicit || l->unroll == 0 || l->unroll == 4 || l->unroll == 8,
               "warp-specialised vadv: levels per step (launch->unroll) must be 4 or 8, got %d",
               l ? l->unroll : 0);
    const int ws_block = ws_explicit && l->unroll == 8 ? 8 : kWsBlock;
    const size_t ws_lds_run = ws_lds_bytes<C>(g->nz, ws_block);
    SB_REQUIRE(variant != 4 || ws_lds_run <= ctx->lds_per_block,
               "nz = %lld needs %zu bytes of LDS per workgroup, device allows %zu",
               (long long)g->nz, ws_lds_run, ctx->lds_per_block);

    CD<C> *gscratch = nullptr;
    if (variant == 2) {
        a.scratch_cols = int64_t(blocks) * 64;
        const size_t bytes = size_t(a.scratch_cols) * size_t(g->nz) * sizeof(CD<C>);
        if (ensure_scratch(ctx, &g_scratch[ctx->device & 63], &g_scratch_bytes[ctx->device & 63],
                           bytes))
            return 1;
        gscratch = static_cast<CD<C> *>(g_scratch[ctx->device & 63]);
    }
    // defaults measured on MI355X: deep prefetch for tall columns, plain
    // loads (non-temporal loads of the streamed fields measured 0-5 % slower
    // on 128²..512² x 32..80 in both storage types, 2 % faster only at
    // 256²x40 FP64)
    const int depth = (l && l->unroll) ? l->unroll : (g->nz >= 128 ? 8 : 4);
    SB_REQUIRE(depth == 4 || depth == 8 || depth == 16, "vadv prefetch depth %d not in {4, 8, 16}",
               depth);
    // launch->bz: bit 0 non-temporal streamed loads, bit 1 helper wave (variant 1),
    // bit 2 reciprocal estimate + Newton steps instead of division (variant 4,
    // float32 storage with FP64 arithmetic only); 0 (the default) and -1 are
    // plain loads
    int flags = l ? l->bz : 0;
    if (flags < 0) flags = 0;
    const bool fast_div = (flags & 4) != 0;
    SB_REQUIRE(!fast_div || (variant == 4 && P == SB_F32_F64),
               "the fast reciprocal needs variant 4 and float32 storage with FP64 arithmetic");
    if (variant == 4) {
        SB_REQUIRE(buffer_ok && uint64_t(last_elem + 1 + g->sk) * sizeof(T) < kWsDropOffset,
                   "warp-specialised vadv needs positive strides and fields below 4 GiB");
        PipeArgs pa{a.nx, a.ny, a.nz, a.si, a.sj, a.sk, a.ncomp, groups,
                    uint32_t(uint64_t(last_elem + 1) * sizeof(T)), uint32_t(ctx->xcds)};
        const int per_cu_limit = l ? l->bx : 0;
        return timed_launch(l, seconds, [&]() -> hipError_t {
            return launch_ws<T, C>(ws_block, flags & 1, fast_div, per_cu_limit, uint64_t(groups),
                                   ws_lds_run, comps, wcon, pa, ctx);
        });
    }
    if (variant == 3) {
        // byte span from a field's interior origin to its last addressable
        // element (halo included); buffer offsets are 32-bit
        const int64_t last = (g->nx - 1 + g->hi) * g->si + (g->ny - 1 + g->hj) * g->sj +
                             (g->nz - 1 + g->hk) * g->sk;
        const uint64_t span = uint64_t(last + 1) * sizeof(T);
        SB_REQUIRE(g->si > 0 && g->sj > 0 && g->sk > 0 && span < (uint64_t(1) << 32),
                   "pipelined vadv needs positive strides and fields below 4 GiB");
        PipeArgs pa{a.nx, a.ny, a.nz, a.si, a.sj, a.sk, a.ncomp, groups, uint32_t(span),
                    uint32_t(ctx->xcds)};
        // persistent grid: launch->bx workgroups per CU (0: as many as LDS allows)
        const int fit = int(size_t(ctx->lds_per_block) / (lds_bytes + kStaticLds));
        const int per_cu = (l && l->bx > 0) ? min(l->bx, max(fit, 1)) : max(fit, 1);
        const PersistentGrid pg = persistent_grid(
            uint64_t(groups), uint64_t(ctx->compute_units) * uint64_t(per_cu), ctx->xcds);
        pa.xcds = pg.xcds;
        const unsigned grid = pg.grid;
        return timed_launch(l, seconds, [&]() -> hipError_t {
            return launch_pipe<T, C>(depth, flags & 1, grid, lds_bytes, comps, wcon, pa, ctx);
        });
    }
    return timed_launch(l, seconds, [&]() -> hipError_t {
        switch (depth) {
            case 4: return launch_depth<4>(variant, flags, blocks, lds_bytes, comps, wcon, gscratch, a, ctx);
            case 16: return launch_depth<16>(variant, flags, blocks, lds_bytes, comps, wcon, gscratch, a, ctx);
            default: return launch_depth<8>(variant, flags, blocks, lds_bytes, comps, wcon, gscratch, a, ctx);
        }
    });
}

template <int P>
int launch_vadv_noalloc(int all_components, void *const *f, const sb_geom *g,
                        const sb_launch *l, double *seconds) {
    using T = typename Prec<P>::T;
    using C = typename Prec<P>::C;
    VadvCall<T> call;
    if (int status = prepare_vadv<T>(all_components, f, g, call)) return status;
    SB_REQUIRE(!l || l->flush == 0,
               "sb_vadv_noalloc: flushing the caches needs a buffer the library allocates; "
               "launch->flush must be 0");
    SB_REQUIRE(!l || l->variant == 0 || l->variant == 2,
               "sb_vadv_noalloc runs the global-scratch solver (variant 0 or 2), got %d",
               l->variant);
    Context *ctx = context();
    if (!ctx) return 1;
    SB_REQUIRE(ctx->noalloc_scratch, "no-allocation scratch unavailable on device %d: %s",
               ctx->device, ctx->noalloc_setup_error);
    if (call.empty) {
        if (seconds) *seconds = 0.0;
        return 0;
    }
    // one slice of 64 columns x nz levels of (c', d') per workgroup
    const size_t slice = size_t(64) * size_t(g->nz) * sizeof(CD<C>);
    SB_REQUIRE(slice <= ctx->noalloc_scratch_bytes,
               "sb_vadv_noalloc: nz = %lld needs %zu bytes of scratch per 64 columns, more "
               "than the %zu bytes it holds",
               (long long)g->nz, slice, ctx->noalloc_scratch_bytes);
    const uint64_t grid = std::min<uint64_t>(call.blocks, ctx->noalloc_scratch_bytes / slice);
    VadvArgs a = call.a;
    a.scratch_cols = int64_t(grid) * 64;
    CD<C> *gscratch = static_cast<CD<C> *>(ctx->noalloc_scratch);
    const int depth = (l && l->unroll) ? l->unroll : (g->nz >= 128 ? 8 : 4);
    SB_REQUIRE(depth == 4 || depth == 8 || depth == 16, "vadv prefetch depth %d not in {4, 8, 16}",
               depth);
    const int flags = (l && l->bz > 0) ? (l->bz & 1) : 0;
    const Components &comps = call.comps;
    const T *wcon = call.wcon;
    return timed_launch(l, seconds, [&]() -> hipError_t {
        switch (depth) {
            case 4: return launch_depth<4>(2, flags, grid, 0, comps, wcon, gscratch, a, ctx);
            case 16: return launch_depth<16>(2, flags, grid, 0, comps, wcon, gscratch, a, ctx);
            default: return launch_depth<8>(2, flags, grid, 0, comps, wcon, gscratch, a, ctx);
        }
    });
}

}  // namespace

void vadv_noalloc_setup(Context &c) {
    void *address = nullptr;
    const hipError_t status = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
    if (status != hipSuccess) {
        std::snprintf(c.noalloc_setup_error, sizeof(c.noalloc_setup_error),
                      "hipGetSymbolAddress failed: %s", hipGetErrorString(status));
        (void)hipGetLastError();
        return;
    }
    c.noalloc_scratch = address;
    c.noalloc_scratch_bytes = kNoAllocScratchBytes;
}

}  // namespace sb

using namespace sb;

extern "C" int sb_vadv_noalloc(int prec, int all_components, void *const *fields,
                               const sb_geom *g, const sb_launch *l, double *seconds) {
    if (check_geom(g, 1, all_components ? 1 : 0, 0)) return 2;
    return SB_DISPATCH_PREC(prec, launch_vadv_noalloc, all_components, fields, g, l, seconds);
}

extern "C" int sb_vadv(int prec, int all_components, void *const *fields, const sb_geom *g,
                       const sb_launch *l, double *seconds) {
    if (check_geom(g, 1, all_components ? 1 : 0, 0)) return 2;
    return SB_DISPATCH_PREC(prec, launch_vadv, all_components, fields, g, l, seconds);
}
