// Batched timesteps over many subdomains of uneven size.
//
// One timestep of a subdomain (the wiring of the batched and served
// workloads; the reference has no counterpart):
//   u_next     = hdiff(u, coeff)                       (interior of u_next)
//   utensstage = vadv(ustage = u_next, upos, utens, utensstage, wcon)
// and the caller swaps u and u_next (parity).  The halo of u / u_next is
// never written: it is a fixed boundary, identical in both buffers.
//
// A sequence of steps starts with one counter-reset kernel; each step is one
// persistent hdiff launch and one or two persistent vadv launches (grids =
// resident workgroups), all drawing items from device work queues.  Subdomains
// are ordered by cost, largest first, so the tail of each queue holds the
// small items.
//   hdiff items: (subdomain, 64 x TY tile, KC-plane chunk), the tile worker
//     of the plain kernel (hdiff_tile.h), drawn from device counters.  The
//     items are split into one contiguous range per XCD, each with its own
//     counter: a workgroup draws from its XCD's range, then from the
//     others' until all are done.  A ticket maps to its subdomain through
//     the per-subdomain prefix sums; the workgroup keeps the item range of
//     its last subdomain, so only a ticket outside it needs the binary
//     search, and it draws the next ticket while the current item runs.
//     Adjacent tiles then run on one XCD, whose L2 fetches their shared
//     halo rows and row-edge cache lines once; with one queue for the chip
//     they ran on any XCD (measured on MI355X, 96-subdomain batch: reads
//     1.38x / 1.63x the algorithmic bytes in FP64 / float32 storage);
//   vadv items:  (subdomain, residue r): the 64-column groups r, r + m,
//     r + 2m, ... of one subdomain, where m (the subdomain's item count) is
//     at most the number of resident workgroups, solved by the
//     warp-specialised column solver of the plain kernel (vadv_ws.h; the
//     single-wave pipeline of vadv_column.h with variant 3) with an LDS
//     buffer sized for the tallest column of its launch.  The host orders
//     the items (subdomains by decreasing item cost from a model of the
//     solver's pipeline, (groups + 1) x (levels / M + 3) barrier steps per
//     item, residues in order), deals them in a snake over the resident
//     workgroup slots and gives each XCD's queue the items of its slots in
//     dealing order (schedule_ws); workgroups draw tickets from their XCD's
//     queue, then from the others', like the hdiff items.  Consecutive
//     residues (adjacent groups of a row) then run on one XCD at about the
//     same time, so the waves of the chip touch the same few k planes at a
//     time, as in the plain kernel; runs of consecutive groups per
//     workgroup scatter the waves over many planes and ran 3x slower.
//     Subdomains whose slot buffer fits twice in a CU's LDS (nz <= 80 with
//     FP64 arithmetic) run in a launch of their own at two workgroups per
//     CU, the taller ones at one (measured on MI355X: 28 % faster for
//     float32-storage columns of nz <= 80).
// The counters (one 128-byte line each) are cleared by a small kernel at
// the start of every sequence (no memset nodes, so the sequence can be
// captured in a hipGraph and replayed).  Every workgroup exits once the
// counters pass their item counts.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <vector>

#include "hdiff_tile.h"
#include "vadv_column.h"
#include "vadv_ws.h"

namespace sb {
namespace {

constexpr int kMinGroupsPerItem = 8;  // amortises the pipeline's fill and drain
// hdiff work queues per step (one per XCD, at most this many XCDs), then
// the vadv counter: counters per step
constexpr int kHdiffQueues = 8;
// warp-specialised vadv work queues per LDS class (one per XCD), then the
// counter of the single-wave pipeline (variant 3)
constexpr int kVadvQueues = 8;
constexpr int kVadvCounters = kHdiffQueues;
constexpr int kPipeCounter = kHdiffQueues + 2 * kVadvQueues;
// Counter slots are kCounterStride words apart: each counter has a 128-byte
// cache line of its own, so the atomics of different queues (different
// XCDs) do not serialise on one line.  One 512x512x160 FP64 subdomain
// (20,480 hdiff items of 2 planes): 272 us per hdiff launch with the
// counters packed, 200 us padded; the 96-subdomain batch is unchanged.
constexpr int kCounterStride = 16;
constexpr int kCountersPerStep = (kPipeCounter + 1) * kCounterStride;
// Groups per item of the warp-specialised solver: its fill and drain cost
// nb + 3 steps per item; with the host's cost-ordered schedule, larger items
// do not unbalance the workgroups (16 measured best on the 96-subdomain
// batch, 8 and 64 slower).
constexpr int kWsGroupsPerItem = 16;

// One scheduled item of the warp-specialised vadv: subdomain rank, residue.
struct VItem {
    int32_t rank, residue;
};

// Device descriptor of one subdomain for one step parity.
struct SubDev {
    const void *inp;   // u (hdiff input)
    void *out;         // u_next (hdiff output, vadv stage)
    const void *coeff;
    const void *pos, *tens;
    void *ts;
    const void *wcon;
    int64_t nx, ny, nz, si, sj, sk;
    int32_t tiles_i, tiles_j;
    uint32_t span_bytes;
    int32_t vstride;  // vadv items of this subdomain = group stride
    // division by tiles_i * tiles_j and by tiles_i (hdiff item decode)
    UDiv hdiv_tiles, hdiv_tiles_i;
};

// A wave-uniform copy of lane 0's value.  Work-queue tickets reach the other
// lanes through this (not a shuffle or an LDS read), so the compiler knows
// the subdomain index and its descriptor are uniform: descriptor fields stay
// in SGPRs and buffer resources need no per-lane waterfall loop around every
// load and store.
__device__ __forceinline__ int64_t uniform64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return int64_t((uint64_t(hi) << 32) | lo);
}

__device__ __forceinline__ int find_sub(const int64_t *prefix, int count, int64_t item) {
    // largest s with prefix[s] <= item (prefix[0] = 0, prefix[count] = total)
    int lo = 0, hi = count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= item)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ void reset_counters(unsigned long long *counters, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counters[i] = 0ull;
}

// `queues` (1 .. kHdiffQueues) counters, kCounterStride words apart, split
// the items into contiguous ranges; workgroup b starts on range b % queues (its XCD when queues is
// the XCD count: workgroups are dispatched to the XCDs round-robin) and
// moves on to the next range once its own is exhausted.
template <typename T, typename C, int TY, int RPT, int KC, int FLAGS>
__global__ __launch_bounds__(64 * (TY / RPT)) void hdiff_batch_kernel(
    const SubDev *__restrict__ subs, const int64_t *__restrict__ prefix, int count,
    unsigned long long *__restrict__ counters, int queues) {
    using Tile = HdiffTile<T, C, TY, RPT, FLAGS>;
    __shared__ C s_in[Tile::IN_ELEMS];
    __shared__ C s_lap[Tile::LAP_ELEMS];
    // tickets, two slots used in turn: item n's ticket is read from one
    // while item n publishes item n+1's into the other
    __shared__ long long s_item[2];
    const int64_t total = prefix[count];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int home = int(blockIdx.x % uint32_t(queues));
    // subdomain of the last item and its item range: a workgroup's tickets
    // mostly stay in one subdomain, which then needs no search
    int sub = 0;
    int64_t sub_lo = 0, sub_hi = 0;
    int slot = 0;
    for (int r = 0; r < queues; ++r) {
        const int q = home + r < queues ? home + r : home + r - queues;
        const int64_t lo = total * q / queues, hi = total * (q + 1) / queues;
        unsigned long long *const counter = counters + q * kCounterStride;
        __syncthreads();  // every wave has read the last ticket of the previous range
        if (tid == 0) s_item[slot] = lo + (long long)atomicAdd(counter, 1ull);
        __syncthreads();
        int64_t item = uniform64(uint64_t(s_item[slot]));
        while (item < hi) {
            // The following item's ticket is drawn as this item starts and
            // published from inside the tile worker once the first plane's
            // loads, issued after the atomic, have landed: the atomic's round
            // trip hides behind them, and no wave waits for the item's stores
            // to read it (the result is used in every lane there, so no later
            // instruction waits on its register).  Every ticket drawn is
            // either worked or ends this range.
            unsigned long long next = 0;
            if (tid == 0) next = atomicAdd(counter, 1ull);
            if (item < sub_lo || item >= sub_hi) {
                sub = find_sub(prefix, count, item);
                sub_lo = prefix[sub];
                sub_hi = prefix[sub + 1];
            }
            const SubDev d = subs[sub];
            // a subdomain's items number below 2^30 (fields below 4 GiB),
            // decoded by the host's multiply-shift constants in scalars
            const uint32_t local = uint32_t(item - sub_lo);
            const uint32_t kc = udiv(local, d.hdiv_tiles);
            const uint32_t t = local - kc * d.hdiv_tiles.d;
            const uint32_t tj = udiv(t, d.hdiv_tiles_i);
            const int64_t i0 = int64_t(t - tj * d.hdiv_tiles_i.d) * Tile::TX;
            const int64_t j0 = int64_t(tj) * TY;
            const int64_t k_begin = int64_t(kc) * KC;
            const int64_t k_end = min(k_begin + int64_t(KC), d.nz);
            const HdiffGeom g{d.nx, d.ny, d.nz, d.si, d.sj, d.sk};
            const int following = slot ^ 1;
            // descriptor pointers have no known address space: pass them as global
            // pointers so the tile's loads are global (flat loads would also count
            // against the LDS wait counter)
            Tile::run(as_global(static_cast<const T *>(d.inp)),
                      as_global(static_cast<const T *>(d.coeff)), as_global(static_cast<T *>(d.out)),
                      g, i0, j0, k_begin, k_end, s_in, s_lap, [&]() {
                          const long long ticket = lo + (long long)next;
                          // an empty statement that reads the ticket in every
                          // lane keeps the add (and the wait for the atomic)
                          // out of the one-lane branch below
                          asm volatile("" : : "v"(ticket));
                          if (tid == 0) s_item[following] = ticket;
                      });
            __syncthreads();  // the tile's LDS is free and the ticket visible
            slot = following;
            item = uniform64(uint64_t(s_item[slot]));
        }
    }
}

template <typename T, typename C, int PF, bool NT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
void vadv_batch_kernel(const SubDev *__restrict__ subs, const int64_t *__restrict__ prefix,
                       int count, unsigned long long *__restrict__ counter) {
    extern __shared__ unsigned char lds_raw[];
    constexpr int AUX = NT ? 2 : 0;
    CD<C> *lds = reinterpret_cast<CD<C> *>(lds_raw) + (threadIdx.x & 63);
    const int64_t total = prefix[count];
    for (;;) {
        unsigned long long ticket = 0;
        if (threadIdx.x == 0) ticket = atomicAdd(counter, 1ull);
        const int64_t item = uniform64(ticket);
        if (item >= total) break;
        const int s = find_sub(prefix, count, item);
        const SubDev d = subs[s];
        const int64_t groups = vadv_groups(d.nx, d.ny);
        const int64_t m = d.vstride;
        const int64_t g0 = item - prefix[s];
        const int64_t n = (groups - g0 + m - 1) / m;  // groups g0, g0 + m, ...
        const PipeArgs a{d.nx, d.ny, d.nz, d.si, d.sj, d.sk, 1, groups, d.span_bytes, 0};
        const Rsrc R{make_rsrc(d.out, d.span_bytes), make_rsrc(d.pos, d.span_bytes),
                     make_rsrc(d.tens, d.span_bytes), make_rsrc(d.ts, d.span_bytes),
                     make_rsrc(d.wcon, d.span_bytes)};
        const int64_t wshift = d.si;  // u component: wcon neighbour at i+1
        C x = C(0);
        T last_pos = T(0);
        pipe_iteration<T, C, PF, AUX, true, false>(R, wshift, a, g0, 0, false, lds, x, last_pos);
        for (int64_t q = 1; q < n; ++q)
            pipe_iteration<T, C, PF, AUX, true, true>(R, wshift, a, g0 + q * m, g0 + (q - 1) * m,
                                                      (q & 1) != 0, lds, x, last_pos);
        pipe_iteration<T, C, 32, AUX, false, true>(R, wshift, a, 0, g0 + (n - 1) * m, (n & 1) != 0,
                                                   lds, x, last_pos);
    }
}

// Items of one LDS class in `queues` ranges off[q] .. off[q+1]-1 (one per
// XCD, in the host's order), each with its own counter (kCounterStride
// words apart): a workgroup draws from range blockIdx.x % queues (its XCD:
// workgroups are dispatched to the XCDs round-robin), then from the others
// until all are done.  The host
// puts consecutive residues of a subdomain (adjacent groups of a row, which
// share the cache lines at their edges) into one range, next to each other.
// The slot buffer fills the LDS at nz = 160, so the ticket is broadcast
// through its first word while no item is in flight.  MB levels per
// barrier step, NW waves per workgroup.
template <typename T, typename C, bool FAST, int MB, int NW>
__global__ __launch_bounds__(64 * NW) void vadv_ws_batch_kernel(
    const SubDev *__restrict__ subs, const int32_t *__restrict__ off,
    const VItem *__restrict__ list, unsigned long long *__restrict__ counters, int queues) {
    extern __shared__ unsigned char lds_raw[];
    C *lane0 = reinterpret_cast<C *>(lds_raw) + (threadIdx.x & 63);
    long long *s_ticket = reinterpret_cast<long long *>(lds_raw);
    const int home = int(blockIdx.x % uint32_t(queues));
    for (int r = 0; r < queues; ++r) {
        const int q = home + r < queues ? home + r : home + r - queues;
        const int32_t lo = off[q], hi = off[q + 1];
        // the next ticket is drawn while the current item runs; every ticket
        // drawn is either worked or ends this range
        unsigned long long next = 0;
        if (threadIdx.x == 0) next = atomicAdd(counters + q * kCounterStride, 1ull);
        for (;;) {
            __syncthreads();  // the previous item's LDS accesses are done
            if (threadIdx.x == 0) *s_ticket = (long long)lo + (long long)next;
            __syncthreads();
            const int64_t i = uniform64(uint64_t(*s_ticket));
            __syncthreads();
            if (i >= hi) break;
            if (threadIdx.x == 0) next = atomicAdd(counters + q * kCounterStride, 1ull);
            const VItem it = list[i];
            const int s = int(__builtin_amdgcn_readfirstlane(it.rank));
            const SubDev d = subs[s];
            const int64_t groups = vadv_groups(d.nx, d.ny);
            const int64_t m = d.vstride;
            const int64_t res = int64_t(__builtin_amdgcn_readfirstlane(it.residue));
            const PipeArgs a{d.nx, d.ny, d.nz, d.si, d.sj, d.sk, 1, groups, d.span_bytes,
                             uint32_t(queues)};
            const WsFields f{d.out, d.pos, d.tens, d.ts, d.wcon, d.span_bytes};
            // u component: wcon neighbour at i+1
            WsSolver<T, C, 0, FAST, MB, NW>::run(f, d.si, a, WsSeq{res, m, (groups - res + m - 1) / m},
                                                 lane0, lane0 + ws_rows(d.nz, MB) * 64);
        }
    }
}

// ---- host side ---------------------------------------------------------------

struct Plan {
    int prec = 0;
    int count = 0;
    int steps = 0;
    int max_nz = 0;
    sb_launch hl{}, vl{};
    // device tables: SubDev arrays for flip 0 / 1, prefix sums, counters
    void *dev = nullptr;
    size_t dev_bytes = 0;
    SubDev *subs[2] = {nullptr, nullptr};
    int64_t *hprefix = nullptr, *vprefix = nullptr;
    unsigned long long *counters = nullptr;
    // warp-specialised vadv: one launch per LDS class (see WsClass)
    struct WsClass {
        int max_nz = 0;     // tallest column of the class (sizes its LDS)
        int per_cu = 0;     // workgroups per CU
        int32_t *off = nullptr;  // items (decreasing cost) and their per-queue ranges
        VItem *list = nullptr;
        std::vector<int32_t> host_off;
        std::vector<VItem> host_list;
    } ws[2];
    std::vector<int8_t> host_vclass;  // LDS class of each subdomain (by rank)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // host copies
    std::vector<SubDev> host_subs[2];
    std::vector<int64_t> host_hprefix, host_vprefix;
    // pinned staging of the device tables, two buffers used in turn: each
    // upload is one asynchronous copy, and a buffer is refilled only after
    // the event recorded behind its previous copy
    void *stage[2] = {nullptr, nullptr};
    size_t stage_bytes[2] = {0, 0};
    hipEvent_t stage_done[2] = {nullptr, nullptr};
    int stage_next = 0;
};

int hdiff_ty(const sb_launch *l) { return (l && l->by) ? l->by : 16; }
int hdiff_kc(const sb_launch *l, int) {
    // 8 planes per item measured best for every precision on the 96-subdomain
    // batch (32..512 squared, nz 40..160)
    return (l && l->bz) ? l->bz : 8;
}
int hdiff_flags(const sb_launch *l) { return l && l->unroll ? (l->unroll & 3) : 2; }
// vadv launch options: variant 4 (default) warp-specialised solver, 3 the
// single-wave pipeline (unroll = prefetch depth 4 or 8, bz > 0 non-temporal
// streamed loads); for variant 4, bz bit 2 selects the fast reciprocal
// (float32 storage with FP64 arithmetic only)
int vadv_variant(const sb_launch *l) { return (l && l->variant) ? l->variant : 4; }
// groups per item of the warp-specialised solver (launch->by, 0 = default)
int vadv_ws_per_item(const sb_launch *l) { return (l && l->by > 0) ? l->by : kWsGroupsPerItem; }
// levels per barrier step of the warp-specialised solver (launch->unroll:
// 4 or 8, 0 = 4; check_lds rejects other values)
int vadv_ws_block(const sb_launch *l) { return (l && l->unroll == 8) ? 8 : kWsBlock; }
bool vadv_ws_block_ok(const sb_launch *l) {
    return !l || l->unroll == 0 || l->unroll == 4 || l->unroll == 8;
}
// waves per workgroup of the warp-specialised solver (launch->bx: 4 or 5,
// 0 = 4; five add a wave for the wcon loads)
int vadv_ws_waves(const sb_launch *l) { return (l && l->bx == 5) ? 5 : kWsWaves; }
bool vadv_ws_waves_ok(const sb_launch *l) { return !l || l->bx == 0 || l->bx == 4 || l->bx == 5; }
int vadv_pf(const sb_launch *l) { return (l && l->unroll) ? l->unroll : 8; }
int vadv_nt(const sb_launch *l) { return (l && l->bz) ? (l->bz > 0 ? 1 : 0) : 0; }
bool vadv_fast(const sb_launch *l) { return l && l->bz > 0 && (l->bz & 4); }

int validate(int prec, const sb_subdomain *subs, int count) {
    SB_REQUIRE(prec == SB_F64 || prec == SB_F32 || prec == SB_F32_F64, "invalid precision %d", prec);
    SB_REQUIRE(subs || count == 0, "NULL subdomain array");
    SB_REQUIRE(count >= 0, "negative subdomain count");
    const size_t esize = prec == SB_F64 ? 8 : 4;
    for (int i = 0; i < count; ++i) {
        const sb_subdomain &s = subs[i];
        const sb_geom &g = s.geom;
        SB_REQUIRE(s.u[0] && s.u[1] && s.coeff && s.upos && s.utens && s.utensstage && s.wcon,
                   "subdomain %d: NULL field pointer", i);
        SB_REQUIRE(g.nx > 0 && g.ny > 0 && g.nz >= 2, "subdomain %d: empty domain or nz < 2", i);
        SB_REQUIRE(g.hi >= 2 && g.hj >= 2, "subdomain %d: horizontal halo must be >= 2", i);
        SB_REQUIRE(g.si > 0 && g.sj > 0 && g.sk > 0, "subdomain %d: strides must be positive", i);
        const int64_t last = (g.nx - 1 + g.hi) * g.si + (g.ny - 1 + g.hj) * g.sj +
                             (g.nz - 1 + g.hk) * g.sk;
        SB_REQUIRE(uint64_t(last + 1 + g.sk) * esize < kWsDropOffset &&
                       hdiff_tile_fits(g.nx, g.ny, g.si, g.sj, esize),
                   "subdomain %d: fields must be below 4 GiB", i);
        SB_REQUIRE(s.parity == 0 || s.parity == 1, "subdomain %d: parity must be 0 or 1", i);
    }
    return 0;
}

template <typename C>
size_t vadv_lds_bytes(int max_nz) {
    return size_t(max_nz - 1) * 64 * sizeof(CD<C>);
}

// LDS per workgroup of the batched vadv (`block`: levels per step of variant 4)
size_t plan_vadv_lds(int prec, int max_nz, int variant, int block) {
    if (variant == 4)
        return prec == SB_F32 ? ws_lds_bytes<float>(max_nz, block)
                              : ws_lds_bytes<double>(max_nz, block);
    return prec == SB_F32 ? vadv_lds_bytes<float>(max_nz) : vadv_lds_bytes<double>(max_nz);
}

// work queues of the warp-specialised vadv: one per XCD
int vadv_queues(const Context *ctx) { return std::max(1, std::min(ctx->xcds, kVadvQueues)); }

// vadv workgroups of one launch for the tallest column: as many as LDS lets
// a CU hold (the warp-specialised kernel: at most 2, its VGPR budget)
unsigned vadv_workers(const Plan &p, const Context *ctx) {
    const int variant = vadv_variant(&p.vl);
    const size_t lds = plan_vadv_lds(p.prec, p.max_nz, variant, vadv_ws_block(&p.vl));
    int fit = std::max(1, int(ctx->lds_per_block / std::max<size_t>(lds, 1)));
    if (variant == 4) fit = std::min(fit, 2);
    return unsigned(fit * ctx->compute_units);
}

// Order the warp-specialised vadv items of one LDS class for the device
// queues.  Cost of an item of n groups of nb blocks: (n + 1) (nb + skew)
// barrier steps (the solver's pipeline with its fill and drain).  Subdomains
// are taken in order of decreasing item cost, the items of one subdomain in
// residue order, and dealt in a snake over `workers` slots (0 .. W-1, then
// W-1 .. 0, and so on), which balances the slots nearly as well as
// longest-first onto the least-loaded one; slot w belongs to queue
// w / (W / queues), and each queue takes its slots' items in dealing order.
// The first pass so hands the workgroups of an XCD consecutive residues
// (adjacent groups), and the workgroups then draw the rest dynamically.
void schedule_ws(Plan &p, int cls, int workers, int queues) {
    std::vector<int64_t> item_cost(size_t(p.count));
    const int block = vadv_ws_block(&p.vl);
    for (int r = 0; r < p.count; ++r) {
        const SubDev &d = p.host_subs[0][r];
        const int64_t groups = vadv_groups(d.nx, d.ny);
        const int64_t nb = (d.nz + block - 1) / block;
        item_cost[size_t(r)] = ((groups + d.vstride - 1) / d.vstride + 1) * (nb + kWsSkew);
    }
    std::vector<int> ranks;
    for (int r = 0; r < p.count; ++r)
        if (p.host_vclass[size_t(r)] == cls) ranks.push_back(r);
    std::stable_sort(ranks.begin(), ranks.end(),
                     [&](int a, int b) { return item_cost[size_t(a)] > item_cost[size_t(b)]; });
    const int64_t per_queue = std::max<int64_t>(1, workers / queues);
    std::vector<std::vector<VItem>> by_queue(static_cast<size_t>(queues));
    int64_t i = 0;
    for (int r : ranks) {
        for (int32_t res = 0; res < p.host_subs[0][r].vstride; ++res, ++i) {
            const int64_t pass = i / workers, o = i % workers;
            const int64_t w = (pass & 1) ? workers - 1 - o : o;
            by_queue[size_t(std::min<int64_t>(w / per_queue, queues - 1))].push_back(VItem{int32_t(r), res});
        }
    }
    Plan::WsClass &c = p.ws[cls];
    c.host_list.clear();
    c.host_off.assign(size_t(queues) + 1, 0);
    for (int q = 0; q < queues; ++q) {
        c.host_list.insert(c.host_list.end(), by_queue[size_t(q)].begin(), by_queue[size_t(q)].end());
        c.host_off[size_t(q) + 1] = int32_t(c.host_list.size());
    }
}

// Build descriptors and prefix sums (largest subdomains first).
void build_tables(int prec, const sb_subdomain *subs, int count, const sb_launch *hl, Plan &p,
                  const Context *ctx) {
    const size_t esize = prec == SB_F64 ? 8 : 4;
    const int ty = hdiff_ty(hl), kc = hdiff_kc(hl, prec);
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    auto cost = [&](int i) {
        const sb_geom &g = subs[i].geom;
        return double(g.nx) * double(g.ny) * double(g.nz);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
    p.max_nz = 2;
    for (int i = 0; i < count; ++i) p.max_nz = std::max<int>(p.max_nz, int(subs[i].geom.nz));
    const int64_t workers = vadv_workers(p, ctx);
    const bool ws = vadv_variant(&p.vl) == 4;
    const int64_t per_item = ws ? vadv_ws_per_item(&p.vl) : kMinGroupsPerItem;
    // warp-specialised solver: columns whose slot buffer fits twice in a
    // CU's LDS (class 0) run in their own launch at two workgroups per CU,
    // the taller ones (class 1) at one
    p.host_vclass.assign(size_t(count), 0);
    for (int c = 0; c < 2; ++c) {
        p.ws[c] = Plan::WsClass{};
    }
    if (ws) {
        const int block = vadv_ws_block(&p.vl);
        for (int r = 0; r < count; ++r) {
            const int nz = int(subs[order[r]].geom.nz);
            const int c = 2 * plan_vadv_lds(prec, nz, 4, block) <= ctx->lds_per_block ? 0 : 1;
            p.host_vclass[size_t(r)] = int8_t(c);
            p.ws[c].max_nz = std::max(p.ws[c].max_nz, nz);
        }
        for (int c = 0; c < 2; ++c) {
            if (!p.ws[c].max_nz) continue;
            const size_t lds = plan_vadv_lds(prec, p.ws[c].max_nz, 4, block);
            p.ws[c].per_cu = std::max(1, std::min(2, int(ctx->lds_per_block / lds)));
        }
    }
    for (int flip = 0; flip < 2; ++flip) p.host_subs[flip].resize(count);
    p.host_hprefix.assign(count + 1, 0);
    p.host_vprefix.assign(count + 1, 0);
    for (int r = 0; r < count; ++r) {
        const sb_subdomain &s = subs[order[r]];
        const sb_geom &g = s.geom;
        const int64_t last = (g.nx - 1 + g.hi) * g.si + (g.ny - 1 + g.hj) * g.sj +
                             (g.nz - 1 + g.hk) * g.sk;
        const int64_t tiles_i = (g.nx + 63) / 64;
        const int64_t groups = vadv_groups(g.nx, g.ny);
        const int64_t class_workers =
            ws ? int64_t(p.ws[p.host_vclass[size_t(r)]].per_cu) * ctx->compute_units : workers;
        const int64_t vitems = std::max<int64_t>(1, std::min(class_workers, groups / per_item));
        for (int flip = 0; flip < 2; ++flip) {
            const int cur = (s.parity + flip) & 1;
            SubDev d{};
            d.inp = s.u[cur];
            d.out = s.u[cur ^ 1];
            d.coeff = s.coeff;
            d.pos = s.upos;
            d.tens = s.utens;
            d.ts = s.utensstage;
            d.wcon = s.wcon;
            d.nx = g.nx;
            d.ny = g.ny;
            d.nz = g.nz;
            d.si = g.si;
            d.sj = g.sj;
            d.sk = g.sk;
            d.tiles_i = int32_t(tiles_i);
            d.tiles_j = int32_t((g.ny + ty - 1) / ty);
            d.hdiv_tiles = make_udiv(uint32_t(d.tiles_i) * uint32_t(d.tiles_j));
            d.hdiv_tiles_i = make_udiv(uint32_t(d.tiles_i));
            d.span_bytes = uint32_t(uint64_t(last + 1) * esize);
            d.vstride = int32_t(vitems);
            p.host_subs[flip][r] = d;
        }
        const int64_t hitems = tiles_i * ((g.ny + ty - 1) / ty) * ((g.nz + kc - 1) / kc);
        p.host_hprefix[r + 1] = p.host_hprefix[r] + hitems;
        p.host_vprefix[r + 1] = p.host_vprefix[r] + vitems;
    }
    if (ws)
        for (int c = 0; c < 2; ++c)
            if (p.ws[c].max_nz)
                schedule_ws(p, c, p.ws[c].per_cu * ctx->compute_units, vadv_queues(ctx));
}

// Enqueue one step (both kernels) on ctx->stream.
template <typename T, typename C>
hipError_t enqueue_step(const Plan &p, int flip, unsigned long long *counters, Context *ctx) {
    const int ty = hdiff_ty(&p.hl), kc = hdiff_kc(&p.hl, p.prec), fl = hdiff_flags(&p.hl);
    const int pf = vadv_pf(&p.vl);
    const bool nt = vadv_nt(&p.vl);
    const SubDev *subs = p.subs[flip];
    // hdiff: resident grid
    int per_cu = 0;
    auto launch_h = [&](auto kernel, int nthreads) -> hipError_t {
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(kernel), nthreads, 0);
        if (e != hipSuccess) return e;
        const unsigned grid = unsigned(std::max(1, per_cu) * ctx->compute_units);
        const int queues = std::max(1, std::min(ctx->xcds, kHdiffQueues));
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(64, nthreads / 64), 0, ctx->stream, subs,
                           p.hprefix, p.count, counters, queues);
        return hipGetLastError();
    };
#define SB_HB(TY_, RPT_, KC_, FL_)                                                          \
    if (ty == TY_ && kc == KC_ && fl == FL_)                                                \
        return_h = launch_h(hdiff_batch_kernel<T, C, TY_, RPT_, KC_, FL_>, 64 * (TY_ / RPT_));
    hipError_t return_h = hipErrorInvalidValue;
    SB_HB(16, 4, 2, 2) else SB_HB(16, 4, 2, 0) else SB_HB(16, 4, 8, 2) else SB_HB(16, 4, 8, 0)
        else SB_HB(8, 2, 1, 2) else SB_HB(32, 8, 2, 2) else SB_HB(16, 4, 4, 2) else SB_HB(16, 4, 1, 2)
#undef SB_HB
    if (return_h != hipSuccess) return return_h;
    // vadv
    const unsigned vgrid = vadv_workers(p, ctx);
    if (vadv_variant(&p.vl) == 4) {
        // one launch per LDS class, taller columns first
        for (int cls = 1; cls >= 0; --cls) {
            const Plan::WsClass &c = p.ws[cls];
            if (!c.max_nz) continue;
            const unsigned grid = unsigned(c.per_cu * ctx->compute_units);
            // the item ranges were cut for this many queues
            const int queues = vadv_queues(ctx);
            if (c.host_off.size() != size_t(queues) + 1) return hipErrorInvalidValue;
            const int block = vadv_ws_block(&p.vl), waves = vadv_ws_waves(&p.vl);
            const size_t wlds = ws_lds_bytes<C>(c.max_nz, block);
            auto launch_w = [&](auto kernel) -> hipError_t {
                (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(wlds));
                hipLaunchKernelGGL(kernel, dim3(grid), dim3(64 * waves), wlds, ctx->stream,
                                   subs, c.off, c.list,
                                   counters + (kVadvCounters + cls * kVadvQueues) * kCounterStride,
                                   queues);
                return hipGetLastError();
            };
            // (fast reciprocal, levels per step, waves) -> kernel
            auto launch_shape = [&](auto fast_tag) -> hipError_t {
                constexpr bool F = decltype(fast_tag)::value;
                if (block == 8)
                    return waves == 5 ? launch_w(vadv_ws_batch_kernel<T, C, F, 8, 5>)
                                      : launch_w(vadv_ws_batch_kernel<T, C, F, 8, kWsWaves>);
                return waves == 5 ? launch_w(vadv_ws_batch_kernel<T, C, F, kWsBlock, 5>)
                                  : launch_w(vadv_ws_batch_kernel<T, C, F, kWsBlock, kWsWaves>);
            };
            hipError_t e = hipSuccess;
            if constexpr (sizeof(T) == 4 && sizeof(C) == 8) {
                e = vadv_fast(&p.vl) ? launch_shape(std::true_type{}) : launch_shape(std::false_type{});
            } else {
                e = launch_shape(std::false_type{});
            }
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const size_t lds = vadv_lds_bytes<C>(p.max_nz);
    auto launch_v = [&](auto kernel) -> hipError_t {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        hipLaunchKernelGGL(kernel, dim3(vgrid), dim3(64), lds, ctx->stream, subs, p.vprefix,
                           p.count, counters + kPipeCounter * kCounterStride);
        return hipGetLastError();
    };
    if (pf == 4) return nt ? launch_v(vadv_batch_kernel<T, C, 4, true>) : launch_v(vadv_batch_kernel<T, C, 4, false>);
    return nt ? launch_v(vadv_batch_kernel<T, C, 8, true>) : launch_v(vadv_batch_kernel<T, C, 8, false>);
}

template <typename T, typename C>
hipError_t enqueue_sequence(const Plan &p, Context *ctx) {
    const int ncounters = kCountersPerStep * p.steps;
    hipLaunchKernelGGL(reset_counters, dim3((ncounters + 255) / 256), dim3(256), 0, ctx->stream,
                       p.counters, ncounters);
    hipError_t e = hipGetLastError();
    for (int s = 0; s < p.steps && e == hipSuccess; ++s)
        e = enqueue_step<T, C>(p, s & 1, p.counters + size_t(kCountersPerStep) * s, ctx);
    return e;
}

hipError_t enqueue_any(const Plan &p, Context *ctx) {
    if (p.prec == SB_F64) return enqueue_sequence<double, double>(p, ctx);
    if (p.prec == SB_F32) return enqueue_sequence<float, float>(p, ctx);
    return enqueue_sequence<float, double>(p, ctx);
}

int upload_tables(Plan &p, Context *ctx) {
    const size_t subs_bytes = sizeof(SubDev) * std::max(1, p.count);
    const size_t pre_bytes = sizeof(int64_t) * (p.count + 1);
    const size_t ctr_bytes = sizeof(unsigned long long) * kCountersPerStep * std::max(1, p.steps);
    const size_t sched = 2 * subs_bytes + 2 * pre_bytes + ctr_bytes;  // 8-byte aligned
    size_t ws_bytes = 0;  // both classes' schedules, each array 8-byte aligned
    for (const Plan::WsClass &c : p.ws)
        ws_bytes += ((sizeof(int32_t) * c.host_off.size() + 7) & ~size_t(7)) +
                    sizeof(VItem) * c.host_list.size();
    const size_t need = sched + ws_bytes + 256;
    if (p.dev_bytes < need) {
        // steps enqueued earlier may still read the old tables
        SB_HIP(hipStreamSynchronize(ctx->stream));
        if (p.dev) SB_HIP(hipFree(p.dev));
        p.dev = nullptr;
        SB_HIP(hipMalloc(&p.dev, need));
        p.dev_bytes = need;
    }
    // staging buffer of this upload (the device tables' image, same offsets)
    const int b = p.stage_next;
    p.stage_next ^= 1;
    if (p.stage_done[b]) {
        SB_HIP(hipEventSynchronize(p.stage_done[b]));  // its previous copy has run
    } else {
        SB_HIP(hipEventCreateWithFlags(&p.stage_done[b], hipEventDisableTiming));
    }
    if (p.stage_bytes[b] < need) {
        if (p.stage[b]) SB_HIP(hipHostFree(p.stage[b]));
        p.stage[b] = nullptr;
        p.stage_bytes[b] = 0;
        SB_HIP(hipHostMalloc(&p.stage[b], need, hipHostMallocDefault));
        p.stage_bytes[b] = need;
    }
    char *host = static_cast<char *>(p.stage[b]);
    char *base = static_cast<char *>(p.dev);
    p.subs[0] = reinterpret_cast<SubDev *>(base);
    p.subs[1] = reinterpret_cast<SubDev *>(base + subs_bytes);
    p.hprefix = reinterpret_cast<int64_t *>(base + 2 * subs_bytes);
    p.vprefix = reinterpret_cast<int64_t *>(base + 2 * subs_bytes + pre_bytes);
    p.counters = reinterpret_cast<unsigned long long *>(base + 2 * subs_bytes + 2 * pre_bytes);
    if (p.count) {
        std::memcpy(host, p.host_subs[0].data(), sizeof(SubDev) * p.count);
        std::memcpy(host + subs_bytes, p.host_subs[1].data(), sizeof(SubDev) * p.count);
    }
    std::memcpy(host + 2 * subs_bytes, p.host_hprefix.data(), pre_bytes);
    std::memcpy(host + 2 * subs_bytes + pre_bytes, p.host_vprefix.data(), pre_bytes);
    // the counters are cleared by a kernel before every step sequence
    std::memset(host + 2 * subs_bytes + 2 * pre_bytes, 0, ctr_bytes);
    size_t at = sched;
    for (Plan::WsClass &c : p.ws) {
        const size_t off_bytes = sizeof(int32_t) * c.host_off.size();
        const size_t list_bytes = sizeof(VItem) * c.host_list.size();
        c.off = reinterpret_cast<int32_t *>(base + at);
        if (off_bytes) std::memcpy(host + at, c.host_off.data(), off_bytes);
        at += (off_bytes + 7) & ~size_t(7);
        c.list = reinterpret_cast<VItem *>(base + at);
        if (list_bytes) std::memcpy(host + at, c.host_list.data(), list_bytes);
        at += list_bytes;
    }
    // one copy, ordered on the stream before the steps that read the tables
    // (and after the previous steps, which read the same device tables)
    SB_HIP(hipMemcpyAsync(p.dev, host, at, hipMemcpyHostToDevice, ctx->stream));
    SB_HIP(hipEventRecord(p.stage_done[b], ctx->stream));
    return 0;
}

int sync_stream(Context *ctx) {
    SB_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int check_lds(const Plan &p, Context *ctx) {
    const int variant = vadv_variant(&p.vl);
    SB_REQUIRE(variant == 3 || variant == 4, "batched vadv variant %d not available", variant);
    SB_REQUIRE(!vadv_fast(&p.vl) || (variant == 4 && p.prec == SB_F32_F64),
               "the fast reciprocal needs variant 4 and float32 storage with FP64 arithmetic");
    SB_REQUIRE(variant != 4 || vadv_ws_block_ok(&p.vl),
               "warp-specialised vadv: levels per step (launch->unroll) must be 4 or 8, got %d",
               p.vl.unroll);
    SB_REQUIRE(variant != 4 || vadv_ws_waves_ok(&p.vl),
               "warp-specialised vadv: waves per workgroup (launch->bx) must be 4 or 5, got %d",
               p.vl.bx);
    const size_t lds = plan_vadv_lds(p.prec, p.max_nz, variant, vadv_ws_block(&p.vl));
    SB_REQUIRE(lds <= ctx->lds_per_block,
               "tallest column nz = %d needs %zu bytes of LDS, device allows %zu", p.max_nz, lds,
               ctx->lds_per_block);
    return 0;
}

int run_plan(Plan &p, int use_graph, double *seconds) {
    Context *ctx = context();
    if (!ctx) return 1;
    if (p.count == 0 || p.steps == 0) {
        if (seconds) *seconds = 0.0;
        return 0;
    }
    if (use_graph && !p.exec) {
        SB_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t e = enqueue_any(p, ctx);
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
        SB_HIP(e);
        SB_HIP(e2);
        p.graph = graph;
        SB_HIP(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
    }
    if (seconds) SB_HIP(hipEventRecord(ctx->start, ctx->stream));
    if (use_graph)
        SB_HIP(hipGraphLaunch(p.exec, ctx->stream));
    else
        SB_HIP(enqueue_any(p, ctx));
    if (seconds) {
        SB_HIP(hipEventRecord(ctx->stop, ctx->stream));
        SB_HIP(hipEventSynchronize(ctx->stop));
        float ms = 0.f;
        SB_HIP(hipEventElapsedTime(&ms, ctx->start, ctx->stop));
        *seconds = double(ms) * 1e-3;
    }
    return 0;
}

void destroy_plan(Plan *p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->dev) (void)hipFree(p->dev);
    for (int b = 0; b < 2; ++b) {
        if (p->stage_done[b]) {
            (void)hipEventSynchronize(p->stage_done[b]);
            (void)hipEventDestroy(p->stage_done[b]);
        }
        if (p->stage[b]) (void)hipHostFree(p->stage[b]);
    }
    delete p;
}

}  // namespace
}  // namespace sb

using namespace sb;

extern "C" {

int sb_batch_plan_create(int prec, const sb_subdomain *subs, int count, int steps,
                         const sb_launch *hdiff_launch, const sb_launch *vadv_launch,
                         void **plan) {
    SB_REQUIRE(plan, "NULL plan pointer");
    SB_REQUIRE(steps >= 0, "negative step count");
    if (validate(prec, subs, count)) return 2;
    Context *ctx = context();
    if (!ctx) return 1;
    std::unique_ptr<Plan> p(new Plan());
    p->prec = prec;
    p->count = count;
    p->steps = steps;
    if (hdiff_launch) p->hl = *hdiff_launch;
    if (vadv_launch) p->vl = *vadv_launch;
    build_tables(prec, subs, count, hdiff_launch, *p, ctx);
    if (check_lds(*p, ctx)) return 2;
    // a plan may be captured into a graph later: start it from an idle stream
    if (upload_tables(*p, ctx) || sync_stream(ctx)) {
        destroy_plan(p.release());
        return 1;
    }
    *plan = p.release();
    return 0;
}

int sb_batch_plan_run(void *plan, int use_graph, double *seconds) {
    SB_REQUIRE(plan, "NULL plan");
    return run_plan(*static_cast<Plan *>(plan), use_graph, seconds);
}

int sb_batch_plan_destroy(void *plan) {
    Context *ctx = context();
    if (ctx) SB_HIP(hipStreamSynchronize(ctx->stream));
    destroy_plan(static_cast<Plan *>(plan));
    return 0;
}

int sb_batch_step(int prec, const sb_subdomain *subs, int count, const sb_launch *hdiff_launch,
                  const sb_launch *vadv_launch, double *seconds) {
    if (validate(prec, subs, count)) return 2;
    Context *ctx = context();
    if (!ctx) return 1;
    // one reusable single-step plan per device, shared by the calling
    // threads (serialised by a lock; its tables are rebuilt per call and
    // uploaded in stream order behind the previous call's kernels), held
    // until the process ends like the device context
    static std::mutex step_mutex;
    static Plan *step_plans[64] = {};
    std::lock_guard<std::mutex> lock(step_mutex);
    Plan *&p = step_plans[ctx->device & 63];
    if (!p) p = new Plan();
    p->prec = prec;
    p->count = count;
    p->steps = 1;
    p->hl = hdiff_launch ? *hdiff_launch : sb_launch{};
    p->vl = vadv_launch ? *vadv_launch : sb_launch{};
    build_tables(prec, subs, count, hdiff_launch, *p, ctx);
    if (check_lds(*p, ctx)) return 2;
    if (upload_tables(*p, ctx)) return 1;
    return run_plan(*p, 0, seconds);
}

}  // extern "C"
