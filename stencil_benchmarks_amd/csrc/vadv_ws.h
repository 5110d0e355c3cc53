// Variant 4 of the vertical-advection solver: warp-specialised persistent
// column solver, shared by the plain kernel (vadv.hip) and the batched
// kernel (batch.hip).  The arithmetic is the one documented in vadv.hip,
// operation by operation.
//
// Why.  With (c', d') of a 64-column group in LDS, nz = 160 leaves room for
// a single group per CU, and one wave doing the whole forward/backward
// work for it issues ~70 instructions per level from one SIMD: the other
// three SIMDs idle and the wave is issue- and latency-bound.  Here the
// work of a level is split over four waves (one per SIMD) that share the
// same LDS slot buffer (slot = (c, d) pair of the arithmetic type):
//   B  (wave 0): backward substitution of group n-1,
//                x(k) = d'(k) - c'(k) x(k+1), tensstage(k) = dtr (x(k) - pos(k));
//   P1 (wave 1): d0(k) = (dtr pos + tens) + tensstage of group n -> slot.d;
//   P2 (wave 2): e(k) = d0(k) + corr(k) -> slot.d and the wcon sum
//                s(k+1) = wcon(i+is, k+1) + wcon(k+1) -> slot.c;
//   C  (wave 3): the recurrence of group n from slot (e, s) alone:
//                a, b, c from s(k), s(k+1), then c'(k), d'(k) -> slot.
// With five waves (a choice of the batched kernel) a fifth wave W takes
// P2's wcon loads: it writes s(k+1) -> slot.c on P1's steps (the rows P1
// fills are free in slot.c too), and P2 reads s from the slot; neither the
// skew nor the barrier count changes.
// C issues no global memory access: it runs the serial dependence chain
// only.  P1, P2 and B are independent per level and keep many loads in
// flight.  A single wave per SIMD issues about one instruction every 7-9
// cycles, so every role keeps its per-level instruction count small:
// blocks without a first, last or partial level take a path with no
// per-level condition, LDS slots are addressed by one base per block plus
// constant offsets, and load offsets advance by one saturating add.
//
// Schedule.  Levels are processed in blocks of M (4 by default, or 8: fewer
// barriers per level against a pipeline fill and drain of the same three
// steps, which suits some heights; the launch chooses); one workgroup
// barrier ends every step.  Iteration n (forward group n, backward group
// n-1) has nb + 3 steps, nb = ceil(nz / M):
//   B  at step t in [0, nb):      block t of group n-1, from the top down;
//   P1 at step t in [1, nb]:      block t-1 of group n;
//   P2 at step t in [2, nb+1]:    block t-2 of group n;
//   C  at step t in [3, nb+2]:    block t-3 of group n.
// Slots.  The buffer has nzp = nb M rows; with pad = nzp - nz, group
// n keeps level k in row pad + k (n even) or nzp - 1 - k (n odd).  Block t
// of the backward group and block t of the forward group are then the same
// rows; every wave touches a different block in a step, each block is
// handed on (B -> P1 -> P2 -> C) across a barrier, and B reads group n's
// rows in iteration n+1, after C has finished.  LDS: 2 nzp x 64 values of
// the arithmetic type (nz <= 160 for FP64 in 160 KB).  Every wave passes
// the same nb + 3 barriers per iteration.
//
// Prefetch.  The loading waves keep a ring of their next levels' loads
// (statically indexed: the block loop is unrolled ring / M times),
// refill an entry right after consuming it, and load the first ring of
// their next group during their idle steps.  Loads and stores are never
// under a branch, which lets the compiler count the outstanding loads
// exactly instead of draining them: past the column, a load re-reads
// level nz-1 (clamped) and a store goes to an out-of-range buffer offset,
// which the hardware drops.
#pragma once

#include "vadv_column.h"

namespace sb {
namespace {

constexpr int kWsWaves = 4;  // default waves per workgroup (the other choice: 5, W below)
constexpr int kWsBlock = 4;  // default levels per step (the other choice: 8)
constexpr int kWsSkew = 3;   // steps from B's read of a block to C's write
constexpr int kWsRing = 16;  // prefetch ring of the loading waves, levels
constexpr uint32_t kWsDropOffset = 0xFFFFFF00u;  // past any field (< 4 GiB - 256 B)

// slot rows for nz levels in blocks of `block` levels
__host__ __device__ constexpr int ws_rows(int64_t nz, int block = kWsBlock) {
    return int((nz + block - 1) / block) * block;
}

// Workgroup barrier ordering LDS only (global loads stay in flight).
__device__ __forceinline__ void ws_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Division and reciprocal of the recurrence.  FAST (FP64 arithmetic only):
// hardware reciprocal estimate refined by two Newton steps, about 1 ulp
// instead of the correctly rounded quotient; used only where the result
// is rounded to float32 afterwards.
template <typename C, bool FAST>
__device__ __forceinline__ C ws_recip(C x) {
    if constexpr (FAST && sizeof(C) == 8) {
        C r = __builtin_amdgcn_rcp(x);
        C e = __builtin_fma(-x, r, C(1));
        r = __builtin_fma(r, e, r);
        e = __builtin_fma(-x, r, C(1));
        return __builtin_fma(r, e, r);
    } else {
        return C(1) / x;
    }
}

template <typename C, bool FAST>
__device__ __forceinline__ C ws_div(C n, C d) {
    if constexpr (FAST && sizeof(C) == 8)
        return n * ws_recip<C, true>(d);
    else
        return n / d;
}

// One column sequence: groups first, first + stride, ..., count of them.
struct WsSeq {
    int64_t first, stride, count;
};

// Field pointers (interior origins) of one component; each role builds the
// buffer resources of the fields it touches itself, after branching on its
// role, which keeps the other roles' descriptors out of its SGPRs.
struct WsFields {
    const void *stage, *pos, *tens;
    void *ts;
    const void *wcon;
    uint32_t span_bytes;
};

__device__ __forceinline__ Rsrc ws_rsrc(const WsFields &f) {
    return Rsrc{make_rsrc(f.stage, f.span_bytes), make_rsrc(f.pos, f.span_bytes),
                make_rsrc(f.tens, f.span_bytes), make_rsrc(f.ts, f.span_bytes),
                make_rsrc(f.wcon, f.span_bytes)};
}

template <typename T, typename C, int AUX, bool FAST, int MB = kWsBlock, int NW = kWsWaves>
struct WsSolver {
    static constexpr int M = MB, R = kWsRing, U = kWsRing / MB;
    static_assert(MB == 4 || MB == 8, "levels per step: 4 or 8");
    static_assert(NW == 4 || NW == 5, "waves per workgroup: 4 or 5");
    // with five waves the wcon loads move from P2 to a wave of their own (W)
    static constexpr bool kSplitW = NW == 5;
    static constexpr int RW = kSplitW ? 1 : R;  // P2's wcon ring (unused when split)

    struct Col {  // one group's addressing
        GroupOff g;
        int nz, nb, nzp, pad;
        uint32_t skb, last;  // level stride and offset of level nz-1 (bytes)
        __device__ __forceinline__ uint32_t at(int k) const { return g.row + uint32_t(k) * skb; }
        __device__ __forceinline__ uint32_t clamped(int k) const {
            return at(min(max(k, 0), nz - 1));
        }
        // saturating advance of a forward load offset (no 32-bit overflow:
        // the launch checks span + skb < 2^32)
        __device__ __forceinline__ uint32_t next(uint32_t so) const { return min(so + skb, last); }
        // first row of forward block b in orientation ODD; level j of the block
        // is at +j (even) or +(M-1-j) (odd)
        template <bool ODD>
        __device__ __forceinline__ int row0(int b) const {
            return ODD ? nzp - b * M - M : pad + b * M;
        }
    };

    __device__ __forceinline__ static Col column(const PipeArgs &a, int64_t group, int64_t wshift) {
        Col c;
        c.g = group_offsets<T>(a, group, wshift);
        c.nz = int(a.nz);
        c.nb = (c.nz + M - 1) / M;
        c.nzp = c.nb * M;
        c.pad = c.nzp - c.nz;
        c.skb = uint32_t(a.sk * int64_t(sizeof(T)));
        c.last = c.at(c.nz - 1);
        return c;
    }

    template <bool ODD>
    __device__ __forceinline__ static constexpr int in_block(int j) {
        return (ODD ? M - 1 - j : j) * 64;
    }

    // n-th group of the sequence (clamped: a prologue past the last group
    // re-loads the last one, so prologue loads need no branch)
    __device__ __forceinline__ static int64_t group(const WsSeq &q, int64_t n) {
        return q.first + min(n, q.count - 1) * q.stride;
    }

    __device__ __forceinline__ static void idle(int steps) {
        for (int t = 0; t < steps; ++t) ws_barrier();
    }

    // ---- B: backward substitution ------------------------------------------
    struct BState {
        T ring[R];  // pos, i-th level from the top at i mod R
        uint32_t so_ld;
    };

    __device__ __forceinline__ static void b_prologue(const Rsrc &Rs, const Col &c, BState &s) {
#pragma unroll
        for (int i = 0; i < R; ++i)
            s.ring[i] = bload<T, 0>(Rs.pos, c.g.lane, c.clamped(c.nz - 1 - i));
        s.so_ld = c.clamped(c.nz - 1 - R);
    }

    // Steps 0 .. nb-1 of an iteration (one barrier per block).  GODD is the
    // slot orientation of the backward group: its block t holds the rows of
    // forward block t of the opposite orientation.
    template <bool GODD>
    __device__ __forceinline__ static void b_blocks(const Rsrc &Rs, const Col &c, BState &s,
                                                    const C *cbuf, const C *dbuf) {
        const C dtr = C(3.0 / 20.0);
        const uint32_t floor_ld = c.g.row + c.skb;
        C x = C(0);
        uint32_t so_st = c.last;  // offset of level kb
        for (int b0 = 0; b0 < c.nb; b0 += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = b0 + u;
                const int rb = c.template row0<!GODD>(b) * 64;
                const C *cb = cbuf + rb;
                const C *db = dbuf + rb;
                C out[M];
                if (b >= 1 && b * M + M <= c.nz) {
                    C dv[M], cv[M];  // block's slot reads first (latency overlaps)
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        dv[j] = db[in_block<!GODD>(j)];
                        cv[j] = cb[in_block<!GODD>(j)];
                    }
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        x = dv[j] - cv[j] * x;
                        out[j] = dtr * (x - C(s.ring[u * M + j]));
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        const int i = b * M + j;  // i-th level from the top
                        if (i < c.nz) {
                            const int o = in_block<!GODD>(j);
                            const C d = db[o];
                            x = i == 0 ? d : d - cb[o] * x;
                        }
                        out[j] = dtr * (x - C(s.ring[u * M + j]));
                    }
                }
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    // past the column: out-of-range offset, the store is dropped
                    const bool live = b * M + j < c.nz;
                    bstore<T>(Rs.ts, live ? c.g.lane : kWsDropOffset, live ? so_st : 0u, T(out[j]));
                    so_st -= c.skb;
                    s.ring[u * M + j] = bload<T, 0>(Rs.pos, c.g.lane, s.so_ld);
                    s.so_ld = s.so_ld >= floor_ld ? s.so_ld - c.skb : c.g.row;
                }
                if (b < c.nb) ws_barrier();
            }
        }
    }

    // Group loops run in pairs (even-slot group, odd-slot group) so a ring
    // flows through straight-line code: a join between the two slot
    // orientations would copy registers with loads in flight and drain them.
    __device__ __forceinline__ static void backward(const WsFields &f, const PipeArgs &a,
                                                    const WsSeq &q, C *cbuf, C *dbuf) {
        const Rsrc Rs = ws_rsrc(f);
        const int nz = int(a.nz), nb = (nz + M - 1) / M;
        BState s;
        // iteration 0: no backward group; load group 0's ring at step nb
        idle(nb);
        Col c = column(a, group(q, 0), 0);
        b_prologue(Rs, c, s);
        idle(kWsSkew);
        int64_t n = 0;
        for (; n + 1 < q.count; n += 2) {
            b_blocks<false>(Rs, c, s, cbuf, dbuf);
            c = column(a, group(q, n + 1), 0);
            b_prologue(Rs, c, s);
            idle(kWsSkew);
            b_blocks<true>(Rs, c, s, cbuf, dbuf);
            c = column(a, group(q, n + 2), 0);
            b_prologue(Rs, c, s);
            idle(kWsSkew);
        }
        if (n < q.count) {
            b_blocks<false>(Rs, c, s, cbuf, dbuf);
            idle(kWsSkew);
        }
    }

    // ---- P1: d0 = (dtr pos + tens) + tensstage ---------------------------------
    struct P1State {
        T rp[R], rt[R], rs[R];  // level k at k mod R
        uint32_t so;            // offset of the next level to load
    };

    __device__ __forceinline__ static void p1_load(const Rsrc &Rs, const Col &c, P1State &s, int r,
                                                   uint32_t so) {
        s.rp[r] = bload<T, 0>(Rs.pos, c.g.lane, so);
        s.rt[r] = bload<T, AUX>(Rs.tens, c.g.lane, so);
        s.rs[r] = bload<T, AUX>(Rs.ts, c.g.lane, so);
    }

    __device__ __forceinline__ static void p1_prologue(const Rsrc &Rs, const Col &c, P1State &s) {
#pragma unroll
        for (int i = 0; i < R; ++i) p1_load(Rs, c, s, i, c.clamped(i));
        s.so = c.clamped(R);
    }

    template <bool ODD>
    __device__ __forceinline__ static void p1_blocks(const Rsrc &Rs, const Col &c, P1State &s,
                                                     C *dbuf) {
        const C dtr = C(3.0 / 20.0);
        for (int b0 = 0; b0 < c.nb; b0 += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = b0 + u;
                C *db = dbuf + c.template row0<ODD>(b) * 64;
                C d0[M];
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    const int r = u * M + j;
                    d0[j] = (dtr * C(s.rp[r]) + C(s.rt[r])) + C(s.rs[r]);
                }
                if (b * M + M <= c.nz) {
#pragma unroll
                    for (int j = 0; j < M; ++j) db[in_block<ODD>(j)] = d0[j];
                } else {
#pragma unroll
                    for (int j = 0; j < M; ++j)
                        if (b * M + j < c.nz) db[in_block<ODD>(j)] = d0[j];
                }
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    p1_load(Rs, c, s, u * M + j, s.so);
                    s.so = c.next(s.so);
                }
                if (b < c.nb) ws_barrier();
            }
        }
    }

    // P1 works at steps 1 .. nb of an iteration and loads the next group's
    // ring in its idle steps nb+1, nb+2.
    __device__ __forceinline__ static void prep_d0(const WsFields &f, const PipeArgs &a,
                                                   const WsSeq &q, C *dbuf) {
        const Rsrc Rs = ws_rsrc(f);
        const int nz = int(a.nz), nb = (nz + M - 1) / M;
        P1State s;
        Col c = column(a, group(q, 0), 0);
        p1_prologue(Rs, c, s);
        int64_t n = 0;
        for (; n + 1 < q.count; n += 2) {
            idle(1);
            p1_blocks<false>(Rs, c, s, dbuf);
            c = column(a, group(q, n + 1), 0);
            p1_prologue(Rs, c, s);
            idle(kWsSkew);  // steps nb+1, nb+2 and step 0 of the next iteration
            p1_blocks<true>(Rs, c, s, dbuf);
            c = column(a, group(q, n + 2), 0);
            p1_prologue(Rs, c, s);
            idle(kWsSkew - 1);
        }
        if (n < q.count) {
            idle(1);
            p1_blocks<false>(Rs, c, s, dbuf);
            idle(kWsSkew - 1);
        }
        idle(nb + kWsSkew);  // last iteration: backward only
    }

    // ---- W (five waves only): s(k+1) = wcon(i+is, k+1) + wcon(k+1) -> slot.c ----
    // at P1's steps (the rows P1 fills with d0 are free in slot.c too), so
    // P2 reads s from the slot instead of loading wcon
    struct WState {
        T rw[R], rwn[R];  // wcon and its neighbour at level k+1, level k at k mod R
        uint32_t so;      // offset of the next level to load
    };

    __device__ __forceinline__ static void w_load(const Rsrc &Rs, const Col &c, WState &s, int r,
                                                  uint32_t so) {
        s.rw[r] = bload<T, 0>(Rs.w, c.g.lane, so);
        s.rwn[r] = bload<T, 0>(Rs.w, c.g.lane_wn, so);
    }

    __device__ __forceinline__ static void w_prologue(const Rsrc &Rs, const Col &c, WState &s) {
#pragma unroll
        for (int i = 0; i < R; ++i) w_load(Rs, c, s, i, c.clamped(i + 1));
        s.so = c.clamped(R + 1);
    }

    template <bool ODD>
    __device__ __forceinline__ static void w_blocks(const Rsrc &Rs, const Col &c, WState &s,
                                                    C *cbuf) {
        for (int b0 = 0; b0 < c.nb; b0 += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = b0 + u;
                C *cb = cbuf + c.template row0<ODD>(b) * 64;
                C sn[M];
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    const int r = u * M + j;
                    sn[j] = C(s.rwn[r]) + C(s.rw[r]);
                }
                if (b * M + M <= c.nz) {
#pragma unroll
                    for (int j = 0; j < M; ++j) cb[in_block<ODD>(j)] = sn[j];
                } else {
#pragma unroll
                    for (int j = 0; j < M; ++j)
                        if (b * M + j < c.nz) cb[in_block<ODD>(j)] = sn[j];
                }
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    w_load(Rs, c, s, u * M + j, s.so);
                    s.so = c.next(s.so);
                }
                if (b < c.nb) ws_barrier();
            }
        }
    }

    // W keeps P1's schedule: steps 1 .. nb, next ring in steps nb+1, nb+2.
    __device__ __forceinline__ static void prep_s(const WsFields &f, const PipeArgs &a,
                                                  int64_t wshift, const WsSeq &q, C *cbuf) {
        const Rsrc Rs = ws_rsrc(f);
        const int nz = int(a.nz), nb = (nz + M - 1) / M;
        WState s;
        Col c = column(a, group(q, 0), wshift);
        w_prologue(Rs, c, s);
        int64_t n = 0;
        for (; n + 1 < q.count; n += 2) {
            idle(1);
            w_blocks<false>(Rs, c, s, cbuf);
            c = column(a, group(q, n + 1), wshift);
            w_prologue(Rs, c, s);
            idle(kWsSkew);
            w_blocks<true>(Rs, c, s, cbuf);
            c = column(a, group(q, n + 2), wshift);
            w_prologue(Rs, c, s);
            idle(kWsSkew - 1);
        }
        if (n < q.count) {
            idle(1);
            w_blocks<false>(Rs, c, s, cbuf);
            idle(kWsSkew - 1);
        }
        idle(nb + kWsSkew);
    }

    // ---- P2: e = d0 + corr, s(k+1) -----------------------------------------------
    struct P2State {
        T rst[R], rw[RW], rwn[RW];  // stage, wcon, neighbour at level k+1
        T st0;
        uint32_t so;                // offset of the next level to load
    };

    __device__ __forceinline__ static void p2_load(const Rsrc &Rs, const Col &c, P2State &s, int r,
                                                   uint32_t so) {
        s.rst[r] = bload<T, AUX>(Rs.stage, c.g.lane, so);
        if constexpr (!kSplitW) {
            s.rw[r] = bload<T, 0>(Rs.w, c.g.lane, so);
            s.rwn[r] = bload<T, 0>(Rs.w, c.g.lane_wn, so);
        }
    }

    __device__ __forceinline__ static void p2_prologue(const Rsrc &Rs, const Col &c, P2State &s) {
        s.st0 = bload<T, AUX>(Rs.stage, c.g.lane, c.at(0));
#pragma unroll
        for (int i = 0; i < R; ++i) p2_load(Rs, c, s, i, c.clamped(i + 1));
        s.so = c.clamped(R + 1);
    }

    // corr = (-as)(stage(k-1) - stage(k)) - cs (stage(k+1) - stage(k)), with
    // as(k) = (-0.25 s(k)) bet_m = -cs(k-1) bit for bit (scaling by a power
    // of two and negation commute with rounding); one-sided at the first and
    // last level.
    template <bool ODD>
    __device__ __forceinline__ static void p2_blocks(const Rsrc &Rs, const Col &c, P2State &s,
                                                     C *cbuf, C *dbuf) {
        const C bet_m = C(0.5);
        C st_m = C(0), st_c = C(s.st0), cs_p = C(0);
        for (int b0 = 0; b0 < c.nb; b0 += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = b0 + u;
                const int rb = c.template row0<ODD>(b) * 64;
                C *cb = cbuf + rb;
                C *db = dbuf + rb;
                if (b >= 1 && b * M + M <= c.nz - 1) {
                    C d0[M], sv[M];  // block's slot reads first (latency overlaps)
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        d0[j] = db[in_block<ODD>(j)];
                        if constexpr (kSplitW) sv[j] = cb[in_block<ODD>(j)];
                    }
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        const int r = u * M + j, o = in_block<ODD>(j);
                        const C st_n = C(s.rst[r]);
                        C s_n;
                        if constexpr (kSplitW)
                            s_n = sv[j];
                        else
                            s_n = C(s.rwn[r]) + C(s.rw[r]);
                        const C cs = (C(0.25) * s_n) * bet_m;
                        const C corr = cs_p * (st_m - st_c) - cs * (st_n - st_c);
                        db[o] = d0[j] + corr;  // e = d0 + corr
                        if constexpr (!kSplitW) cb[o] = s_n;
                        st_m = st_c;
                        st_c = st_n;
                        cs_p = cs;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        const int r = u * M + j, o = in_block<ODD>(j);
                        const int k = b * M + j;
                        const C st_n = C(s.rst[r]);
                        if (k < c.nz) {  // slot rows exist for levels below nz only
                            C s_n;
                            if constexpr (kSplitW)
                                s_n = cb[o];
                            else
                                s_n = C(s.rwn[r]) + C(s.rw[r]);
                            const C cs = (C(0.25) * s_n) * bet_m;
                            C corr;
                            if (k == 0)
                                corr = -cs * (st_n - st_c);
                            else if (k < c.nz - 1)
                                corr = cs_p * (st_m - st_c) - cs * (st_n - st_c);
                            else
                                corr = cs_p * (st_m - st_c);
                            db[o] = db[o] + corr;
                            if constexpr (!kSplitW) cb[o] = s_n;
                            cs_p = cs;
                        }
                        st_m = st_c;
                        st_c = st_n;
                    }
                }
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    p2_load(Rs, c, s, u * M + j, s.so);
                    s.so = c.next(s.so);
                }
                if (b < c.nb) ws_barrier();
            }
        }
    }

    // P2 works at steps 2 .. nb+1 and loads the next ring in its idle step nb+2.
    __device__ __forceinline__ static void prep_e(const WsFields &f, const PipeArgs &a,
                                                  int64_t wshift, const WsSeq &q, C *cbuf,
                                                  C *dbuf) {
        const Rsrc Rs = ws_rsrc(f);
        const int nz = int(a.nz), nb = (nz + M - 1) / M;
        P2State s;
        Col c = column(a, group(q, 0), wshift);
        p2_prologue(Rs, c, s);
        int64_t n = 0;
        for (; n + 1 < q.count; n += 2) {
            idle(2);
            p2_blocks<false>(Rs, c, s, cbuf, dbuf);
            c = column(a, group(q, n + 1), wshift);
            p2_prologue(Rs, c, s);
            idle(kWsSkew);  // step nb+2 and steps 0, 1 of the next iteration
            p2_blocks<true>(Rs, c, s, cbuf, dbuf);
            c = column(a, group(q, n + 2), wshift);
            p2_prologue(Rs, c, s);
            idle(kWsSkew - 2);
        }
        if (n < q.count) {
            idle(2);
            p2_blocks<false>(Rs, c, s, cbuf, dbuf);
            idle(kWsSkew - 2);
        }
        idle(nb + kWsSkew);
    }

    // ---- C: the recurrence -------------------------------------------------------
    struct Chain {
        C cv, cc, dd;  // c(k-1) = (0.25 s(k)) bet_p, c'(k-1), d'(k-1)
    };

    // level in [1, nz-2] with e(k), s(k+1) read; a(k) = -c(k-1) bit for bit
    __device__ __forceinline__ static void c_middle(C e, C s_n, Chain &h, C &cc_out, C &dd_out) {
        const C dtr = C(3.0 / 20.0), bet_p = C(0.5);
        const C av = -h.cv;
        const C cv = (C(0.25) * s_n) * bet_p;
        const C bv = (dtr - av) - cv;
        const C rr = ws_recip<C, FAST>(bv - h.cc * av);
        h.cc = cv * rr;
        h.dd = (e - h.dd * av) * rr;
        h.cv = cv;
        cc_out = h.cc;
        dd_out = h.dd;
    }

    template <bool ODD>
    __device__ __forceinline__ static void c_any(int k, int nz, C *cb, C *db, int o, Chain &h) {
        const C dtr = C(3.0 / 20.0), bet_p = C(0.5);
        if (k == 0) {
            const C e = db[o], s_n = cb[o];
            const C cv = (C(0.25) * s_n) * bet_p;
            const C bv = dtr - cv;
            h.cc = ws_div<C, FAST>(cv, bv);
            h.dd = ws_div<C, FAST>(e, bv);
            h.cv = cv;
            cb[o] = h.cc;
            db[o] = h.dd;
        } else if (k < nz - 1) {
            C cc, dd;
            c_middle(db[o], cb[o], h, cc, dd);
            cb[o] = cc;
            db[o] = dd;
        } else {
            const C e = db[o];
            const C av = -h.cv;
            const C bv = dtr - av;
            db[o] = ws_div<C, FAST>(e - h.dd * av, bv - h.cc * av);
        }
    }

    // C works at steps 3 .. nb+2: first block, middle blocks, last block.
    template <bool ODD>
    __device__ __forceinline__ static void c_blocks(const Col &c, C *cbuf, C *dbuf) {
        Chain h{C(0), C(0), C(0)};
        {
            const int rb = c.template row0<ODD>(0) * 64;
            for (int j = 0; j < M; ++j)
                if (j < c.nz) c_any<ODD>(j, c.nz, cbuf + rb, dbuf + rb, in_block<ODD>(j), h);
            ws_barrier();
        }
        for (int b = 1; b < c.nb - 1; ++b) {
            const int rb = c.template row0<ODD>(b) * 64;
            C *cb = cbuf + rb;
            C *db = dbuf + rb;
            // read the block's inputs first: the LDS latency then overlaps
            // the recurrence instead of adding to every level
            C e[M], sn[M], cc[M], dd[M];
#pragma unroll
            for (int j = 0; j < M; ++j) {
                e[j] = db[in_block<ODD>(j)];
                sn[j] = cb[in_block<ODD>(j)];
            }
#pragma unroll
            for (int j = 0; j < M; ++j) c_middle(e[j], sn[j], h, cc[j], dd[j]);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                cb[in_block<ODD>(j)] = cc[j];
                db[in_block<ODD>(j)] = dd[j];
            }
            ws_barrier();
        }
        if (c.nb >= 2) {
            const int b = c.nb - 1;
            const int rb = c.template row0<ODD>(b) * 64;
            for (int j = 0; j < M; ++j) {
                const int k = b * M + j;
                if (k < c.nz) c_any<ODD>(k, c.nz, cbuf + rb, dbuf + rb, in_block<ODD>(j), h);
            }
            ws_barrier();
        }
    }

    __device__ __forceinline__ static void chain(const PipeArgs &a, const WsSeq &q, C *cbuf,
                                                 C *dbuf) {
        const int nz = int(a.nz), nb = (nz + M - 1) / M;
        const Col c = column(a, q.first, 0);  // geometry only (no loads)
        for (int64_t n = 0; n < q.count; ++n) {
            idle(kWsSkew);
            if ((n & 1) != 0)
                c_blocks<true>(c, cbuf, dbuf);
            else
                c_blocks<false>(c, cbuf, dbuf);
        }
        idle(nb + kWsSkew);
    }

    // All NW waves of the workgroup call this with the same arguments.
    __device__ __forceinline__ static void run(const WsFields &f, int64_t wshift, const PipeArgs &a,
                                               const WsSeq &q, C *cbuf, C *dbuf) {
        const int role = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
        if (role == 0)
            backward(f, a, q, cbuf, dbuf);
        else if (role == 1)
            prep_d0(f, a, q, dbuf);
        else if (role == 2)
            prep_e(f, a, wshift, q, cbuf, dbuf);
        else if (role == 3)
            chain(a, q, cbuf, dbuf);
        else if constexpr (kSplitW)
            prep_s(f, a, wshift, q, cbuf);
    }
};

template <typename C>
constexpr size_t ws_lds_bytes(int64_t nz, int block = kWsBlock) {
    return size_t(2 * ws_rows(nz, block)) * 64 * sizeof(C);
}

}  // namespace
}  // namespace sb
