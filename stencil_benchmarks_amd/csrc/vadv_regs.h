// Variant 5 of the vertical-advection solver: one wave per column group,
// (c', d') of the lowest levels in registers and of the rest in LDS.
//
// Why.  The warp-specialised solver keeps all of (c', d') in LDS and
// pipelines four role waves over a sequence of column groups; with few groups
// per workgroup (256 x 256 x 80 gives two per workgroup at two workgroups per
// CU) its pipeline fill and drain are a third of its steps.  Here one wave
// solves one 64-column group from its first level to its last and back.  The
// lowest NR levels of (c', d') live in registers (4 NR 32-bit registers with
// FP64 arithmetic), the levels above in LDS, so a short column takes a
// fraction of the LDS a whole column would: at nz = 80 and NR = 40 a wave
// needs 40 KB, four waves (one per SIMD) share a CU, and every group of a
// 256 x 256 x 80 domain is in flight at once with no fill or drain.
//
// Register indices must be compile-time constants, so the register levels
// are the lowest ones: the forward sweep runs levels 0 .. NR-1 fully unrolled
// and then the levels above in blocks of PF through LDS (row k - NR); the
// backward sweep takes the LDS levels from the top down and then the
// register levels, again unrolled.  nz is a run-time value: levels at or
// past nz - 1 run the same instructions on clamped inputs (loads re-read
// level nz-1; stores go to an out-of-range offset the hardware drops), and
// the forward state entering level nz - 1 is captured by selects, so no load
// or store sits under a branch and the compiler can keep the prefetch rings
// in flight.  The arithmetic is the one documented in vadv.hip, operation by
// operation (bit-identical to the other variants).
#pragma once

#include "vadv_column.h"

namespace sb {
namespace {

constexpr uint32_t kRegDropOffset = 0xFFFFFF00u;  // past any field (< 4 GiB - 256 B)

// LDS rows (levels NR .. nz-2, rounded up to whole blocks of PF) of one wave.
__host__ __device__ constexpr int64_t hybrid_lds_rows(int64_t nz, int nr, int pf) {
    return nz - 1 > nr ? (nz - 1 - nr + pf - 1) / pf * pf : 0;
}

template <typename T, typename C, int NR, int PF, int AUX>
struct HybridColumn {
    static_assert(NR % PF == 0 && NR >= PF, "register levels: whole prefetch blocks");
    // The forward sweep keeps pos of the lowest NK levels in registers (all
    // NR register levels with float32 fields; 16 with FP64 fields, which
    // fill the wave's 512 registers without scratch, where 8, 12 and 20
    // spilled), so the backward sweep reads pos from memory for the levels
    // above only.
    static constexpr int NK = sizeof(T) == 4 ? NR : 16;

    // forward state entering a level: stage(k-1), stage(k), wcon and its
    // neighbour at k, c'(k-1), d'(k-1)
    struct State {
        C st_m, st_c, w_c, wn_c, c_prev, d_prev;
    };

    __device__ __forceinline__ static State hold(bool at_last, const State &s, const State &h) {
        return State{at_last ? s.st_m : h.st_m,   at_last ? s.st_c : h.st_c,
                     at_last ? s.w_c : h.w_c,     at_last ? s.wn_c : h.wn_c,
                     at_last ? s.c_prev : h.c_prev, at_last ? s.d_prev : h.d_prev};
    }

    // level in [1, nz-2]: c'(k), d'(k) from the level record and the state
    __device__ __forceinline__ static void middle(const Level<T> &r, State &s, C &c_out, C &d_out) {
        const C dtr = C(3.0 / 20.0), bet_m = C(0.5), bet_p = C(0.5);
        const C gav = C(-0.25) * (s.wn_c + s.w_c);
        const C w_n = C(r.w), wn_n = C(r.wn), st_n = C(r.st);
        const C gcv = C(0.25) * (wn_n + w_n);
        const C as = gav * bet_m;
        const C cs = gcv * bet_m;
        const C acol = gav * bet_p;
        C c = gcv * bet_p;
        const C b = (dtr - acol) - c;
        const C corr = -as * (s.st_m - s.st_c) - cs * (st_n - s.st_c);
        C d = ((dtr * C(r.pos) + C(r.tens)) + C(r.ts)) + corr;
        const C divided = C(1.0) / (b - s.c_prev * acol);
        c = c * divided;
        d = (d - s.d_prev * acol) * divided;
        c_out = c;
        d_out = d;
        s = State{s.st_c, st_n, w_n, wn_n, c, d};
    }

    // `lds` is this wave's lane column: row r at lds[64 r].
    __device__ __forceinline__ static void solve(const Rsrc &R, int64_t wshift, const PipeArgs &a,
                                                 int64_t g, CD<C> *lds) {
        const int nz = int(a.nz);
        const int kl = nz - 1;
        const uint32_t skb = uint32_t(a.sk * int64_t(sizeof(T)));
        const GroupOff F = group_offsets<T>(a, g, wshift);
        const C dtr = C(3.0 / 20.0), bet_m = C(0.5), bet_p = C(0.5);
        const uint32_t last = F.row + uint32_t(kl) * skb;  // offset of level nz-1
        auto off = [&](int k) { return F.row + uint32_t(min(max(k, 0), kl)) * skb; };

        // level record k: pos, tens, tensstage at k; stage, wcon and its
        // neighbour at k + 1 (clamped to nz - 1)
        auto load = [&](int k, Level<T> &r) {
            const uint32_t so = off(k), sn = off(k + 1);
            r.pos = bload<T, 0>(R.pos, F.lane, so);
            r.tens = bload<T, AUX>(R.tens, F.lane, so);
            r.ts = bload<T, AUX>(R.ts, F.lane, so);
            r.st = bload<T, AUX>(R.stage, F.lane, sn);
            r.w = bload<T, AUX>(R.w, F.lane, sn);
            r.wn = bload<T, AUX>(R.w, F.lane_wn, sn);
        };

        // the last level's own inputs, issued first
        const T last_pos = bload<T, 0>(R.pos, F.lane, last);
        const T last_tens = bload<T, AUX>(R.tens, F.lane, last);
        const T last_ts = bload<T, AUX>(R.ts, F.lane, last);
        const T st0 = bload<T, AUX>(R.stage, F.lane, F.row);
        Level<T> ring[PF];
#pragma unroll
        for (int p = 0; p < PF; ++p) load(p, ring[p]);

        C cc[NR], dd[NR];
        T pk[NK];  // pos of the lowest NK levels
        State s{}, h{};
        // ---- level 0 ---------------------------------------------------------
        {
            const Level<T> r = ring[0];
            pk[0] = r.pos;
            load(PF, ring[0]);
            const C gcv = C(0.25) * (C(r.wn) + C(r.w));
            const C cs = gcv * bet_m;
            const C c = gcv * bet_p;
            const C b = dtr - c;
            const C corr = -cs * (C(r.st) - C(st0));
            const C d = ((dtr * C(r.pos) + C(r.tens)) + C(r.ts)) + corr;
            cc[0] = c / b;
            dd[0] = d / b;
            s = State{C(st0), C(r.st), C(r.w), C(r.wn), cc[0], dd[0]};
        }
        // ---- levels 1 .. NR-1: registers ---------------------------------------
#pragma unroll
        for (int k = 1; k < NR; ++k) {
            h = hold(k == kl, s, h);
            const Level<T> r = ring[k % PF];
            if (k < NK) pk[k] = r.pos;
            load(k + PF, ring[k % PF]);
            middle(r, s, cc[k], dd[k]);
        }
        // ---- levels NR .. nz-2: LDS row k - NR, in blocks of PF -----------------
        const int rows = int(hybrid_lds_rows(nz, NR, PF));
        for (int kb = NR; kb < NR + rows; kb += PF) {
#pragma unroll
            for (int p = 0; p < PF; ++p) {
                const int k = kb + p;
                h = hold(k == kl, s, h);
                const Level<T> r = ring[p];
                load(k + PF, ring[p]);
                C c, d;
                middle(r, s, c, d);
                lds[int64_t(k - NR) * 64] = CD<C>{c, d};
            }
        }
        h = hold(NR + rows == kl, s, h);  // nz - 1 = NR + rows exactly

        // ---- level nz-1 from the held state ------------------------------------
        C x;
        {
            const C gav = C(-0.25) * (h.wn_c + h.w_c);
            const C as = gav * bet_m;
            const C acol = gav * bet_p;
            const C b = dtr - acol;
            const C corr = -as * (h.st_m - h.st_c);
            const C d = ((dtr * C(last_pos) + C(last_tens)) + C(last_ts)) + corr;
            x = (d - h.d_prev * acol) / (b - h.c_prev * acol);
        }
        bstore<T>(R.ts, F.lane, last, T(dtr * (x - C(last_pos))));

        // ---- backward: LDS levels from the top, then the register levels ------
        // slot p of the pos ring holds the level of in-block position p
        T pr[PF];
        // a level below NK is kept in pk, so its load is dropped (an
        // out-of-range offset reads nothing and returns 0)
        auto pos_load = [&](int k) {
            const bool kept = k < NK;
            return bload<T, 0>(R.pos, kept ? kRegDropOffset : F.lane, kept ? 0u : off(k));
        };
#pragma unroll
        for (int p = 0; p < PF; ++p) pr[p] = pos_load(NR + rows - PF + p);
        for (int kb = NR + rows - PF; kb >= NR; kb -= PF) {
#pragma unroll
            for (int p = PF - 1; p >= 0; --p) {
                const int k = kb + p;
                const T pos = pr[p];
                pr[p] = pos_load(k - PF);
                const bool live = k < kl;
                const CD<C> cd = lds[int64_t(k - NR) * 64];
                const C xn = cd.d - cd.c * x;
                x = live ? xn : x;
                bstore<T>(R.ts, live ? F.lane : kRegDropOffset, live ? off(k) : 0u,
                          T(dtr * (x - C(pos))));
            }
        }
#pragma unroll
        for (int k = NR - 1; k >= 0; --k) {
            const int p = k % PF;
            T pos;
            if (k < NK) {
                pos = pk[k];
            } else {
                pos = pr[p];
                pr[p] = pos_load(k - PF);
            }
            const bool live = k < kl;
            const C xn = dd[k] - cc[k] * x;
            x = live ? xn : x;
            bstore<T>(R.ts, live ? F.lane : kRegDropOffset, live ? off(k) : 0u,
                      T(dtr * (x - C(pos))));
        }
    }
};

}  // namespace
}  // namespace sb
