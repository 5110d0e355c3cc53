// Column machinery of the vertical-advection Thomas solver shared by the
// plain kernels (vadv.hip) and the batched persistent kernel (batch.hip):
// field/level types and the pipelined group iteration.  The arithmetic is
// documented in vadv.hip.
#pragma once

#include "sb_internal.h"

namespace sb {
namespace {

struct VadvArgs {
    int64_t nx, ny, nz, si, sj, sk;
    int32_t tiles_i;
    int32_t ncomp;
    int64_t scratch_cols;  // columns of the global scratch (variant 2)
    int64_t tiles;         // 64-column tiles (tiles_i x ny); workgroups stride over them
};

struct Component {
    const void *stage, *pos, *tens;
    void *tensstage;
    int64_t wshift;  // element offset of the wcon neighbour (is * si + js * sj)
};

struct Components {
    Component c[3];
};

template <typename C>
struct alignas(2 * sizeof(C)) CD {
    C c, d;
};

// Inputs of one interior level k: pos, tens, tensstage at k and stage,
// wcon, wcon neighbour at k + 1 (stored in the field type T).
template <typename T>
struct Level {
    T pos, tens, ts;
    T st, w, wn;
};

template <bool NT, typename T>
__device__ __forceinline__ T stream_load(const T *p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// ---------------------------------------------------------------------------
// Variant 3: persistent pipelined column solver.
//
// A workgroup (one wave) processes a sequence of 64-column groups g_0, g_1,
// ... (g_n = w + n * gridDim for the workgroup's logical index w, so the
// groups in flight across the chip are neighbours in memory).  The forward sweep of group n runs in the same
// loop as the backward substitution of group n-1: at step s the backward
// handles level nz-1-s of group n-1 and the forward level s-1 of group n.
// Both use ONE LDS buffer of nz-1 (c', d') slots per lane: group n stores
// level k at slot k when n is even and at slot nz-2-k when n is odd, so at
// every step the backward reads exactly the slot the forward writes next
// (read first, then write).  The backward's pos loads and tensstage stores
// join the forward's loads in one prefetch pipeline, and the forward and
// backward recurrences interleave.  Arithmetic is the same as above.
// ---------------------------------------------------------------------------
struct PipeArgs {
    int64_t nx, ny, nz, si, sj, sk;
    int32_t ncomp;
    int64_t ngroups;  // column groups: ceil(nx * ny / 64)
    uint32_t span_bytes;  // addressable bytes from each field's interior origin
    uint32_t xcds;        // XCDs of the device: consecutive groups share one XCD's L2
};

// Buffer-resource addressing: every field is one descriptor based at its
// interior origin (out-of-range reads return 0, writes are dropped), a lane
// adds its 32-bit column byte offset and a level adds the uniform byte
// offset j*sj + k*sk, so a load is a single buffer instruction.
constexpr int kRsrcWord3 = 0x00020000;

template <typename T, int AUX>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
}

// Stores of the solvers' output (tensstage), with the non-temporal bit: the
// kernel never reads the output again.  Measured on MI355X against plain
// stores: 512x512x160 FP64 +1.5-3.5 %, float32 storage +0.5-1 %, 256x256x80
// float32 storage +2-3 % (FP64 within noise), 96-subdomain batched FP64 step
// +2 %; a column-walk probe +1-2 % (non-temporal loads instead cost it 15 %).
constexpr int kStoreAux = 2;

template <typename T>
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, T v) {
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(
            __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r,
            voff, soff, kStoreAux);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, voff, soff, kStoreAux);
}

struct Rsrc {
    __amdgpu_buffer_rsrc_t stage, pos, tens, ts, w;
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), short(0), int(bytes), kRsrcWord3);
}

template <typename C>
struct FwdState {
    C st_m, st_c, w_c, wn_c, c_prev, d_prev;
};

// Column-group geometry: lane byte offsets of the group's columns relative
// to a uniform row byte offset (the row of the group's first column), which
// level offsets add as the buffer instructions' scalar offset.
struct GroupOff {
    uint32_t lane, lane_wn, row;
};

// Number of column groups of an nx x ny domain.
__host__ __device__ constexpr int64_t vadv_groups(int64_t nx, int64_t ny) {
    return (nx * ny + 63) / 64;
}

// Group g holds columns 64 g .. 64 g + 63 of the domain's nx * ny columns in
// i-fastest order: a group runs on into the next row where nx is not a
// multiple of 64, so only the domain's last group can have idle lanes.
// Idle lanes duplicate the last column: same inputs, same outputs, so their
// (identical) stores are harmless.  Column indices fit 32 bits (every field
// is below 4 GiB).
template <typename T>
__device__ __forceinline__ GroupOff group_offsets(const PipeArgs &a, int64_t g, int64_t wshift) {
    const uint32_t nx = uint32_t(a.nx), cols = uint32_t(a.nx * a.ny);
    const uint32_t first = min(uint32_t(g) * 64u, cols - 1);
    const uint32_t j0 = first / nx;  // uniform: the group's first row
    uint32_t col = first + (threadIdx.x & 63);
    col = col < cols ? col : cols - 1;
    const uint32_t j = col / nx, i = col - j * nx;
    const int64_t e = int64_t(i) * a.si + int64_t(j - j0) * a.sj;
    GroupOff o;
    o.lane = uint32_t(e * int64_t(sizeof(T)));
    o.lane_wn = uint32_t((e + wshift) * int64_t(sizeof(T)));
    o.row = uint32_t(int64_t(j0) * a.sj * int64_t(sizeof(T)));
    return o;
}

// One group-iteration: forward of group `gf` (if FWD) fused with the
// backward of group `gb` (if BWD).  `odd` is the slot orientation of the
// forward group; the backward group has the opposite one.
template <typename T, typename C, int PF, int AUX, bool FWD, bool BWD>
__device__ __forceinline__ void pipe_iteration(const Rsrc &R, int64_t wshift, const PipeArgs &a,
                                               int64_t gf, int64_t gb, bool odd, CD<C> *lds,
                                               C &x, T &last_pos) {
    const int nz = int(a.nz);
    const int kl = nz - 1;
    const uint32_t skb = uint32_t(a.sk * int64_t(sizeof(T)));
    const C dtr = C(3.0 / 20.0);
    const C bet_m = C(0.5), bet_p = C(0.5);
    GroupOff F{}, B{};
    if constexpr (FWD) F = group_offsets<T>(a, gf, wshift);
    if constexpr (BWD) B = group_offsets<T>(a, gb, wshift);
    auto fslot = [&](int k) { return odd ? (kl - 1 - k) : k; };
    auto bslot = [&](int k) { return odd ? k : (kl - 1 - k); };

    auto load_level = [&](int k, Level<T> &r) {  // k in [0, nz-2]
        const uint32_t so = F.row + uint32_t(k) * skb;
        r.pos = bload<T, 0>(R.pos, F.lane, so);
        r.tens = bload<T, AUX>(R.tens, F.lane, so);
        r.ts = bload<T, AUX>(R.ts, F.lane, so);
        r.st = bload<T, AUX>(R.stage, F.lane, so + skb);
        r.w = bload<T, AUX>(R.w, F.lane, so + skb);
        r.wn = bload<T, AUX>(R.w, F.lane_wn, so + skb);
    };
    auto load_bpos = [&](int k) { return bload<T, 0>(R.pos, B.lane, B.row + uint32_t(k) * skb); };
    auto store_out = [&](int k, C value) {
        bstore<T>(R.ts, B.lane, B.row + uint32_t(k) * skb, T(value));
    };

    FwdState<C> f{};
    // ---- step 0: backward level kl (register-resident) -------------------
    if constexpr (BWD) store_out(kl, dtr * (x - C(last_pos)));
    // ---- step 1: forward level 0, backward level kl-1 --------------------
    Level<T> cur[PF], nxt[PF];
    T bcur[PF], bnxt[PF];
    {
        Level<T> first{};
        T st0 = T(0);
        if constexpr (FWD) {
            load_level(0, first);
            st0 = bload<T, AUX>(R.stage, F.lane, F.row);
#pragma unroll
            for (int p = 0; p < PF; ++p) load_level(min(1 + p, kl - 1), cur[p]);
        }
        T bpos = T(0);
        if constexpr (BWD) {
            bpos = load_bpos(kl - 1);
#pragma unroll
            for (int p = 0; p < PF; ++p) bcur[p] = load_bpos(max(kl - 2 - p, 0));
            const CD<C> cd = lds[int64_t(bslot(kl - 1)) * 64];
            x = cd.d - cd.c * x;
            store_out(kl - 1, dtr * (x - C(bpos)));
        }
        if constexpr (FWD) {
            const C gcv = C(0.25) * (C(first.wn) + C(first.w));
            const C cs = gcv * bet_m;
            C cc = gcv * bet_p;
            const C bcol = dtr - cc;
            const C corr = -cs * (C(first.st) - C(st0));
            C dd = ((dtr * C(first.pos) + C(first.tens)) + C(first.ts)) + corr;
            cc = cc / bcol;
            dd = dd / bcol;
            lds[int64_t(fslot(0)) * 64] = CD<C>{cc, dd};
            f = FwdState<C>{C(st0), C(first.st), C(first.w), C(first.wn), cc, dd};
        }
    }

    // ---- steps 2 .. nz-1: forward level s-1 (middle), backward kl-s -----
    auto step = [&](int s, const Level<T> &r, T bp) {
        if constexpr (BWD) {
            const int kb = kl - s;
            const CD<C> cd = lds[int64_t(bslot(kb)) * 64];
            x = cd.d - cd.c * x;
            store_out(kb, dtr * (x - C(bp)));
        }
        if constexpr (FWD) {
            const int k = s - 1;
            const C gav = C(-0.25) * (f.wn_c + f.w_c);
            const C w_n = C(r.w), wn_n = C(r.wn), st_n = C(r.st);
            const C gcv = C(0.25) * (wn_n + w_n);
            const C as = gav * bet_m;
            const C cs = gcv * bet_m;
            const C acol = gav * bet_p;
            C cc = gcv * bet_p;
            const C bcol = (dtr - acol) - cc;
            const C corr = -as * (f.st_m - f.st_c) - cs * (st_n - f.st_c);
            C dd = ((dtr * C(r.pos) + C(r.tens)) + C(r.ts)) + corr;
            const C divided = C(1.0) / (bcol - f.c_prev * acol);
            cc = cc * divided;
            dd = (dd - f.d_prev * acol) * divided;
            lds[int64_t(fslot(k)) * 64] = CD<C>{cc, dd};
            f = FwdState<C>{f.st_c, st_n, w_n, wn_n, cc, dd};
        }
    };
    auto prefetch = [&](int s0, Level<T> *lv, T *bp) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            if constexpr (FWD) load_level(min(s0 + p - 1, kl - 1), lv[p]);
            if constexpr (BWD) bp[p] = load_bpos(max(kl - (s0 + p), 0));
        }
    };
    auto block = [&](int s0, const Level<T> *lv, const T *bp) {
#pragma unroll
        for (int p = 0; p < PF; ++p) step(s0 + p, lv[p], bp[p]);
    };

    // blocks alternate between the two register buffers (no copies)
    int s = 2;
    for (; s + 2 * PF - 1 <= kl; s += 2 * PF) {
        prefetch(s + PF, nxt, bnxt);
        block(s, cur, bcur);
        prefetch(s + 2 * PF, cur, bcur);
        block(s + PF, nxt, bnxt);
    }
    T lp = T(0), lt = T(0), lts = T(0);
    if (s + PF - 1 <= kl) {
        prefetch(s + PF, nxt, bnxt);
        if constexpr (FWD) {
            const uint32_t so = F.row + uint32_t(kl) * skb;
            lp = bload<T, 0>(R.pos, F.lane, so);
            lt = bload<T, AUX>(R.tens, F.lane, so);
            lts = bload<T, AUX>(R.ts, F.lane, so);
        }
        block(s, cur, bcur);
        s += PF;
#pragma unroll
        for (int p = 0; p < PF; ++p)
            if (s + p <= kl) step(s + p, nxt[p], bnxt[p]);
    } else {
        if constexpr (FWD) {
            const uint32_t so = F.row + uint32_t(kl) * skb;
            lp = bload<T, 0>(R.pos, F.lane, so);
            lt = bload<T, AUX>(R.tens, F.lane, so);
            lts = bload<T, AUX>(R.ts, F.lane, so);
        }
#pragma unroll
        for (int p = 0; p < PF; ++p)
            if (s + p <= kl) step(s + p, cur[p], bcur[p]);
    }

    // ---- step nz: forward level kl (kept in registers) -------------------
    if constexpr (FWD) {
        const C gav = C(-0.25) * (f.wn_c + f.w_c);
        const C as = gav * bet_m;
        const C acol = gav * bet_p;
        const C bcol = dtr - acol;
        const C corr = -as * (f.st_m - f.st_c);
        C dd = ((dtr * C(lp) + C(lt)) + C(lts)) + corr;
        x = (dd - f.d_prev * acol) / (bcol - f.c_prev * acol);
        last_pos = lp;
    }
}

}  // namespace
}  // namespace sb
