// Horizontal diffusion tile worker shared by the plain kernel (hdiff.hip)
// and the batched persistent kernel (batch.hip).  See hdiff.hip for the
// arithmetic and the tile schedule.
#pragma once

#include "sb_internal.h"

namespace sb {
namespace {

// Pointer into device global memory.  The tile worker takes its fields as
// such, so its loads and stores compile to global (not flat) instructions
// whatever the caller's pointers were read from.
template <typename T>
using global_ptr = __attribute__((address_space(1))) T *;

template <typename T>
__device__ __forceinline__ global_ptr<T> as_global(T *p) {
    return (global_ptr<T>)p;
}

struct HdiffGeom {
    int64_t nx, ny, nz, si, sj, sk;
};

// Whether the tile worker's 32-bit plane offsets cover a field: positive i
// and j strides, and point (nx + 1, ny + 1) counted from (-2, -2) below
// 4 GiB.
inline bool hdiff_tile_fits(int64_t nx, int64_t ny, int64_t si, int64_t sj, size_t itemsize) {
    if (si <= 0 || sj <= 0) return false;
    const double span = (double(nx + 3) * double(si) + double(ny + 3) * double(sj)) * double(itemsize);
    return span < 4294967296.0;
}

// Byte offset of point (gi, gj) within a k-plane of field element type T.
template <typename T>
__device__ __forceinline__ uint32_t plane_offset_t(int64_t gi, int64_t gj, const HdiffGeom &a) {
    return uint32_t((gi * a.si + gj * a.sj) * int64_t(sizeof(T)));
}

// `base` advanced by a non-negative byte offset.
template <typename P>
__device__ __forceinline__ P at(P base, uint32_t byte_offset) {
    using B = __attribute__((address_space(1))) const char *;
    return (P)((B)base + byte_offset);
}

template <typename C>
__device__ __forceinline__ C limited(C flux, C gradient) {
    return flux * gradient > C(0) ? C(0) : flux;
}

template <bool NT, typename T>
__device__ __forceinline__ T ld_nt(global_ptr<const T> p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// The tile worker's default hook: nothing.
struct NoTileHook {
    __device__ void operator()() const {}
};

template <bool NT, typename T>
__device__ __forceinline__ void st_nt(global_ptr<T> p, T v) {
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// Tile worker: the calling workgroup (64 x TY/RPT threads) computes the
// 64 x TY output tile at (i0, j0) for planes [k_begin, k_end).  Staging:
// each of the NW waves loads rows w, w+NW, ... of the tile's 64 main
// columns (one coalesced 64-element row segment per load), and threads
// 0..4*LH-1 load the four halo columns; out-of-range indices are clamped
// into the 2-point halo, so every load is in bounds and the clamped
// duplicates only feed outputs outside the domain, which are not stored.
// The loads for plane k+1 are issued into registers before plane k is
// computed and written to the other half of the double-buffered LDS tile
// afterwards.  FLAGS bit 0: non-temporal out stores; bit 1: non-temporal
// coeff loads (neither is re-read; inp is, by the neighbouring halos).
// `hook` runs once per call, in every thread, right after the first plane
// is staged (its loads have been waited on), before any barrier.
template <typename T, typename C, int TY, int RPT, int FLAGS>
struct HdiffTile {
    static constexpr int TX = 64;
    static constexpr int NW = TY / RPT;   // waves per workgroup
    static constexpr int LW = TX + 4;     // staged columns: tile i = -2 .. TX+1
    static constexpr int LH = TY + 4;     // staged rows:    tile j = -2 .. TY+1
    static constexpr int PW = TX + 2;     // published Laplacian columns: -1 .. TX
    static constexpr int MAIN = (LH + NW - 1) / NW;
    static constexpr int IN_ELEMS = 2 * LH * LW;  // double-buffered input tile
    static constexpr int LAP_ELEMS = TY * PW;

    // Laplacian, limited fluxes and update of the calling thread's RPT output
    // points (column tx, rows ty*RPT ..) from one staged plane `sin` (LW
    // elements of type S per row, rows -2 .. TY+1, columns -2 .. TX+1; each
    // read widened to the arithmetic type C, so a plane staged in the storage
    // type gives the same results as one staged in C).  Each thread
    // computes its column's Laplacian for rows -1..RPT in registers and
    // publishes the interior ones to s_lap for its i-neighbours; the
    // workgroup barrier inside separates those writes from the flux reads.
    // The caller must separate one call's s_lap reads from the next call's
    // writes by a barrier.  The same arithmetic as the plane loop of run(),
    // which keeps its own copy: calling this from there changes the register
    // allocation of the float32-storage kernel (126 -> 130 VGPRs, one wave
    // per SIMD fewer, 100 -> 110 us at 512x512x160).
    template <typename S>
    __device__ static void plane(const S *sin, C *s_lap, const C (&cf)[RPT], C (&res)[RPT], int tx,
                                 int ty, int tid) {
        const int col = tx + 2;         // staged column of this thread
        const int row0 = ty * RPT + 2;  // staged row of its first output
        C ic[RPT + 4], il[RPT + 2], ir[RPT + 2], lap[RPT + 2];
    #pragma unroll
        for (int q = 0; q < RPT + 4; ++q) ic[q] = C(sin[(row0 - 2 + q) * LW + col]);
    #pragma unroll
        for (int q = 0; q < RPT + 2; ++q) {
            il[q] = C(sin[(row0 - 1 + q) * LW + col - 1]);
            ir[q] = C(sin[(row0 - 1 + q) * LW + col + 1]);
        }
    #pragma unroll
        for (int q = 0; q < RPT + 2; ++q)
            lap[q] = C(4) * ic[q + 1] - (((ir[q] + il[q]) + ic[q + 2]) + ic[q]);
    #pragma unroll
        for (int q = 0; q < RPT; ++q) s_lap[(ty * RPT + q) * PW + tx + 1] = lap[q + 1];
        if (tid < 2 * TY) {
            // Laplacian of the tile's halo columns -1 and TX.
            const int side = tid / TY, r = tid - side * TY;
            const int c = side ? TX + 2 : 1;
            const int rr = r + 2;
            const C v = C(4) * C(sin[rr * LW + c]) -
                        (((C(sin[rr * LW + c + 1]) + C(sin[rr * LW + c - 1])) +
                          C(sin[(rr + 1) * LW + c])) +
                         C(sin[(rr - 1) * LW + c]));
            s_lap[r * PW + (side ? TX + 1 : 0)] = v;
        }
        __syncthreads();
    #pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = ty * RPT + q;
            const C c0 = ic[q + 2];
            const C lc = lap[q + 1];
            const C flx = limited(s_lap[r * PW + tx + 2] - lc, ir[q + 1] - c0);
            const C flx_m = limited(lc - s_lap[r * PW + tx], c0 - il[q + 1]);
            const C fly = limited(lap[q + 2] - lc, ic[q + 3] - c0);
            const C fly_m = limited(lc - lap[q], c0 - ic[q + 1]);
            res[q] = c0 - cf[q] * (((flx - flx_m) + fly) - fly_m);
        }
    }

    template <typename Hook = NoTileHook>
    __device__ static void run(global_ptr<const T> __restrict__ inp,
                               global_ptr<const T> __restrict__ coeff, global_ptr<T> __restrict__ out,
                               const HdiffGeom &a, int64_t i0, int64_t j0,
                               int64_t k_begin, int64_t k_end, C *s_in_base, C *s_lap,
                               Hook hook = Hook{}) {
        C(*s_in)[LH * LW] = reinterpret_cast<C(*)[LH * LW]>(s_in_base);
        // the wave index read as wave-uniform (every caller launches rows of
        // 64 threads): branches on it and row offsets derived from it are scalar
        const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
        const int tid = ty * TX + tx;
        // ---- per-thread staging offsets (plane-independent) -----------------
        // Byte offsets within a k-plane, 32-bit (a plane of a field below
        // 4 GiB), added to the plane's uniform base: the loads take the
        // scalar-base form with one VGPR of offset each.  Input offsets count
        // from point (-2, -2), the lowest halo point a tile reads, so none is
        // negative; coeff and out offsets count from (0, 0).
        uint32_t off_main[MAIN];
        int lds_main[MAIN];
    #pragma unroll
        for (int m = 0; m < MAIN; ++m) {
            const int r = ty + NW * m;
            const int rr = r < LH ? r : LH - 1;
            const int64_t gi = min(i0 + tx, a.nx + 1);
            const int64_t gj = min(j0 + rr - 2, a.ny + 1);
            off_main[m] = plane_offset_t<T>(gi + 2, gj + 2, a);
            lds_main[m] = r < LH ? rr * LW + tx + 2 : -1;
        }
        const bool has_halo = tid < 4 * LH;
        uint32_t off_halo = 0;
        int lds_halo = -1;
        if (has_halo) {
            const int r = tid >> 2, q = tid & 3;
            const int c = q < 2 ? q : TX + q;  // 0, 1, TX+2, TX+3
            const int64_t gi = min(i0 + c - 2, a.nx + 1);
            const int64_t gj = min(j0 + r - 2, a.ny + 1);
            off_halo = plane_offset_t<T>(gi + 2, gj + 2, a);
            lds_halo = r * LW + c;
        }
        const int64_t gi_out = i0 + tx;
        // coeff (clamped into the domain) and, where in range, out
        uint32_t off_point[RPT];
    #pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t gi = min(gi_out, a.nx - 1);
            const int64_t gj = min(j0 + ty * RPT + q, a.ny - 1);
            off_point[q] = plane_offset_t<T>(gi, gj, a);
        }

        T pm[MAIN], ph = T(0), pc[RPT];
        auto issue = [&](int64_t k) {
            global_ptr<const T> in_k = inp + (k * a.sk - 2 * a.si - 2 * a.sj);
    #pragma unroll
            for (int m = 0; m < MAIN; ++m) pm[m] = *at(in_k, off_main[m]);
            if (has_halo) ph = *at(in_k, off_halo);
            global_ptr<const T> cf_k = coeff + k * a.sk;
    #pragma unroll
            for (int q = 0; q < RPT; ++q) pc[q] = ld_nt<(FLAGS & 2) != 0>(at(cf_k, off_point[q]));
        };
        auto commit = [&](int buf) {
    #pragma unroll
            for (int m = 0; m < MAIN; ++m)
                if (lds_main[m] >= 0) s_in[buf][lds_main[m]] = C(pm[m]);
            if (has_halo) s_in[buf][lds_halo] = C(ph);
        };

        if (k_begin >= k_end) {
            hook();
            return;
        }
        issue(k_begin);
        commit(0);
        hook();
        const int col = tx + 2;         // staged column of this thread
        const int row0 = ty * RPT + 2;  // staged row of its first output
        for (int64_t k = k_begin; k < k_end; ++k) {
            const int cur = int(k - k_begin) & 1;
            C cf[RPT];
    #pragma unroll
            for (int q = 0; q < RPT; ++q) cf[q] = C(pc[q]);
            __syncthreads();
            const bool more = k + 1 < k_end;
            if (more) issue(k + 1);

            const C *sin = s_in[cur];
            C ic[RPT + 4], il[RPT + 2], ir[RPT + 2], lap[RPT + 2];
    #pragma unroll
            for (int q = 0; q < RPT + 4; ++q) ic[q] = sin[(row0 - 2 + q) * LW + col];
    #pragma unroll
            for (int q = 0; q < RPT + 2; ++q) {
                il[q] = sin[(row0 - 1 + q) * LW + col - 1];
                ir[q] = sin[(row0 - 1 + q) * LW + col + 1];
            }
    #pragma unroll
            for (int q = 0; q < RPT + 2; ++q)
                lap[q] = C(4) * ic[q + 1] - (((ir[q] + il[q]) + ic[q + 2]) + ic[q]);
    #pragma unroll
            for (int q = 0; q < RPT; ++q) s_lap[(ty * RPT + q) * PW + tx + 1] = lap[q + 1];
            if (tid < 2 * TY) {
                // Laplacian of the tile's halo columns -1 and TX.
                const int side = tid / TY, r = tid - side * TY;
                const int c = side ? TX + 2 : 1;
                const int rr = r + 2;
                const C v = C(4) * sin[rr * LW + c] -
                            (((sin[rr * LW + c + 1] + sin[rr * LW + c - 1]) + sin[(rr + 1) * LW + c]) +
                             sin[(rr - 1) * LW + c]);
                s_lap[r * PW + (side ? TX + 1 : 0)] = v;
            }
            __syncthreads();

            C res[RPT];
    #pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int r = ty * RPT + q;
                const C c0 = ic[q + 2];
                const C lc = lap[q + 1];
                const C flx = limited(s_lap[r * PW + tx + 2] - lc, ir[q + 1] - c0);
                const C flx_m = limited(lc - s_lap[r * PW + tx], c0 - il[q + 1]);
                const C fly = limited(lap[q + 2] - lc, ic[q + 3] - c0);
                const C fly_m = limited(lc - lap[q], c0 - ic[q + 1]);
                res[q] = c0 - cf[q] * (((flx - flx_m) + fly) - fly_m);
            }
            if (gi_out < a.nx) {
                global_ptr<T> out_k = out + k * a.sk;
    #pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int64_t gj = j0 + ty * RPT + q;
                    // in range, the clamped point offset is the point's own
                    if (gj < a.ny) st_nt<(FLAGS & 1) != 0>(at(out_k, off_point[q]), T(res[q]));
                }
            }
            if (more) commit(cur ^ 1);
        }
    }
};

}  // namespace
}  // namespace sb
