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
// One lane owns one column; a workgroup is one wave of 64 consecutive i at
// fixed j, so every field access of a wave is one contiguous 64-element row
// segment.  The forward sweep streams the inputs PF levels ahead of the
// recurrence (the loads do not depend on it), keeps stage/wcon of the
// previous level in registers, and stores (c', d') per level either in LDS
// (variant 1: nz x 64 x 16 bytes per wave for FP64 arithmetic) or in a
// library-owned FP64 global scratch laid out [k][column] (variant 2).  The
// backward substitution runs in the same kernel from that storage.  All
// arithmetic is in C (double for SB_F64 and SB_F32_F64); the ccol, dcol and
// datacol fields of the reference's argument list are not used.
#include "sb_internal.h"

namespace sb {
namespace {

constexpr int PF = 8;  // prefetch depth in levels

struct VadvArgs {
    int64_t nx, ny, nz, si, sj, sk;
    int32_t tiles_i;
    int32_t ncomp;
    int64_t scratch_cols;  // columns of the global scratch (variant 2)
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

template <typename C>
struct Level {
    C pos, tens, ts;  // at k
    C st, w, wn;      // stage, wcon, wcon neighbour at k+1
};

template <typename T, typename C, bool LDS>
__device__ __forceinline__ void solve_column(const Component &cp, const T *__restrict__ wcon,
                                             int64_t base, int64_t sk, int nz, CD<C> *scratch,
                                             int64_t sstride, bool active) {
    const T *__restrict__ stage = static_cast<const T *>(cp.stage);
    const T *__restrict__ pos = static_cast<const T *>(cp.pos);
    const T *__restrict__ tens = static_cast<const T *>(cp.tens);
    T *__restrict__ tstage = static_cast<T *>(cp.tensstage);
    const T *__restrict__ wn = wcon + cp.wshift;

    const C dtr = C(3.0 / 20.0);
    const C bet_m = C(0.5), bet_p = C(0.5);

    auto load = [&](int k, Level<C> &r) {
        const int64_t p = base + int64_t(k) * sk;
        r.pos = C(pos[p]);
        r.tens = C(tens[p]);
        r.ts = C(tstage[p]);
        if (k + 1 < nz) {
            r.st = C(stage[p + sk]);
            r.w = C(wcon[p + sk]);
            r.wn = C(wn[p + sk]);
        } else {
            r.st = r.w = r.wn = C(0);
        }
    };

    C st_m = C(0), st_c = C(stage[base]), w_c = C(wcon[base]), wn_c = C(wn[base]);
    C c_prev = C(0), d_prev = C(0);

    Level<C> cur[PF] = {}, nxt[PF] = {};
#pragma unroll
    for (int p = 0; p < PF; ++p)
        if (p < nz) load(p, cur[p]);

    for (int kb = 0; kb < nz; kb += PF) {
#pragma unroll
        for (int p = 0; p < PF; ++p)
            if (kb + PF + p < nz) load(kb + PF + p, nxt[p]);
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            const int k = kb + p;
            if (k >= nz) break;
            const Level<C> &r = cur[p];
            C cc, dd;
            if (k == 0) {
                const C gcv = C(0.25) * (r.wn + r.w);
                const C cs = gcv * bet_m;
                cc = gcv * bet_p;
                const C bcol = dtr - cc;
                const C corr = -cs * (r.st - st_c);
                dd = ((dtr * r.pos + r.tens) + r.ts) + corr;
                cc = cc / bcol;
                dd = dd / bcol;
            } else if (k < nz - 1) {
                const C gav = C(-0.25) * (wn_c + w_c);
                const C gcv = C(0.25) * (r.wn + r.w);
                const C as = gav * bet_m;
                const C cs = gcv * bet_m;
                const C acol = gav * bet_p;
                cc = gcv * bet_p;
                const C bcol = (dtr - acol) - cc;
                const C corr = -as * (st_m - st_c) - cs * (r.st - st_c);
                dd = ((dtr * r.pos + r.tens) + r.ts) + corr;
                const C divided = C(1.0) / (bcol - c_prev * acol);
                cc = cc * divided;
                dd = (dd - d_prev * acol) * divided;
            } else {
                const C gav = C(-0.25) * (wn_c + w_c);
                const C as = gav * bet_m;
                const C acol = gav * bet_p;
                const C bcol = dtr - acol;
                const C corr = -as * (st_m - st_c);
                dd = ((dtr * r.pos + r.tens) + r.ts) + corr;
                dd = (dd - d_prev * acol) / (bcol - c_prev * acol);
                cc = C(0);
            }
            scratch[int64_t(k) * sstride] = CD<C>{cc, dd};
            c_prev = cc;
            d_prev = dd;
            st_m = st_c;
            st_c = r.st;
            w_c = r.w;
            wn_c = r.wn;
        }
#pragma unroll
        for (int p = 0; p < PF; ++p) cur[p] = nxt[p];
    }

    // Backward substitution, PF levels of pos fetched ahead of use.
    C x = C(0);
    for (int kt = nz - 1; kt >= 0; kt -= PF) {
        C pk[PF];
#pragma unroll
        for (int p = 0; p < PF; ++p)
            pk[p] = (kt - p >= 0) ? C(pos[base + int64_t(kt - p) * sk]) : C(0);
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            const int k = kt - p;
            if (k < 0) break;
            const CD<C> cd = scratch[int64_t(k) * sstride];
            x = (k == nz - 1) ? cd.d : cd.d - cd.c * x;
            if (active) tstage[base + int64_t(k) * sk] = T(dtr * (x - pk[p]));
        }
    }
}

template <typename T, typename C, bool LDS>
__global__ __launch_bounds__(64) void vadv_kernel(Components comps, const T *__restrict__ wcon,
                                                  CD<C> *__restrict__ gscratch, VadvArgs a) {
    extern __shared__ unsigned char lds_raw[];
    const int lane = threadIdx.x;
    const int64_t tile_i = blockIdx.x % a.tiles_i;
    const int64_t j = blockIdx.x / a.tiles_i;
    const int64_t i = tile_i * 64 + lane;
    const bool active = i < a.nx;
    const int64_t ic = active ? i : a.nx - 1;  // idle lanes re-read a valid column
    const int64_t base = ic * a.si + j * a.sj;
    CD<C> *scratch;
    int64_t sstride;
    if constexpr (LDS) {
        scratch = reinterpret_cast<CD<C> *>(lds_raw) + lane;
        sstride = 64;
    } else {
        scratch = gscratch + int64_t(blockIdx.x) * 64 + lane;
        sstride = a.scratch_cols;
    }
    for (int c = 0; c < a.ncomp; ++c)
        solve_column<T, C, LDS>(comps.c[c], wcon, base, a.sk, int(a.nz), scratch, sstride,
                                active);
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

template <int P>
int launch_vadv(int all_components, void *const *f, const sb_geom *g, const sb_launch *l,
                double *seconds) {
    using T = typename Prec<P>::T;
    using C = typename Prec<P>::C;
    SB_REQUIRE(f, "NULL field array");
    const int ncomp = all_components ? 3 : 1;
    for (int q = 0; q < 4 * ncomp + 1; ++q) SB_REQUIRE(f[q], "NULL field pointer (argument %d)", q);
    Context *ctx = context();
    if (!ctx) return 1;
    if (g->nx == 0 || g->ny == 0 || g->nz == 0) {
        if (seconds) *seconds = 0.0;
        return 0;
    }
    SB_REQUIRE(g->nz >= 2, "vertical advection needs nz >= 2 (got %lld)", (long long)g->nz);
    SB_REQUIRE(g->nz <= (1 << 24), "nz too large");

    Components comps{};
    const int64_t shifts[3] = {g->si, g->sj, 0};
    for (int q = 0; q < ncomp; ++q)
        comps.c[q] = Component{f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3], shifts[q]};
    const T *wcon = static_cast<const T *>(f[4 * ncomp]);

    VadvArgs a{};
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.si = g->si;
    a.sj = g->sj;
    a.sk = g->sk;
    a.tiles_i = int32_t((g->nx + 63) / 64);
    a.ncomp = ncomp;
    const uint64_t blocks = uint64_t(a.tiles_i) * uint64_t(g->ny);
    SB_REQUIRE(blocks < (uint64_t(1) << 26), "domain too large for one launch");

    const size_t lds_bytes = size_t(g->nz) * 64 * sizeof(CD<C>);
    int variant = l ? l->variant : 0;
    if (variant == 0) variant = lds_bytes <= ctx->lds_per_block ? 1 : 2;
    SB_REQUIRE(variant == 1 || variant == 2, "vadv variant %d not available", variant);
    SB_REQUIRE(variant != 1 || lds_bytes <= ctx->lds_per_block,
               "nz = %lld needs %zu bytes of LDS per wave, device allows %zu",
               (long long)g->nz, lds_bytes, ctx->lds_per_block);

    CD<C> *gscratch = nullptr;
    if (variant == 2) {
        a.scratch_cols = int64_t(blocks) * 64;
        const size_t bytes = size_t(a.scratch_cols) * size_t(g->nz) * sizeof(CD<C>);
        if (ensure_scratch(ctx, &g_scratch[ctx->device & 63], &g_scratch_bytes[ctx->device & 63],
                           bytes))
            return 1;
        gscratch = static_cast<CD<C> *>(g_scratch[ctx->device & 63]);
    }
    if (variant == 1) {
        (void)hipFuncSetAttribute(
            reinterpret_cast<const void *>(&vadv_kernel<T, C, true>),
            hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_bytes));
    }
    return timed_launch(l, seconds, [&]() -> hipError_t {
        if (variant == 1)
            hipLaunchKernelGGL((vadv_kernel<T, C, true>), dim3(unsigned(blocks)), dim3(64),
                               lds_bytes, ctx->stream, comps, wcon, gscratch, a);
        else
            hipLaunchKernelGGL((vadv_kernel<T, C, false>), dim3(unsigned(blocks)), dim3(64), 0,
                               ctx->stream, comps, wcon, gscratch, a);
        return hipGetLastError();
    });
}

}  // namespace
}  // namespace sb

using namespace sb;

extern "C" int sb_vadv(int prec, int all_components, void *const *fields, const sb_geom *g,
                       const sb_launch *l, double *seconds) {
    if (check_geom(g, 1, all_components ? 1 : 0, 0)) return 2;
    return SB_DISPATCH_PREC(prec, launch_vadv, all_components, fields, g, l, seconds);
}
