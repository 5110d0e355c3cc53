// Horizontal diffusion: 4th-order Laplacian -> limited fluxes -> update.
//
// Arithmetic (operation order of stencils/base.py:284-307, which the FP64
// path reproduces bit for bit; compile with -ffp-contract=off):
//   lap(i,j) = 4 in(i,j) - (((in(i+1,j) + in(i-1,j)) + in(i,j+1)) + in(i,j-1))
//   flx(i,j) = lap(i+1,j) - lap(i,j), set to 0 if flx * (in(i+1,j) - in(i,j)) > 0
//   fly(i,j) = lap(i,j+1) - lap(i,j), set to 0 if fly * (in(i,j+1) - in(i,j)) > 0
//   out      = in - coeff * (((flx(i,j) - flx(i-1,j)) + fly(i,j)) - fly(i,j-1))
//
// Variant 1 (LDS tile, default).  A workgroup of 64 x (TY/RPT) threads owns
// a 64 x TY output tile of one k-plane.  It stages the (64+4) x (TY+4) input
// tile (2-point halo on every side) in LDS with coalesced loads, each
// thread then walks RPT consecutive rows of one column: it computes the
// Laplacian of its column for rows -1..RPT in registers, publishes its RPT
// interior Laplacians to an LDS tile (the i-neighbours need them), and
// after one barrier forms the four limited fluxes and the update from
// registers plus two LDS reads.  Input, Laplacian and fluxes never touch
// HBM; each output point costs one inp read (plus the halo share), one
// coeff read and one out write.
//
// Variant 3 (on the fly).  One thread per point reading its 13 inputs
// through the caches; the schedule of the reference's OnTheFly body
// (stencils/cuda_hip/horizontal_diffusion.py:50-96) with the oracle's
// arithmetic order.  Kept as an independent implementation to cross-check
// variant 1.
#include "sb_internal.h"

namespace sb {
namespace {

struct HdiffArgs {
    int64_t nx, ny, nz, si, sj, sk;
    int32_t tiles_i, tiles_j;
    int32_t xcds;
    int32_t pad;
};

template <typename C>
__device__ __forceinline__ C limited(C flux, C gradient) {
    return flux * gradient > C(0) ? C(0) : flux;
}

// Variant 1 kernel.  Staging: each of the NW waves loads rows w, w+NW, ...
// of the tile's 64 main columns (one coalesced 64-element row segment per
// load), and threads 0..4*LH-1 load the four halo columns; out-of-range
// indices are clamped into the 2-point halo, so every load is in bounds and
// the clamped duplicates only feed outputs outside the domain, which are
// not stored.  A workgroup walks KC consecutive k-planes of its (i, j)
// tile: the loads for plane k+1 are issued into registers before plane k
// is computed and written to the other half of a double-buffered LDS tile
// afterwards, so each wave always has a plane's worth of loads in flight.
template <typename T, typename C, int TY, int RPT, int KC>
__global__ __launch_bounds__(64 * (TY / RPT)) void hdiff_lds_kernel(
    const T *__restrict__ inp, const T *__restrict__ coeff, T *__restrict__ out, HdiffArgs a) {
    constexpr int TX = 64;
    constexpr int NW = TY / RPT;          // waves per workgroup
    constexpr int LW = TX + 4;            // staged columns: tile i = -2 .. TX+1
    constexpr int LH = TY + 4;            // staged rows:    tile j = -2 .. TY+1
    constexpr int PW = TX + 2;            // published Laplacian columns: -1 .. TX
    constexpr int MAIN = (LH + NW - 1) / NW;
    __shared__ C s_in[2][LH * LW];
    __shared__ C s_lap[TY * PW];

    const uint32_t tiles = uint32_t(a.tiles_i) * uint32_t(a.tiles_j);
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x, a.xcds);
    const uint32_t kc = b / tiles;
    const uint32_t t = b - kc * tiles;
    const int64_t i0 = int64_t(t % a.tiles_i) * TX;
    const int64_t j0 = int64_t(t / a.tiles_i) * TY;
    const int64_t k_begin = int64_t(kc) * KC;
    const int64_t k_end = min(k_begin + int64_t(KC), a.nz);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = ty * TX + tx;

    // ---- per-thread staging offsets (plane-independent) -----------------
    int64_t off_main[MAIN];
    int lds_main[MAIN];
#pragma unroll
    for (int m = 0; m < MAIN; ++m) {
        const int r = ty + NW * m;
        const int rr = r < LH ? r : LH - 1;
        const int64_t gi = min(i0 + tx, a.nx + 1);
        const int64_t gj = min(j0 + rr - 2, a.ny + 1);
        off_main[m] = gi * a.si + gj * a.sj;
        lds_main[m] = r < LH ? rr * LW + tx + 2 : -1;
    }
    const bool has_halo = tid < 4 * LH;
    int64_t off_halo = 0;
    int lds_halo = -1;
    if (has_halo) {
        const int r = tid >> 2, q = tid & 3;
        const int c = q < 2 ? q : TX + q;  // 0, 1, TX+2, TX+3
        const int64_t gi = min(i0 + c - 2, a.nx + 1);
        const int64_t gj = min(j0 + r - 2, a.ny + 1);
        off_halo = gi * a.si + gj * a.sj;
        lds_halo = r * LW + c;
    }
    const int64_t gi_out = i0 + tx;
    int64_t off_coeff[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int64_t gi = min(gi_out, a.nx - 1);
        const int64_t gj = min(j0 + ty * RPT + q, a.ny - 1);
        off_coeff[q] = gi * a.si + gj * a.sj;
    }

    T pm[MAIN], ph = T(0), pc[RPT];
    auto issue = [&](int64_t k) {
        const T *in_k = inp + k * a.sk;
#pragma unroll
        for (int m = 0; m < MAIN; ++m) pm[m] = in_k[off_main[m]];
        if (has_halo) ph = in_k[off_halo];
        const T *cf_k = coeff + k * a.sk;
#pragma unroll
        for (int q = 0; q < RPT; ++q) pc[q] = cf_k[off_coeff[q]];
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int m = 0; m < MAIN; ++m)
            if (lds_main[m] >= 0) s_in[buf][lds_main[m]] = C(pm[m]);
        if (has_halo) s_in[buf][lds_halo] = C(ph);
    };

    if (k_begin >= k_end) return;
    issue(k_begin);
    commit(0);
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
            T *out_k = out + k * a.sk + gi_out * a.si;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int64_t gj = j0 + ty * RPT + q;
                if (gj < a.ny) out_k[gj * a.sj] = T(res[q]);
            }
        }
        if (more) commit(cur ^ 1);
    }
}

template <typename T, typename C>
__global__ __launch_bounds__(256) void hdiff_otf_kernel(const T *__restrict__ inp,
                                                        const T *__restrict__ coeff,
                                                        T *__restrict__ out, HdiffArgs a) {
    const int64_t i = int64_t(blockIdx.x) * 64 + threadIdx.x;
    const int64_t j = int64_t(blockIdx.y) * 4 + threadIdx.y;
    const int64_t k = blockIdx.z;
    if (i >= a.nx || j >= a.ny) return;
    const int64_t x = a.si, y = a.sj;
    const int64_t p = i * a.si + j * a.sj + k * a.sk;
    auto in = [&](int64_t off) { return C(inp[p + off]); };
    const C c = in(0), im = in(-x), ip = in(x), jm = in(-y), jp = in(y);
    const C lap_c = C(4) * c - (((ip + im) + jp) + jm);
    const C lap_im = C(4) * im - (((c + in(-2 * x)) + in(-x + y)) + in(-x - y));
    const C lap_ip = C(4) * ip - (((in(2 * x) + c) + in(x + y)) + in(x - y));
    const C lap_jm = C(4) * jm - (((in(x - y) + in(-x - y)) + c) + in(-2 * y));
    const C lap_jp = C(4) * jp - (((in(x + y) + in(-x + y)) + in(2 * y)) + c);
    const C flx = limited(lap_ip - lap_c, ip - c);
    const C flx_m = limited(lap_c - lap_im, c - im);
    const C fly = limited(lap_jp - lap_c, jp - c);
    const C fly_m = limited(lap_c - lap_jm, c - jm);
    out[p] = T(c - C(coeff[p]) * (((flx - flx_m) + fly) - fly_m));
}

template <int TY, int RPT, int KC, typename T, typename C>
hipError_t launch_lds(const T *inp, const T *coeff, T *out, HdiffArgs a, Context *ctx) {
    a.tiles_i = int32_t((a.nx + 63) / 64);
    a.tiles_j = int32_t((a.ny + TY - 1) / TY);
    const uint64_t blocks =
        uint64_t(a.tiles_i) * uint64_t(a.tiles_j) * uint64_t((a.nz + KC - 1) / KC);
    hipLaunchKernelGGL((hdiff_lds_kernel<T, C, TY, RPT, KC>), dim3(unsigned(blocks)),
                       dim3(64, TY / RPT), 0, ctx->stream, inp, coeff, out, a);
    return hipGetLastError();
}

template <int TY, int RPT, typename T, typename C>
hipError_t dispatch_kc_ty(int kc, const T *inp, const T *coeff, T *out, HdiffArgs a,
                          Context *ctx) {
    switch (kc) {
        case 1: return launch_lds<TY, RPT, 1, T, C>(inp, coeff, out, a, ctx);
        case 2: return launch_lds<TY, RPT, 2, T, C>(inp, coeff, out, a, ctx);
        case 4: return launch_lds<TY, RPT, 4, T, C>(inp, coeff, out, a, ctx);
        case 8: return launch_lds<TY, RPT, 8, T, C>(inp, coeff, out, a, ctx);
        case 16: return launch_lds<TY, RPT, 16, T, C>(inp, coeff, out, a, ctx);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, typename C>
hipError_t dispatch_kc(int ty, int kc, const T *inp, const T *coeff, T *out, HdiffArgs a,
                       Context *ctx) {
    switch (ty) {
        case 8: return dispatch_kc_ty<8, 2, T, C>(kc, inp, coeff, out, a, ctx);
        case 16: return dispatch_kc_ty<16, 4, T, C>(kc, inp, coeff, out, a, ctx);
        case 32: return dispatch_kc_ty<32, 8, T, C>(kc, inp, coeff, out, a, ctx);
        default: return hipErrorInvalidValue;
    }
}

template <int P>
int launch_hdiff(const void *inp_v, const void *coeff_v, void *out_v, const sb_geom *g,
                 const sb_launch *l, double *seconds) {
    using T = typename Prec<P>::T;
    using C = typename Prec<P>::C;
    const T *inp = static_cast<const T *>(inp_v);
    const T *coeff = static_cast<const T *>(coeff_v);
    T *out = static_cast<T *>(out_v);
    SB_REQUIRE(inp && coeff && out, "NULL field pointer");
    Context *ctx = context();
    if (!ctx) return 1;
    if (g->nx == 0 || g->ny == 0 || g->nz == 0) {
        if (seconds) *seconds = 0.0;
        return 0;
    }
    HdiffArgs a{g->nx, g->ny, g->nz, g->si, g->sj, g->sk, 0, 0, ctx->xcds, 0};
    const int variant = l ? l->variant : 0;
    const int ty = (l && l->by) ? l->by : 16;
    const int kc = (l && l->bz) ? l->bz : 8;
    SB_REQUIRE(variant == 0 || variant == 1 || variant == 3, "hdiff variant %d not available",
               variant);
    if (variant == 3) {
        SB_REQUIRE(g->nz <= 65535, "on-the-fly hdiff supports nz <= 65535");
        const dim3 grid(unsigned((g->nx + 63) / 64), unsigned((g->ny + 3) / 4), unsigned(g->nz));
        return timed_launch(l, seconds, [&]() -> hipError_t {
            hipLaunchKernelGGL((hdiff_otf_kernel<T, C>), grid, dim3(64, 4), 0, ctx->stream, inp,
                               coeff, out, a);
            return hipGetLastError();
        });
    }
    SB_REQUIRE(kc == 1 || kc == 2 || kc == 4 || kc == 8 || kc == 16,
               "hdiff planes per workgroup %d not in {1, 2, 4, 8, 16}", kc);
    const uint64_t blocks = uint64_t((g->nx + 63) / 64) * uint64_t((g->ny + ty - 1) / ty) *
                            uint64_t((g->nz + kc - 1) / kc);
    SB_REQUIRE(blocks < (uint64_t(1) << 31), "domain too large for one launch");
    return timed_launch(l, seconds, [&]() -> hipError_t {
        return dispatch_kc<T, C>(ty, kc, inp, coeff, out, a, ctx);
    });
}

}  // namespace
}  // namespace sb

using namespace sb;

extern "C" int sb_hdiff(int prec, const void *inp, const void *coeff, void *out,
                        const sb_geom *g, const sb_launch *l, double *seconds) {
    if (check_geom(g, 2, 2, 0)) return 2;
    if (l && l->by) {
        SB_REQUIRE(l->by == 8 || l->by == 16 || l->by == 32, "hdiff tile height %d not in {8, 16, 32}",
                   l->by);
    }
    return SB_DISPATCH_PREC(prec, launch_hdiff, inp, coeff, out, g, l, seconds);
}
