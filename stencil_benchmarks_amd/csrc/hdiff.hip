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
// Variant 2 (LDS-DMA tile; default wherever it applies).  The same arithmetic (HdiffTile::plane) on a
// 64 x TY tile of one k-plane per workgroup, its (64+4) x (TY+4) input tile
// staged straight into LDS by global-to-LDS loads: the tile is one
// contiguous LDS array, so wave instruction n stages units 64n .. 64n+63
// of it and no VGPR holds the staged data.  A unit is a 16-byte chunk (two
// adjacent doubles of a row) for FP64 fields and one 4-byte element for
// float32 fields.  With 64 x 8 FP64 tiles and 2 rows per thread a workgroup
// needs 54 VGPRs and 11.4 KB of LDS, so eight share a CU and overlap one
// another's staging and arithmetic.  It takes whole tiles only (nx a
// multiple of 64, ny of TY) and unit i stride, and for FP64 fields even row
// and plane strides and a 16-byte aligned interior origin; other domains
// take variant 1.
//
// Variant 3 (on the fly).  One thread per point reading its 13 inputs
// through the caches; the schedule of the reference's OnTheFly body
// (stencils/cuda_hip/horizontal_diffusion.py:50-96) with the oracle's
// arithmetic order.  Kept as an independent implementation to cross-check
// variant 1.
#include <type_traits>

#include "hdiff_tile.h"

namespace sb {
namespace {

struct HdiffArgs {
    int64_t nx, ny, nz, si, sj, sk;
    int32_t tiles_i, tiles_j;
    int32_t xcds;
    int32_t pad;
};


// Variant 1 kernel.  Staging: each of the NW waves loads rows w, w+NW, ...
// of the tile's 64 main columns (one coalesced 64-element row segment per
// load), and threads 0..4*LH-1 load the four halo columns; out-of-range
// indices are clamped into the 2-point halo, so every load is in bounds and
// the clamped duplicates only feed outputs outside the domain, which are
// not stored.  A workgroup walks KC consecutive k-planes of its (i, j)
// tile: the loads for plane k+1 are issued into registers before plane k
// is computed and written to the other half of a double-buffered LDS tile
// afterwards, so each wave always has a plane's worth of loads in flight.


// FLAGS bit 0: non-temporal stores of out; bit 1: non-temporal loads of coeff
// (neither is re-read; inp is, by the neighbouring tiles' halos).
template <typename T, typename C, int TY, int RPT, int KC, int FLAGS>
__global__ __launch_bounds__(64 * (TY / RPT)) void hdiff_lds_kernel(
    const T *__restrict__ inp, const T *__restrict__ coeff, T *__restrict__ out, HdiffArgs a) {
    using Tile = HdiffTile<T, C, TY, RPT, FLAGS>;
    __shared__ C s_in[Tile::IN_ELEMS];
    __shared__ C s_lap[Tile::LAP_ELEMS];
    const uint32_t tiles = uint32_t(a.tiles_i) * uint32_t(a.tiles_j);
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x, a.xcds);
    const uint32_t kc = b / tiles;
    const uint32_t t = b - kc * tiles;
    const int64_t i0 = int64_t(t % a.tiles_i) * Tile::TX;
    const int64_t j0 = int64_t(t / a.tiles_i) * TY;
    const int64_t k_begin = int64_t(kc) * KC;
    const int64_t k_end = min(k_begin + int64_t(KC), a.nz);
    const HdiffGeom g{a.nx, a.ny, a.nz, a.si, a.sj, a.sk};
    Tile::run(as_global(inp), as_global(coeff), as_global(out), g, i0, j0, k_begin, k_end, s_in,
              s_lap);
}

// Variant 2 kernel (see the file comment).  The staged tile is one
// contiguous LDS array counted from point (i0-2, j0-2), LW elements a row;
// lane l of wave instruction n stages unit 64n + l of it.  A unit is a
// 16-byte chunk (two adjacent doubles of a row) for 8-byte fields, whose
// tile corner is then 16-byte aligned, and a 4-byte word (one element) for
// float32 fields, which need no alignment beyond the element's.  The coeff
// loads are non-temporal (not re-read), the stores plain.
template <typename T, typename C, int TY, int RPT>
__global__ __launch_bounds__(64 * (TY / RPT)) void hdiff_dma_kernel(
    const T *__restrict__ inp, const T *__restrict__ coeff, T *__restrict__ out, HdiffArgs a,
    TileGrid tg) {
    using Tile = HdiffTile<T, C, TY, RPT, 2>;
    constexpr int LW = Tile::LW, LH = Tile::LH, NW = Tile::NW;
    constexpr int UNIT = sizeof(T) == 8 ? 16 : 4;   // bytes a lane stages per instruction
    constexpr int PER = UNIT / int(sizeof(T));      // elements per unit: 2 or 1
    constexpr int UNITS = LH * LW / PER;            // units of the staged tile
    constexpr int STEPS = (UNITS + 63) / 64;        // wave instructions that stage it
    static_assert(LW % PER == 0, "a unit must not straddle two rows");
    __shared__ __attribute__((aligned(16))) T s_tile[STEPS * 64 * PER];
    __shared__ C s_lap[Tile::LAP_ELEMS];
    uint32_t ti, tj, k;
    tile_of(blockIdx.x, tg, ti, tj, k);
    const int64_t i0 = int64_t(ti) * Tile::TX;
    const int64_t j0 = int64_t(tj) * TY;
    // the wave index read as wave-uniform: the branches on it and the row
    // offsets derived from it are scalar
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int tid = ty * Tile::TX + tx;
    using GB = __attribute__((address_space(1))) const char *;
    using GP = __attribute__((address_space(1))) const void *;
    using LP = __attribute__((address_space(3))) void *;
    const GB corner =
        (GB)(inp) + ((int64_t(k) * a.sk + (j0 - 2) * a.sj + (i0 - 2)) * int64_t(sizeof(T)));
    // row pitch below 2^24 bytes (checked on the host): a staged unit's byte
    // offset from the tile corner is one 24-bit multiply and an add
    const uint32_t pitch = uint32_t(a.sj) * uint32_t(sizeof(T));
#pragma unroll
    for (int n = 0; n < STEPS; ++n) {
        if (n % NW != ty) continue;  // instruction n belongs to wave n % NW
        const int c = n * 64 + tx;
        if (c < UNITS) {
            const uint32_t row = uint32_t(c) / (LW / PER);
            const uint32_t unit = uint32_t(c) - __umul24(row, uint32_t(LW / PER));
            const GP src = (GP)(corner + (__umul24(row, pitch) + unit * uint32_t(UNIT)));
            if constexpr (UNIT == 16)
                __builtin_amdgcn_global_load_lds(src, (LP)(s_tile + n * 64 * PER), 16, 0, 0);
            else
                __builtin_amdgcn_global_load_lds(src, (LP)(s_tile + n * 64 * PER), 4, 0, 0);
        }
    }
    using GT = __attribute__((address_space(1))) const T *;
    const int64_t point = int64_t(k) * a.sk + (j0 + ty * RPT) * a.sj + i0 + tx;
    C cf[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) cf[q] = C(__builtin_nontemporal_load((GT)(coeff) + point + q * a.sj));
    __syncthreads();  // every wave's staging has landed
    C res[RPT];
    Tile::plane(s_tile, s_lap, cf, res, tx, ty, tid);
    using GO = __attribute__((address_space(1))) T *;
#pragma unroll
    for (int q = 0; q < RPT; ++q) *((GO)(out) + point + q * a.sj) = T(res[q]);
}

// Whether the variant-2 kernel with TY-row tiles takes this domain: whole
// tiles, unit i stride, a positive row pitch below 16 MiB, and for 8-byte fields
// (16-byte units) even row and plane strides and a 16-byte aligned interior
// origin.
bool hdiff_dma_fits(const sb_geom *g, const void *inp, int ty, size_t itemsize) {
    const bool tiles = g->si == 1 && g->nx % 64 == 0 && g->ny % ty == 0 && g->sj > 0 &&
                       g->sj * int64_t(itemsize) < (int64_t(1) << 24);
    if (itemsize != 8) return tiles;
    return tiles && g->sj % 2 == 0 && g->sk % 2 == 0 && reinterpret_cast<uintptr_t>(inp) % 16 == 0;
}

template <typename T, typename C, int TY, int RPT>
hipError_t launch_dma(const T *inp, const T *coeff, T *out, HdiffArgs a, Context *ctx) {
    a.tiles_i = int32_t(a.nx / 64);
    a.tiles_j = int32_t(a.ny / TY);
    const uint64_t blocks = uint64_t(a.tiles_i) * uint64_t(a.tiles_j) * uint64_t(a.nz);
    if (blocks >= (uint64_t(1) << 31)) return hipErrorInvalidValue;
    const TileGrid tg = make_tile_grid(uint32_t(blocks), uint32_t(a.tiles_i), uint32_t(a.tiles_j),
                                       ctx->xcds);
    enqueue_kernel((hdiff_dma_kernel<T, C, TY, RPT>), dim3(unsigned(blocks)), dim3(64, TY / RPT), 0,
                   ctx->stream, inp, coeff, out, a, tg);
    return hipGetLastError();
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
hipError_t launch_lds(const T *inp, const T *coeff, T *out, HdiffArgs a, Context *ctx, int flags) {
    a.tiles_i = int32_t((a.nx + 63) / 64);
    a.tiles_j = int32_t((a.ny + TY - 1) / TY);
    const uint64_t blocks =
        uint64_t(a.tiles_i) * uint64_t(a.tiles_j) * uint64_t((a.nz + KC - 1) / KC);
    switch (flags & 3) {
        case 0: enqueue_kernel((hdiff_lds_kernel<T, C, TY, RPT, KC, 0>), dim3(unsigned(blocks)),
                                   dim3(64, TY / RPT), 0, ctx->stream, inp, coeff, out, a); break;
        case 1: enqueue_kernel((hdiff_lds_kernel<T, C, TY, RPT, KC, 1>), dim3(unsigned(blocks)),
                                   dim3(64, TY / RPT), 0, ctx->stream, inp, coeff, out, a); break;
        case 2: enqueue_kernel((hdiff_lds_kernel<T, C, TY, RPT, KC, 2>), dim3(unsigned(blocks)),
                                   dim3(64, TY / RPT), 0, ctx->stream, inp, coeff, out, a); break;
        default: enqueue_kernel((hdiff_lds_kernel<T, C, TY, RPT, KC, 3>), dim3(unsigned(blocks)),
                                    dim3(64, TY / RPT), 0, ctx->stream, inp, coeff, out, a); break;
    }
    return hipGetLastError();
}

template <int TY, int RPT, typename T, typename C>
hipError_t dispatch_kc_ty(int kc, const T *inp, const T *coeff, T *out, HdiffArgs a,
                          Context *ctx, int flags) {
    switch (kc) {
        case 1: return launch_lds<TY, RPT, 1, T, C>(inp, coeff, out, a, ctx, flags);
        case 2: return launch_lds<TY, RPT, 2, T, C>(inp, coeff, out, a, ctx, flags);
        case 4: return launch_lds<TY, RPT, 4, T, C>(inp, coeff, out, a, ctx, flags);
        case 8: return launch_lds<TY, RPT, 8, T, C>(inp, coeff, out, a, ctx, flags);
        case 16: return launch_lds<TY, RPT, 16, T, C>(inp, coeff, out, a, ctx, flags);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, typename C>
hipError_t dispatch_kc(int ty, int kc, const T *inp, const T *coeff, T *out, HdiffArgs a,
                       Context *ctx, int flags) {
    switch (ty) {
        case 8: return dispatch_kc_ty<8, 2, T, C>(kc, inp, coeff, out, a, ctx, flags);
        case 16: return dispatch_kc_ty<16, 4, T, C>(kc, inp, coeff, out, a, ctx, flags);
        case 32: return dispatch_kc_ty<32, 8, T, C>(kc, inp, coeff, out, a, ctx, flags);
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
    // the LDS tile (variant 1, default) addresses a plane with 32-bit offsets
    const bool tile_fits = hdiff_tile_fits(g->nx, g->ny, g->si, g->sj, sizeof(T));
    // LDS-DMA tile rows when none are given: 8 for 8-byte fields, 16 for
    // float32 ones (512 x 512 x 160, FP32: 80.7 us with 16 rows, 105 us
    // with 8; float32 storage with FP64 arithmetic 86.0 / 116.4 us)
    constexpr int dma_rows = sizeof(T) == 8 ? 8 : 16;
    // tile options (rows, planes, cache hints) select the register-staged tile
    const bool tile_options = l && (l->by || l->bz || l->unroll);
    int variant = l ? l->variant : 0;
    if (variant == 0) {
        // the LDS-DMA tile where it applies (measured on MI355X at 512 x 512
        // x 160 against variant 1's defaults: FP64 156.5-157.3 us against
        // 168.7 us, FP32 80.7 us against 93.0-93.4 us, float32 storage with
        // FP64 arithmetic 86.0 us against 97.5-99.2 us)
        if (!tile_options && hdiff_dma_fits(g, inp, dma_rows, sizeof(T)))
            variant = 2;
        else
            variant = tile_fits ? 1 : 3;
    }
    if (variant == 2) {
        const int ty2 = (l && l->by) ? l->by : dma_rows;
        SB_REQUIRE(ty2 == 8 || ty2 == 16, "LDS-DMA hdiff tile height %d not in {8, 16}", ty2);
        SB_REQUIRE(!l || (l->bz <= 1 && l->unroll == 0),
                   "the LDS-DMA hdiff tile takes one k-plane per workgroup and no cache-hint options");
        SB_REQUIRE(hdiff_dma_fits(g, inp, ty2, sizeof(T)),
                   "the LDS-DMA hdiff tile needs nx a multiple of 64, ny a multiple of %d, unit i "
                   "stride, a positive row pitch below 16 MiB, and for FP64 fields even j and k "
                   "strides "
                   "and a 16-byte aligned interior origin",
                   ty2);
        return timed_launch(l, seconds, [&]() -> hipError_t {
            return ty2 == 8 ? launch_dma<T, C, 8, 2>(inp, coeff, out, a, ctx)
                            : launch_dma<T, C, 16, 4>(inp, coeff, out, a, ctx);
        });
    }
    SB_REQUIRE(variant != 1 || tile_fits,
               "the LDS-tiled hdiff needs positive i and j strides and a k-plane below 4 GiB");
    const int ty = (l && l->by) ? l->by : 16;
    // defaults measured on MI355X (512 x 512 x 160): 64 x 16 tiles, 2 planes
    // per workgroup in FP64 arithmetic, 8 with FP32 arithmetic
    const int kc = (l && l->bz) ? l->bz : (sizeof(C) == 8 && sizeof(T) == 8 ? 2 : 8);
    // launch->unroll: 0 = default (non-temporal coeff loads, measured best on
    // MI355X), else 4 | flags with bit 0 non-temporal out stores and bit 1
    // non-temporal coeff loads
    const int flags = (l && l->unroll) ? (l->unroll & 3) : 2;
    SB_REQUIRE(variant == 1 || variant == 3, "hdiff variant %d not available", variant);
    if (variant == 3) {
        SB_REQUIRE(g->nz <= 65535, "on-the-fly hdiff supports nz <= 65535");
        const dim3 grid(unsigned((g->nx + 63) / 64), unsigned((g->ny + 3) / 4), unsigned(g->nz));
        return timed_launch(l, seconds, [&]() -> hipError_t {
            enqueue_kernel((hdiff_otf_kernel<T, C>), grid, dim3(64, 4), 0, ctx->stream, inp,
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
        return dispatch_kc<T, C>(ty, kc, inp, coeff, out, a, ctx, flags);
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
