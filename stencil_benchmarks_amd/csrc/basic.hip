// Basic stencils over the interior: copy, one-sided and symmetric average,
// Laplacian.
//
// The host permutes the three logical axes so that kernel axis 0 is the one
// with unit stride.  Each thread computes V consecutive points along it:
// the centre and the off-axis neighbours are V-wide aligned vector loads
// (one dwordx4 for V*sizeof(T) == 16), the two along-axis neighbours that
// fall outside the vector are scalar loads, the rest are register shuffles.
// A workgroup sweeps R consecutive rows (points with equal axis-1/axis-2
// index; launch->by = 4, 8 or 16, default 4) so every thread keeps several
// rows' loads in flight, and row blocks are remapped so that consecutive
// blocks run on the same XCD: the off-axis neighbour rows a block shares with
// the next one then hit that XCD's L2 instead of being fetched again.
//
// Arithmetic follows the verification oracles of stencils/base.py:
//   onesided  (inp[x+e] + inp[x]) / 2                       (base.py:204)
//   symmetric (inp[x+e] + inp[x-e]) / 2                     (base.py:221)
//   laplacian r = 0; r += (2 inp - inp[x+e]) - inp[x-e] per enabled axis in
//             x, y, z order                                 (base.py:243-253)
#include <algorithm>
#include <type_traits>

#include "sb_internal.h"

namespace sb {
namespace {

enum { OP_COPY = 0, OP_AVG1 = 1, OP_AVG2 = 2, OP_LAP = 3 };

struct BasicArgs {
    int64_t n[3];       // extents of the permuted axes
    int64_t s[3];       // element strides of the permuted axes (s[0] == 1 if V > 1)
    int32_t axis;       // permuted averaging axis
    int32_t lap_mask;   // bit o: Laplacian along ORIGINAL axis o
    int32_t perm[3];    // permuted index of original axis o
    int32_t xcds;       // XCDs of the device (row-block remap)
};

constexpr int BLOCK = 256;

template <typename T, int V>
struct alignas(sizeof(T) * V) Vec {
    T e[V];
};

template <typename T, typename C, int OP, int V>
__device__ __forceinline__ C scalar_point(const T *__restrict__ inp, int64_t p,
                                          const BasicArgs &a) {
    if constexpr (OP == OP_COPY) {
        return C(inp[p]);
    } else if constexpr (OP == OP_AVG1) {
        const int64_t s = a.s[a.axis];
        return (C(inp[p + s]) + C(inp[p])) / C(2);
    } else if constexpr (OP == OP_AVG2) {
        const int64_t s = a.s[a.axis];
        return (C(inp[p + s]) + C(inp[p - s])) / C(2);
    } else {
        const C c = C(inp[p]);
        C r = C(0);
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            if (a.lap_mask & (1 << o)) {
                const int64_t s = a.s[a.perm[o]];
                r += (C(2) * c - C(inp[p + s])) - C(inp[p - s]);
            }
        }
        return r;
    }
}

template <typename T, typename C, int OP, int V>
__device__ __forceinline__ void vector_points(const T *__restrict__ inp, T *__restrict__ out,
                                              int64_t p, const BasicArgs &a) {
    using VT = Vec<T, V>;
    const VT cv = *reinterpret_cast<const VT *>(inp + p);
    VT res;
    if constexpr (OP == OP_COPY) {
        res = cv;
    } else if constexpr (OP == OP_AVG1 || OP == OP_AVG2) {
        C plus[V], minus[V];
        if (a.axis == 0) {
            const C right = C(inp[p + V]);
            C left = C(0);
            if constexpr (OP == OP_AVG2) left = C(inp[p - 1]);
#pragma unroll
            for (int e = 0; e < V; ++e) {
                plus[e] = e + 1 < V ? C(cv.e[e + 1]) : right;
                minus[e] = e > 0 ? C(cv.e[e - 1]) : left;
            }
        } else {
            const int64_t s = a.s[a.axis];
            const VT pv = *reinterpret_cast<const VT *>(inp + p + s);
#pragma unroll
            for (int e = 0; e < V; ++e) plus[e] = C(pv.e[e]);
            if constexpr (OP == OP_AVG2) {
                const VT mv = *reinterpret_cast<const VT *>(inp + p - s);
#pragma unroll
                for (int e = 0; e < V; ++e) minus[e] = C(mv.e[e]);
            }
        }
#pragma unroll
        for (int e = 0; e < V; ++e) {
            if constexpr (OP == OP_AVG1)
                res.e[e] = T((plus[e] + C(cv.e[e])) / C(2));
            else
                res.e[e] = T((plus[e] + minus[e]) / C(2));
        }
    } else {
        C r[V];
#pragma unroll
        for (int e = 0; e < V; ++e) r[e] = C(0);
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            if (!(a.lap_mask & (1 << o))) continue;
            const int q = a.perm[o];
            C plus[V], minus[V];
            if (q == 0) {
                const C right = C(inp[p + V]);
                const C left = C(inp[p - 1]);
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    plus[e] = e + 1 < V ? C(cv.e[e + 1]) : right;
                    minus[e] = e > 0 ? C(cv.e[e - 1]) : left;
                }
            } else {
                const int64_t s = a.s[q];
                const VT pv = *reinterpret_cast<const VT *>(inp + p + s);
                const VT mv = *reinterpret_cast<const VT *>(inp + p - s);
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    plus[e] = C(pv.e[e]);
                    minus[e] = C(mv.e[e]);
                }
            }
#pragma unroll
            for (int e = 0; e < V; ++e) r[e] += (C(2) * C(cv.e[e]) - plus[e]) - minus[e];
        }
#pragma unroll
        for (int e = 0; e < V; ++e) res.e[e] = T(r[e]);
    }
    *reinterpret_cast<VT *>(out + p) = res;
}

template <typename T, typename C, int OP, int V, int ROWS>
__global__ __launch_bounds__(BLOCK) void basic_kernel(const T *__restrict__ inp,
                                                      T *__restrict__ out, BasicArgs a) {
    const int64_t i = (int64_t(blockIdx.y) * BLOCK + threadIdx.x) * V;
    if (i >= a.n[0]) return;
    const int64_t rows = a.n[1] * a.n[2];
    const bool full = (i + V <= a.n[0]);
    const uint32_t rb = xcd_remap(blockIdx.x, gridDim.x, a.xcds);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int64_t row = int64_t(rb) * ROWS + r;
        if (row >= rows) break;
        const int64_t j = row % a.n[1];
        const int64_t k = row / a.n[1];
        const int64_t p = i * a.s[0] + j * a.s[1] + k * a.s[2];
        if (V > 1 && full) {
            vector_points<T, C, OP, V>(inp, out, p, a);
        } else {
            const int64_t last = a.n[0] - i < V ? a.n[0] - i : V;
            for (int64_t e = 0; e < last; ++e)
                out[p + e * a.s[0]] = T(scalar_point<T, C, OP, V>(inp, p + e * a.s[0], a));
        }
    }
}

// Marching variant for stencils whose terms lie along permuted axes 0 and 1
// only: a thread owns V consecutive points of axis 0 and walks JB rows of
// axis 1 (one plane of axis 2), keeping rows j-1, j, j+1 in registers, so
// every input row is loaded once per thread (plus the two halo rows of its
// run) instead of once per neighbouring output row; the two axis-0
// neighbours outside the vector are scalar loads of the adjacent lanes'
// lines.  Terms are summed in original-axis order, as the row kernel does.
template <typename T, typename C, int OP, int V, int JB>
__global__ __launch_bounds__(BLOCK) void march_kernel(const T *__restrict__ inp,
                                                      T *__restrict__ out, BasicArgs a) {
    using VT = Vec<T, V>;
    const int64_t i = (int64_t(blockIdx.y) * blockDim.x + threadIdx.x) * V;
    if (i >= a.n[0]) return;
    const int64_t chunks = (a.n[1] + JB - 1) / JB;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x, a.xcds);
    const int64_t j0 = int64_t(b % uint32_t(chunks)) * JB;
    const int64_t k = int64_t(b / uint32_t(chunks));
    const int64_t jn = a.n[1] - j0 < JB ? a.n[1] - j0 : JB;
    const int64_t s1 = a.s[1];
    int64_t p = i * a.s[0] + j0 * s1 + k * a.s[2];
    if (i + V > a.n[0]) {  // ragged end of axis 0: scalar points
        const int64_t last = a.n[0] - i;
        for (int64_t jj = 0; jj < jn; ++jj, p += s1)
            for (int64_t e = 0; e < last; ++e)
                out[p + e * a.s[0]] = T(scalar_point<T, C, OP, V>(inp, p + e * a.s[0], a));
        return;
    }
    // axis-1 neighbours needed: below (j-1) and above (j+1)
    bool below = false, above = false;
    if constexpr (OP == OP_LAP) {
        below = above = false;
#pragma unroll
        for (int o = 0; o < 3; ++o)
            if ((a.lap_mask & (1 << o)) && a.perm[o] == 1) below = above = true;
    } else {
        above = a.axis == 1;
        below = OP == OP_AVG2 && a.axis == 1;
    }
    VT prev{}, cur, next{};
    if (below) prev = *reinterpret_cast<const VT *>(inp + p - s1);
    cur = *reinterpret_cast<const VT *>(inp + p);
#pragma unroll 4
    for (int64_t jj = 0; jj < jn; ++jj, p += s1) {
        if (above) next = *reinterpret_cast<const VT *>(inp + p + s1);
        VT res;
        if constexpr (OP == OP_LAP) {
            C r[V];
#pragma unroll
            for (int e = 0; e < V; ++e) r[e] = C(0);
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                if (!(a.lap_mask & (1 << o))) continue;
                C plus[V], minus[V];
                if (a.perm[o] == 0) {
                    const C right = C(inp[p + V]);
                    const C left = C(inp[p - 1]);
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        plus[e] = e + 1 < V ? C(cur.e[e + 1]) : right;
                        minus[e] = e > 0 ? C(cur.e[e - 1]) : left;
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        plus[e] = C(next.e[e]);
                        minus[e] = C(prev.e[e]);
                    }
                }
#pragma unroll
                for (int e = 0; e < V; ++e) r[e] += (C(2) * C(cur.e[e]) - plus[e]) - minus[e];
            }
#pragma unroll
            for (int e = 0; e < V; ++e) res.e[e] = T(r[e]);
        } else {
            C plus[V], minus[V];
            if (a.axis == 0) {
                const C right = C(inp[p + V]);
                C left = C(0);
                if constexpr (OP == OP_AVG2) left = C(inp[p - 1]);
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    plus[e] = e + 1 < V ? C(cur.e[e + 1]) : right;
                    minus[e] = e > 0 ? C(cur.e[e - 1]) : left;
                }
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    plus[e] = C(next.e[e]);
                    minus[e] = C(prev.e[e]);
                }
            }
#pragma unroll
            for (int e = 0; e < V; ++e) {
                if constexpr (OP == OP_AVG1)
                    res.e[e] = T((plus[e] + C(cur.e[e])) / C(2));
                else
                    res.e[e] = T((plus[e] + minus[e]) / C(2));
            }
        }
        *reinterpret_cast<VT *>(out + p) = res;
        prev = cur;
        cur = next;
        if (!above && jj + 1 < jn) cur = *reinterpret_cast<const VT *>(inp + p + s1);
    }
}

// Laplacian along i and j in FP64 (variant 3, default where it applies): a
// 64 x TY tile of one k-plane per workgroup, its (64+2) x (TY+2) input tile
// staged straight into LDS by 4-byte global-to-LDS loads (one contiguous LDS
// array; wave instruction n stages dwords 64n .. 64n+63 of it, so the tile's
// 8-byte aligned corner at (i0-1, j0-1) needs no 16-byte alignment and no
// second halo column), then each thread forms RPT points of one column:
// r = 0; r += (2 c - in(i+1)) - in(i-1); r += (2 c - in(j+1)) - in(j-1).
// One plane per workgroup and a few VGPRs: many workgroups per CU overlap one
// another's staging, and the grid has many rounds, so no tail of idle CUs.
struct LapDmaArgs {
    int64_t sj, sk;
    int32_t tiles_i, tiles_j, xcds, pad;
};

template <int TY, int RPT>
__global__ __launch_bounds__(64 * (TY / RPT)) void lap_dma_kernel(const double *__restrict__ inp,
                                                                  double *__restrict__ out,
                                                                  LapDmaArgs a, TileGrid tg) {
    constexpr int TX = 64, NW = TY / RPT, LW = TX + 2, LH = TY + 2;
    constexpr int DWORDS = LH * LW * 2;          // staged tile in 4-byte words
    constexpr int STEPS = (DWORDS + 63) / 64;    // wave instructions that stage it
    __shared__ __attribute__((aligned(16))) double s_tile[STEPS * 32];
    uint32_t ti, tj, k;
    tile_of(blockIdx.x, tg, ti, tj, k);
    const int64_t i0 = int64_t(ti) * TX;
    const int64_t j0 = int64_t(tj) * TY;
    // the wave index read as wave-uniform: the branches on it and the row
    // offsets derived from it are scalar
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
    using GB = __attribute__((address_space(1))) const char *;
    using GP = __attribute__((address_space(1))) const void *;
    using LP = __attribute__((address_space(3))) void *;
    const GB corner = (GB)(inp) + ((int64_t(k) * a.sk + (j0 - 1) * a.sj + (i0 - 1)) * 8);
    // row pitch below 2^24 bytes (checked on the host)
    const uint32_t pitch = uint32_t(a.sj) * 8u;
#pragma unroll
    for (int n = 0; n < STEPS; ++n) {
        if (n % NW != ty) continue;  // instruction n belongs to wave n % NW
        const int d = n * 64 + tx;
        if (d < DWORDS) {
            const uint32_t row = uint32_t(d) / (2 * LW);
            const uint32_t word = uint32_t(d) - __umul24(row, uint32_t(2 * LW));
            __builtin_amdgcn_global_load_lds((GP)(corner + (__umul24(row, pitch) + word * 4u)),
                                             (LP)(s_tile + n * 32), 4, 0, 0);
        }
    }
    __syncthreads();  // every wave's staging has landed
    using GO = __attribute__((address_space(1))) double *;
    const int64_t point = int64_t(k) * a.sk + (j0 + ty * RPT) * a.sj + i0 + tx;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int at = (ty * RPT + q + 1) * LW + tx + 1;
        const double c = s_tile[at];
        double r = 0.0;
        r += (2.0 * c - s_tile[at + 1]) - s_tile[at - 1];
        r += (2.0 * c - s_tile[at + LW]) - s_tile[at - LW];
        *((GO)(out) + point + q * a.sj) = r;
    }
}

// Whether the variant-3 Laplacian with TY-row tiles takes this call.
bool lap_dma_fits(const sb_geom *g, int lap_mask, int ty) {
    return lap_mask == 3 && g->si == 1 && g->nx % 64 == 0 && g->ny % ty == 0 && g->nz > 0 &&
           g->sj > 0 && g->sj * 8 < (int64_t(1) << 24);
}

template <int TY, int RPT>
hipError_t launch_lap_dma(const double *inp, double *out, const sb_geom *g, Context *ctx) {
    LapDmaArgs a{g->sj, g->sk, int32_t(g->nx / 64), int32_t(g->ny / TY), ctx->xcds, 0};
    const uint64_t blocks = uint64_t(a.tiles_i) * uint64_t(a.tiles_j) * uint64_t(g->nz);
    if (blocks >= (uint64_t(1) << 31)) return hipErrorInvalidValue;
    const TileGrid tg = make_tile_grid(uint32_t(blocks), uint32_t(a.tiles_i), uint32_t(a.tiles_j),
                                       ctx->xcds);
    enqueue_kernel((lap_dma_kernel<TY, RPT>), dim3(unsigned(blocks)), dim3(64, TY / RPT), 0,
                   ctx->stream, inp, out, a, tg);
    return hipGetLastError();
}

template <int P, int OP>
int launch_basic_prec(const void *inp_v, void *out_v, const sb_geom *g, int axis, int lap_mask,
                      const sb_launch *l, double *seconds) {
    using T = typename Prec<P>::T;
    using C = typename Prec<P>::C;
    const T *inp = static_cast<const T *>(inp_v);
    T *out = static_cast<T *>(out_v);
    SB_REQUIRE(inp && out, "NULL field pointer");

    // Permute axes: unit (smallest |stride|) axis first, others by stride.
    const int64_t n_orig[3] = {g->nx, g->ny, g->nz};
    const int64_t s_orig[3] = {g->si, g->sj, g->sk};
    int order[3] = {0, 1, 2};
    for (int x = 0; x < 3; ++x)
        for (int y = x + 1; y < 3; ++y) {
            const int64_t sx = s_orig[order[x]] < 0 ? -s_orig[order[x]] : s_orig[order[x]];
            const int64_t sy = s_orig[order[y]] < 0 ? -s_orig[order[y]] : s_orig[order[y]];
            if (sy < sx) {
                const int t = order[x];
                order[x] = order[y];
                order[y] = t;
            }
        }
    BasicArgs a{};
    for (int q = 0; q < 3; ++q) {
        a.n[q] = n_orig[order[q]];
        a.s[q] = s_orig[order[q]];
        a.perm[order[q]] = q;
    }
    a.axis = a.perm[axis];
    a.lap_mask = lap_mask;
    if (a.n[0] == 0 || a.n[1] == 0 || a.n[2] == 0) {
        if (seconds) *seconds = 0.0;
        return 0;
    }

    sb_geom pg = *g;
    pg.si = a.s[0];
    pg.sj = a.s[1];
    pg.sk = a.s[2];
    const void *ptrs[2] = {inp, out};
    const int V = vector_width(sizeof(T), ptrs, 2, &pg);

    Context *ctx = context();
    if (!ctx) return 1;
    a.xcds = ctx->xcds;
    // launch->variant: 0 default, 1 the row kernel, 2 the marching kernel.
    // The marching kernel needs vector access along axis 0 and no term along
    // axis 2; by default it runs where a term lies along axis 1, whose rows
    // it reuses (measured on MI355X, 512x512x160 FP64: Laplacian ij 3.72 ->
    // 4.76 TB/s); with terms along axis 0 only the row kernel is faster
    // (more workgroups in flight, nothing to reuse).
    bool marchable = OP != OP_COPY && V > 1, axis1 = false;
    if constexpr (OP == OP_LAP) {
        for (int o = 0; o < 3; ++o) {
            if ((lap_mask & (1 << o)) && a.perm[o] == 2) marchable = false;
            if ((lap_mask & (1 << o)) && a.perm[o] == 1) axis1 = true;
        }
    } else {
        if (a.axis == 2) marchable = false;
        axis1 = a.axis == 1;
    }
    const int variant = l ? l->variant : 0;
    SB_REQUIRE(variant >= 0 && variant <= 3, "basic stencil variant %d not in {0, 1, 2, 3}", variant);
    // launch->variant 3: the LDS-DMA Laplacian (FP64, terms along i and j,
    // unit i stride, whole 64 x by tiles; launch->by = tile rows 8, 16 or 32,
    // 0 = 16 where ny allows, else 8), the default for such calls without
    // row options.  512x512x160: 107 us with 16 rows, 110 us with 32, 125 us
    // with 8, against 131 us for the marching kernel.
    constexpr bool lap64 = OP == OP_LAP && std::is_same_v<T, double> && std::is_same_v<C, double>;
    SB_REQUIRE(variant != 3 || lap64,
               "basic stencil variant 3 (LDS-DMA tile) takes FP64 Laplacians only");
    if constexpr (lap64) {
        const int ty = (l && l->by) ? l->by : (g->ny % 16 == 0 ? 16 : 8);
        const bool row_options = l && (l->by || l->bz);
        if (variant == 3 || (variant == 0 && !row_options && lap_dma_fits(g, lap_mask, ty))) {
            SB_REQUIRE(ty == 8 || ty == 16 || ty == 32, "LDS-DMA Laplacian tile rows %d not in {8, 16, 32}",
                       ty);
            SB_REQUIRE(lap_dma_fits(g, lap_mask, ty),
                       "the LDS-DMA Laplacian takes terms along i and j only, unit i stride, a "
                       "positive row pitch below 16 MiB, nx a multiple of 64 and ny a multiple of %d",
                       ty);
            return timed_launch(l, seconds, [&]() -> hipError_t {
                if (ty == 8) return launch_lap_dma<8, 2>(inp, out, g, ctx);
                if (ty == 16) return launch_lap_dma<16, 4>(inp, out, g, ctx);
                return launch_lap_dma<32, 8>(inp, out, g, ctx);
            });
        }
    }
    SB_REQUIRE(variant != 2 || marchable,
               "the marching kernel needs vector access along the unit-stride axis and no "
               "term along the largest-stride axis");
    const int64_t iblocks = (a.n[0] + int64_t(BLOCK) * V - 1) / (int64_t(BLOCK) * V);
    if (variant == 2 || (variant == 0 && marchable && axis1)) {
        // launch->bz: rows per thread (8, 16, 32, 64).  0: the shortest run
        // with which every workgroup is resident at once (a short last round
        // of workgroups idles most of the chip), else 64.  Each run re-reads
        // its two halo rows, which neighbouring runs rarely still hold in L2:
        // at 512x512x160 FP64, 8-row runs fetch 1.66x the input bytes.
        // threads per workgroup: one row segment of up to BLOCK x V points,
        // no wider than the row (whole waves)
        const int threads =
            int(std::min<int64_t>(BLOCK, ((a.n[0] + V - 1) / V + 63) / 64 * 64));
        const int64_t mblocks = (a.n[0] + int64_t(threads) * V - 1) / (int64_t(threads) * V);
        int JB = (l && l->bz) ? l->bz : 0;
        if (JB == 0) {
            int per_cu = 0;
            const void *probe = reinterpret_cast<const void *>(
                V == 4 && sizeof(T) == 4 ? &march_kernel<T, C, OP, (sizeof(T) == 4 ? 4 : 2), 8>
                                         : &march_kernel<T, C, OP, 2, 8>);
            SB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, probe, threads, 0));
            const int64_t resident = int64_t(std::max(per_cu, 1)) * ctx->compute_units;
            JB = 64;
            for (int rows : {8, 16, 32})
                if (((a.n[1] + rows - 1) / rows) * a.n[2] * mblocks <= resident) {
                    JB = rows;
                    break;
                }
        }
        SB_REQUIRE(JB == 8 || JB == 16 || JB == 32 || JB == 64,
                   "rows per thread %d not in {8, 16, 32, 64}", JB);
        const int64_t chunks = (a.n[1] + JB - 1) / JB;
        SB_REQUIRE(chunks * a.n[2] < (int64_t(1) << 31), "domain too large for one launch");
        const dim3 grid(unsigned(chunks * a.n[2]), unsigned(mblocks));
        auto launch_j = [&](auto v_tag) {
            constexpr int VV = decltype(v_tag)::value;
            auto go = [&](auto kernel) {
                enqueue_kernel(kernel, grid, dim3(threads), 0, ctx->stream, inp, out, a);
            };
            if (JB == 8) go(march_kernel<T, C, OP, VV, 8>);
            else if (JB == 16) go(march_kernel<T, C, OP, VV, 16>);
            else if (JB == 64) go(march_kernel<T, C, OP, VV, 64>);
            else go(march_kernel<T, C, OP, VV, 32>);
        };
        return timed_launch(l, seconds, [&]() -> hipError_t {
            if (V == 4 && sizeof(T) == 4)
                launch_j(std::integral_constant<int, (sizeof(T) == 4 ? 4 : 2)>{});
            else
                launch_j(std::integral_constant<int, 2>{});
            return hipGetLastError();
        });
    }
    const int64_t rows = a.n[1] * a.n[2];
    const int R = (l && l->by) ? l->by : 4;
    SB_REQUIRE(R == 4 || R == 8 || R == 16, "rows per workgroup %d not in {4, 8, 16}", R);
    const dim3 grid(unsigned((rows + R - 1) / R), unsigned(iblocks));
    auto launch_v = [&](auto kernel) {
        enqueue_kernel(kernel, grid, dim3(BLOCK), 0, ctx->stream, inp, out, a);
    };
    auto launch_r = [&](auto v_tag) {
        constexpr int VV = decltype(v_tag)::value;
        if (R == 4) launch_v(basic_kernel<T, C, OP, VV, 4>);
        else if (R == 16) launch_v(basic_kernel<T, C, OP, VV, 16>);
        else launch_v(basic_kernel<T, C, OP, VV, 8>);
    };
    return timed_launch(l, seconds, [&]() -> hipError_t {
        switch (V) {
            case 4:
                launch_r(std::integral_constant<int, (sizeof(T) == 4 ? 4 : 1)>{});
                break;
            case 2:
                launch_r(std::integral_constant<int, 2>{});
                break;
            default:
                launch_r(std::integral_constant<int, 1>{});
                break;
        }
        return hipGetLastError();
    });
}

template <int P>
int launch_copy(const void *inp, void *out, const sb_geom *g, const sb_launch *l,
                double *seconds) {
    return launch_basic_prec<P, OP_COPY>(inp, out, g, 0, 0, l, seconds);
}
template <int P>
int launch_avg1(const void *inp, void *out, const sb_geom *g, int axis, const sb_launch *l,
                double *seconds) {
    return launch_basic_prec<P, OP_AVG1>(inp, out, g, axis, 0, l, seconds);
}
template <int P>
int launch_avg2(const void *inp, void *out, const sb_geom *g, int axis, const sb_launch *l,
                double *seconds) {
    return launch_basic_prec<P, OP_AVG2>(inp, out, g, axis, 0, l, seconds);
}
template <int P>
int launch_lap(const void *inp, void *out, const sb_geom *g, int mask, const sb_launch *l,
               double *seconds) {
    return launch_basic_prec<P, OP_LAP>(inp, out, g, 0, mask, l, seconds);
}

}  // namespace
}  // namespace sb

using namespace sb;

extern "C" {

int sb_copy(int prec, const void *inp, void *out, const sb_geom *g, const sb_launch *l,
            double *seconds) {
    if (check_geom(g, 0, 0, 0)) return 2;
    return SB_DISPATCH_PREC(prec, launch_copy, inp, out, g, l, seconds);
}

int sb_onesided_average(int prec, int axis, const void *inp, void *out, const sb_geom *g,
                        const sb_launch *l, double *seconds) {
    SB_REQUIRE(axis >= 0 && axis < 3, "axis %d not in {0, 1, 2}", axis);
    if (check_geom(g, axis == 0, axis == 1, axis == 2)) return 2;
    return SB_DISPATCH_PREC(prec, launch_avg1, inp, out, g, axis, l, seconds);
}

int sb_symmetric_average(int prec, int axis, const void *inp, void *out, const sb_geom *g,
                         const sb_launch *l, double *seconds) {
    SB_REQUIRE(axis >= 0 && axis < 3, "axis %d not in {0, 1, 2}", axis);
    if (check_geom(g, axis == 0, axis == 1, axis == 2)) return 2;
    return SB_DISPATCH_PREC(prec, launch_avg2, inp, out, g, axis, l, seconds);
}

int sb_laplacian(int prec, int along_x, int along_y, int along_z, const void *inp, void *out,
                 const sb_geom *g, const sb_launch *l, double *seconds) {
    const int mask = (along_x ? 1 : 0) | (along_y ? 2 : 0) | (along_z ? 4 : 0);
    if (check_geom(g, along_x ? 1 : 0, along_y ? 1 : 0, along_z ? 1 : 0)) return 2;
    return SB_DISPATCH_PREC(prec, launch_lap, inp, out, g, mask, l, seconds);
}

}  // extern "C"
