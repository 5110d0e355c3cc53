/*
 * ORACLE - test infrastructure only.  Nothing in the product links, loads
 * or calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker.
 *
 * Plain C restatement of the reference's verification oracles
 * (stencil_benchmarks/benchmarks_collection/stencils/base.py) and of the
 * GPU STREAM recurrence (benchmarks_collection/stream/cuda_hip.j2), scalar
 * loops over strided fields.  Every arithmetic operation is written in the
 * order numpy evaluates the reference's expressions, and the file is built
 * with -O2 -ffp-contract=off on x86-64 (SSE2 arithmetic, no x87 excess
 * precision, no FMA contraction), so float64 results equal the reference
 * oracle's bit for bit.
 *
 * Conventions: field pointers address the first interior point; si, sj, sk
 * are element strides of the logical axes i, j, k; nx, ny, nz the interior
 * extents.  Each function exists for float (suffix _f32, float arithmetic
 * like numpy on float32 arrays), double (_f64) and, for the Thomas solve,
 * long double arithmetic on double fields (_ld) for error bars.
 */
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

typedef ptrdiff_t idx;

#define AT(p, i, j, k) (p)[(idx)(i) * si + (idx)(j) * sj + (idx)(k) * sk]

/* Copy (base.py:180-188): out = inp on the interior. */
#define DEFINE_COPY(SUF, T)                                                    \
    void oracle_copy_##SUF(const T *inp, T *out, idx nx, idx ny, idx nz,       \
                           idx si, idx sj, idx sk) {                           \
        for (idx k = 0; k < nz; ++k)                                           \
            for (idx j = 0; j < ny; ++j)                                       \
                for (idx i = 0; i < nx; ++i) AT(out, i, j, k) = AT(inp, i, j, k); \
    }

/* One-sided / symmetric averages (base.py:191-222):
 * (inp[x + e] + inp[x]) / 2 and (inp[x + e] + inp[x - e]) / 2. */
#define DEFINE_AVERAGES(SUF, T)                                                \
    void oracle_onesided_average_##SUF(const T *inp, T *out, int axis, idx nx, \
                                       idx ny, idx nz, idx si, idx sj,         \
                                       idx sk) {                               \
        const idx e = axis == 0 ? si : axis == 1 ? sj : sk;                    \
        for (idx k = 0; k < nz; ++k)                                           \
            for (idx j = 0; j < ny; ++j)                                       \
                for (idx i = 0; i < nx; ++i) {                                 \
                    const T *p = &AT(inp, i, j, k);                            \
                    AT(out, i, j, k) = (T)((p[e] + p[0]) / (T)2);              \
                }                                                              \
    }                                                                          \
    void oracle_symmetric_average_##SUF(const T *inp, T *out, int axis,        \
                                        idx nx, idx ny, idx nz, idx si,        \
                                        idx sj, idx sk) {                      \
        const idx e = axis == 0 ? si : axis == 1 ? sj : sk;                    \
        for (idx k = 0; k < nz; ++k)                                           \
            for (idx j = 0; j < ny; ++j)                                       \
                for (idx i = 0; i < nx; ++i) {                                 \
                    const T *p = &AT(inp, i, j, k);                            \
                    AT(out, i, j, k) = (T)((p[e] + p[-e]) / (T)2);             \
                }                                                              \
    }

/* Laplacian (base.py:238-254): result starts at 0 and accumulates
 * (2 inp - inp[+e]) - inp[-e] for each enabled axis in x, y, z order. */
#define DEFINE_LAPLACIAN(SUF, T)                                               \
    void oracle_laplacian_##SUF(const T *inp, T *out, int along_x,             \
                                int along_y, int along_z, idx nx, idx ny,      \
                                idx nz, idx si, idx sj, idx sk) {              \
        const int on[3] = {along_x, along_y, along_z};                         \
        const idx st[3] = {si, sj, sk};                                        \
        for (idx k = 0; k < nz; ++k)                                           \
            for (idx j = 0; j < ny; ++j)                                       \
                for (idx i = 0; i < nx; ++i) {                                 \
                    const T *p = &AT(inp, i, j, k);                            \
                    T r = (T)0;                                                \
                    for (int a = 0; a < 3; ++a)                                \
                        if (on[a]) {                                           \
                            const T term = (T)((T)((T)2 * p[0] - p[st[a]]) -   \
                                               p[-st[a]]);                     \
                            r = (T)(r + term);                                 \
                        }                                                      \
                    AT(out, i, j, k) = r;                                      \
                }                                                              \
    }

/* Horizontal diffusion (base.py:284-307).  lap at (i, j) is
 * 4 in - (((in(i+1) + in(i-1)) + in(j+1)) + in(j-1)); fluxes are Laplacian
 * differences zeroed where flux * (input difference) > 0; the update is
 * in - coeff * (((flx(i) - flx(i-1)) + fly(j)) - fly(j-1)).  Laplacians are
 * recomputed per use, which yields the same values as the oracle's arrays. */
#define DEFINE_HDIFF(SUF, T)                                                   \
    static T lap_##SUF(const T *p, idx si, idx sj) {                           \
        return (T)((T)4 * p[0] -                                               \
                   (T)((T)((T)(p[si] + p[-si]) + p[sj]) + p[-sj]));            \
    }                                                                          \
    static T limit_##SUF(T flux, T gradient) {                                 \
        return (T)(flux * gradient) > (T)0 ? (T)0 : flux;                      \
    }                                                                          \
    void oracle_hdiff_##SUF(const T *inp, const T *coeff, T *out, idx nx,      \
                            idx ny, idx nz, idx si, idx sj, idx sk) {          \
        for (idx k = 0; k < nz; ++k)                                           \
            for (idx j = 0; j < ny; ++j)                                       \
                for (idx i = 0; i < nx; ++i) {                                 \
                    const T *p = &AT(inp, i, j, k);                            \
                    const T lc = lap_##SUF(p, si, sj);                         \
                    const T flx = limit_##SUF((T)(lap_##SUF(p + si, si, sj) - lc), \
                                              (T)(p[si] - p[0]));              \
                    const T flx_m = limit_##SUF((T)(lc - lap_##SUF(p - si, si, sj)), \
                                                (T)(p[0] - p[-si]));           \
                    const T fly = limit_##SUF((T)(lap_##SUF(p + sj, si, sj) - lc), \
                                              (T)(p[sj] - p[0]));              \
                    const T fly_m = limit_##SUF((T)(lc - lap_##SUF(p - sj, si, sj)), \
                                                (T)(p[0] - p[-sj]));           \
                    const T sum = (T)((T)((T)(flx - flx_m) + fly) - fly_m);    \
                    AT(out, i, j, k) = (T)(p[0] - (T)(AT(coeff, i, j, k) * sum)); \
                }                                                              \
    }

/* Vertical advection, one component (base.py:410-473).  Fields hold
 * storage type S, arithmetic runs in T.  tensstage is read in the forward
 * sweep and overwritten with the result.  (ishift, jshift) selects the wcon
 * neighbour: (1, 0) u, (0, 1) v, (0, 0) w. */
#define DEFINE_VADV(SUF, S, T)                                                 \
    int oracle_vadv_##SUF(const S *stage, const S *pos, const S *tens,         \
                          S *tensstage, const S *wcon, int ishift, int jshift, \
                          idx nx, idx ny, idx nz, idx si, idx sj, idx sk) {    \
        const T dtr = (T)(3.0 / 20.0);                                         \
        const T bet_m = (T)0.5, bet_p = (T)0.5;                                \
        const idx ws = ishift * si + jshift * sj;                              \
        T *cc = (T *)malloc(sizeof(T) * (size_t)nz);                           \
        T *dd = (T *)malloc(sizeof(T) * (size_t)nz);                           \
        if (!cc || !dd || nz < 2) {                                            \
            free(cc);                                                          \
            free(dd);                                                          \
            return 1;                                                          \
        }                                                                      \
        for (idx j = 0; j < ny; ++j)                                           \
            for (idx i = 0; i < nx; ++i) {                                     \
                const S *w = &AT(wcon, i, j, 0);                               \
                const S *st = &AT(stage, i, j, 0);                             \
                const S *ps = &AT(pos, i, j, 0);                               \
                const S *tn = &AT(tens, i, j, 0);                              \
                S *ts = &AT(tensstage, i, j, 0);                               \
                for (idx k = 0; k < nz; ++k) {                                 \
                    const idx o = k * sk;                                      \
                    const T d0 = (T)((T)((T)(dtr * (T)ps[o]) + (T)tn[o]) + (T)ts[o]); \
                    if (k == 0) {                                              \
                        const T gcv = (T)((T)0.25 * (T)((T)w[o + sk + ws] + (T)w[o + sk])); \
                        const T cs = (T)(gcv * bet_m);                         \
                        const T c = (T)(gcv * bet_p);                          \
                        const T b = (T)(dtr - c);                              \
                        const T corr = (T)(-cs * (T)((T)st[o + sk] - (T)st[o])); \
                        cc[k] = (T)(c / b);                                    \
                        dd[k] = (T)((T)(d0 + corr) / b);                       \
                    } else if (k < nz - 1) {                                   \
                        const T gav = (T)((T)-0.25 * (T)((T)w[o + ws] + (T)w[o])); \
                        const T gcv = (T)((T)0.25 * (T)((T)w[o + sk + ws] + (T)w[o + sk])); \
                        const T as = (T)(gav * bet_m), cs = (T)(gcv * bet_m);  \
                        const T a = (T)(gav * bet_p), c = (T)(gcv * bet_p);    \
                        const T b = (T)((T)(dtr - a) - c);                     \
                        const T corr = (T)((T)(-as * (T)((T)st[o - sk] - (T)st[o])) - \
                                           (T)(cs * (T)((T)st[o + sk] - (T)st[o]))); \
                        const T r = (T)((T)1 / (T)(b - (T)(cc[k - 1] * a)));   \
                        cc[k] = (T)(c * r);                                    \
                        dd[k] = (T)((T)((T)(d0 + corr) - (T)(dd[k - 1] * a)) * r); \
                    } else {                                                   \
                        const T gav = (T)((T)-0.25 * (T)((T)w[o + ws] + (T)w[o])); \
                        const T as = (T)(gav * bet_m), a = (T)(gav * bet_p);   \
                        const T b = (T)(dtr - a);                              \
                        const T corr = (T)(-as * (T)((T)st[o - sk] - (T)st[o])); \
                        dd[k] = (T)((T)((T)(d0 + corr) - (T)(dd[k - 1] * a)) / \
                                    (T)(b - (T)(cc[k - 1] * a)));              \
                        cc[k] = (T)0;                                          \
                    }                                                          \
                }                                                              \
                T x = (T)0;                                                    \
                for (idx k = nz - 1; k >= 0; --k) {                            \
                    x = (k == nz - 1) ? dd[k] : (T)(dd[k] - (T)(cc[k] * x));   \
                    ts[k * sk] = (S)(T)(dtr * (T)(x - (T)ps[k * sk]));          \
                }                                                              \
            }                                                                  \
        free(cc);                                                              \
        free(dd);                                                              \
        return 0;                                                              \
    }

DEFINE_COPY(f64, double)
DEFINE_COPY(f32, float)
DEFINE_AVERAGES(f64, double)
DEFINE_AVERAGES(f32, float)
DEFINE_LAPLACIAN(f64, double)
DEFINE_LAPLACIAN(f32, float)
DEFINE_HDIFF(f64, double)
DEFINE_HDIFF(f32, float)
DEFINE_VADV(f64, double, double)
DEFINE_VADV(f32, float, float)
DEFINE_VADV(ld, double, long double)

/* GPU STREAM recurrence (stream/cuda_hip.j2:147-173, :329-340), elementwise:
 * c = a; b = s c; c = a + b; a = b + s c, repeated ntimes. */
#define DEFINE_STREAM(SUF, T)                                                  \
    void oracle_stream_##SUF(T *a, T *b, T *c, idx n, int ntimes, double s) {  \
        const T sc = (T)s;                                                     \
        for (idx i = 0; i < n; ++i) {                                          \
            T ai = a[i], bi = b[i], ci = c[i];                                 \
            for (int t = 0; t < ntimes; ++t) {                                 \
                ci = ai;                                                       \
                bi = (T)(sc * ci);                                             \
                ci = (T)(ai + bi);                                             \
                ai = (T)(bi + (T)(sc * ci));                                   \
            }                                                                  \
            a[i] = ai;                                                         \
            b[i] = bi;                                                         \
            c[i] = ci;                                                         \
        }                                                                      \
    }

DEFINE_STREAM(f64, double)
DEFINE_STREAM(f32, float)
