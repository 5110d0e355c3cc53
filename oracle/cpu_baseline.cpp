// ORACLE - test infrastructure only (CPU baseline timed by bench.py).
//
// OpenMP C++ port of the reference's CPU backend variants, for the
// cpu_baseline figure reported next to the GPU numbers.  The reference
// renders these from Jinja2 templates at run time
// (benchmarks_collection/stencils/openmp/templates/*.j2), which cannot be
// built without that generator; this file restates the schedules:
//   hdiff "Classic" (openmp/templates/horizontal_diffusion_classic.j2:37-93):
//     (kb, jb, ib) blocks of (8, 8, 1) points distributed with OpenMP, per
//     block Laplacian, limited fluxes and update through small local arrays;
//   vadv "KMiddle" (openmp/templates/vertical_advection_kmiddle.j2:257-351):
//     (j, i-block of 8) distributed with OpenMP, k loop outside a simd i loop,
//     Thomas coefficients in per-block scratch;
//   Laplacian / copy "1D" (openmp/templates/basic_1d.j2:37-51) over the
//     interior rows.
// Built like the reference's CPU backend (openmp/mixin.py:88-101):
//   g++ -fopenmp -Ofast -march=native -mtune=native.
// -Ofast may reassociate, so results are checked with the reference's
// tolerances, not bit for bit.  Each entry point runs the kernel once and
// returns the wall time of that run (openmp/templates/base.j2:232-248).
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <vector>

using idx = std::ptrdiff_t;

namespace {

template <typename F>
double timed(int threads, F &&body) {
    if (threads > 0) omp_set_num_threads(threads);
    const auto start = std::chrono::steady_clock::now();
    body();
    const auto stop = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(stop - start).count();
}

template <typename T>
void hdiff_classic(const T *__restrict inp, const T *__restrict coeff, T *__restrict out, idx nx,
                   idx ny, idx nz, idx si, idx sj, idx sk) {
    constexpr idx BX = 8, BY = 8;
#pragma omp parallel
    {
        T lap[BY + 2][BX + 2];
        T flx[BY][BX + 1];
        T fly[BY + 1][BX];
#pragma omp for collapse(3) schedule(static)
        for (idx k = 0; k < nz; ++k)
            for (idx jb = 0; jb < ny; jb += BY)
                for (idx ib = 0; ib < nx; ib += BX) {
                    const idx imax = std::min(BX, nx - ib), jmax = std::min(BY, ny - jb);
                    const T *in = inp + k * sk + ib * si + jb * sj;
                    for (idx j = -1; j < jmax + 1; ++j)
#pragma omp simd
                        for (idx i = -1; i < imax + 1; ++i) {
                            const T *p = in + i * si + j * sj;
                            lap[j + 1][i + 1] = T(4) * p[0] - (((p[si] + p[-si]) + p[sj]) + p[-sj]);
                        }
                    for (idx j = 0; j < jmax; ++j)
#pragma omp simd
                        for (idx i = -1; i < imax; ++i) {
                            const T *p = in + i * si + j * sj;
                            const T f = lap[j + 1][i + 2] - lap[j + 1][i + 1];
                            flx[j][i + 1] = f * (p[si] - p[0]) > T(0) ? T(0) : f;
                        }
                    for (idx j = -1; j < jmax; ++j)
#pragma omp simd
                        for (idx i = 0; i < imax; ++i) {
                            const T *p = in + i * si + j * sj;
                            const T f = lap[j + 2][i + 1] - lap[j + 1][i + 1];
                            fly[j + 1][i] = f * (p[sj] - p[0]) > T(0) ? T(0) : f;
                        }
                    T *o = out + k * sk + ib * si + jb * sj;
                    const T *cf = coeff + k * sk + ib * si + jb * sj;
                    for (idx j = 0; j < jmax; ++j)
#pragma omp simd
                        for (idx i = 0; i < imax; ++i)
                            o[i * si + j * sj] =
                                in[i * si + j * sj] -
                                cf[i * si + j * sj] *
                                    (((flx[j][i + 1] - flx[j][i]) + fly[j + 1][i]) - fly[j][i]);
                }
    }
}

template <typename T>
void vadv_kmiddle(const T *__restrict stage, const T *__restrict pos, const T *__restrict tens,
                  T *__restrict tensstage, const T *__restrict wcon, idx nx, idx ny, idx nz,
                  idx si, idx sj, idx sk) {
    constexpr idx BX = 8;
    const T dtr = T(3.0 / 20.0), bet_m = T(0.5), bet_p = T(0.5);
#pragma omp parallel
    {
        std::vector<T> cc(nz * BX), dd(nz * BX);
#pragma omp for collapse(2) schedule(static)
        for (idx j = 0; j < ny; ++j)
            for (idx ib = 0; ib < nx; ib += BX) {
                const idx imax = std::min(BX, nx - ib);
                const idx base = ib * si + j * sj;
                for (idx k = 0; k < nz; ++k) {
#pragma omp simd
                    for (idx i = 0; i < imax; ++i) {
                        const idx p = base + i * si + k * sk;
                        const T d0 = dtr * pos[p] + tens[p] + tensstage[p];
                        if (k == 0) {
                            const T gcv = T(0.25) * (wcon[p + sk + si] + wcon[p + sk]);
                            const T c = gcv * bet_p, b = dtr - c;
                            const T corr = -(gcv * bet_m) * (stage[p + sk] - stage[p]);
                            cc[k * BX + i] = c / b;
                            dd[k * BX + i] = (d0 + corr) / b;
                        } else if (k < nz - 1) {
                            const T gav = T(-0.25) * (wcon[p + si] + wcon[p]);
                            const T gcv = T(0.25) * (wcon[p + sk + si] + wcon[p + sk]);
                            const T a = gav * bet_p, c = gcv * bet_p, b = dtr - a - c;
                            const T corr = -(gav * bet_m) * (stage[p - sk] - stage[p]) -
                                           (gcv * bet_m) * (stage[p + sk] - stage[p]);
                            const T r = T(1) / (b - cc[(k - 1) * BX + i] * a);
                            cc[k * BX + i] = c * r;
                            dd[k * BX + i] = (d0 + corr - dd[(k - 1) * BX + i] * a) * r;
                        } else {
                            const T gav = T(-0.25) * (wcon[p + si] + wcon[p]);
                            const T a = gav * bet_p, b = dtr - a;
                            const T corr = -(gav * bet_m) * (stage[p - sk] - stage[p]);
                            dd[k * BX + i] = (d0 + corr - dd[(k - 1) * BX + i] * a) /
                                             (b - cc[(k - 1) * BX + i] * a);
                        }
                    }
                }
                T x[BX];
                for (idx k = nz - 1; k >= 0; --k) {
#pragma omp simd
                    for (idx i = 0; i < imax; ++i) {
                        const idx p = base + i * si + k * sk;
                        x[i] = (k == nz - 1) ? dd[k * BX + i]
                                             : dd[k * BX + i] - cc[k * BX + i] * x[i];
                        tensstage[p] = dtr * (x[i] - pos[p]);
                    }
                }
            }
    }
}

template <typename T>
void laplacian_1d(const T *__restrict inp, T *__restrict out, idx nx, idx ny, idx nz, idx si,
                  idx sj, idx sk) {
#pragma omp parallel for collapse(2) schedule(static)
    for (idx k = 0; k < nz; ++k)
        for (idx j = 0; j < ny; ++j) {
            const T *p = inp + j * sj + k * sk;
            T *o = out + j * sj + k * sk;
#pragma omp simd
            for (idx i = 0; i < nx; ++i) {
                const idx q = i * si;
                o[q] = ((T(2) * p[q] - p[q + si]) - p[q - si]) + ((T(2) * p[q] - p[q + sj]) - p[q - sj]);
            }
        }
}

template <typename T>
void copy_1d(const T *__restrict inp, T *__restrict out, idx nx, idx ny, idx nz, idx si, idx sj,
             idx sk) {
#pragma omp parallel for collapse(2) schedule(static)
    for (idx k = 0; k < nz; ++k)
        for (idx j = 0; j < ny; ++j) {
            const T *p = inp + j * sj + k * sk;
            T *o = out + j * sj + k * sk;
#pragma omp simd
            for (idx i = 0; i < nx; ++i) o[i * si] = p[i * si];
        }
}

}  // namespace

extern "C" {

int cpu_max_threads(void) { return omp_get_max_threads(); }

#define GEOM idx nx, idx ny, idx nz, idx si, idx sj, idx sk
#define GARGS nx, ny, nz, si, sj, sk

double cpu_hdiff_f64(const double *inp, const double *coeff, double *out, GEOM, int threads) {
    return timed(threads, [&] { hdiff_classic(inp, coeff, out, GARGS); });
}
double cpu_hdiff_f32(const float *inp, const float *coeff, float *out, GEOM, int threads) {
    return timed(threads, [&] { hdiff_classic(inp, coeff, out, GARGS); });
}
double cpu_vadv_f64(const double *stage, const double *pos, const double *tens, double *tensstage,
                    const double *wcon, GEOM, int threads) {
    return timed(threads, [&] { vadv_kmiddle(stage, pos, tens, tensstage, wcon, GARGS); });
}
double cpu_vadv_f32(const float *stage, const float *pos, const float *tens, float *tensstage,
                    const float *wcon, GEOM, int threads) {
    return timed(threads, [&] { vadv_kmiddle(stage, pos, tens, tensstage, wcon, GARGS); });
}
double cpu_laplacian_f64(const double *inp, double *out, GEOM, int threads) {
    return timed(threads, [&] { laplacian_1d(inp, out, GARGS); });
}
double cpu_copy_f64(const double *inp, double *out, GEOM, int threads) {
    return timed(threads, [&] { copy_1d(inp, out, GARGS); });
}

}  // extern "C"
