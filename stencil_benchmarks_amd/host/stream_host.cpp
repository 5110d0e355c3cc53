// Host (CPU) STREAM: copy, scale, add and triad over three arrays with
// OpenMP, timed per kernel and iteration and checked against the closed form
// of the kernel sequence.  The reference generates, compiles and runs
// McCalpin's stream.c at run time (benchmarks_collection/stream/mc_calpin.py,
// original.j2, native.j2); this is a fixed build of the same measurement:
//   mode 0 ("original"): plain loops, vectorised by the compiler;
//   mode 1 ("native"):   explicit vector loads and non-temporal stores
//                        (AVX-512 when the CPU has it, else AVX2), which
//                        skip the read-for-ownership of the destination.
// The arrays are allocated 2 MiB-aligned and first touched by the thread
// that streams them (static schedule), as McCalpin's code does.
#include <immintrin.h>
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

thread_local std::string g_error;

template <typename T>
struct Arrays {
    T *a = nullptr, *b = nullptr, *c = nullptr;
    int64_t n = 0;
};

// [lo, hi) of thread t of `threads` over n elements, split at multiples of
// `align` elements so every thread's vector body starts aligned.
inline void chunk(int64_t n, int t, int threads, int64_t align, int64_t &lo, int64_t &hi) {
    const int64_t blocks = (n + align - 1) / align;
    const int64_t per = blocks / threads, extra = blocks % threads;
    const int64_t b0 = t * per + (t < extra ? t : extra);
    const int64_t b1 = b0 + per + (t < extra ? 1 : 0);
    lo = b0 * align < n ? b0 * align : n;
    hi = b1 * align < n ? b1 * align : n;
}

// ---- mode 0: plain loops ----------------------------------------------------
template <typename T>
void copy0(Arrays<T> &x) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < x.n; ++j) x.c[j] = x.a[j];
}
template <typename T>
void scale0(Arrays<T> &x, T s) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < x.n; ++j) x.b[j] = s * x.c[j];
}
template <typename T>
void add0(Arrays<T> &x) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < x.n; ++j) x.c[j] = x.a[j] + x.b[j];
}
template <typename T>
void triad0(Arrays<T> &x, T s) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < x.n; ++j) x.a[j] = x.b[j] + s * x.c[j];
}

// ---- mode 1: vector loads, non-temporal stores ------------------------------
// V: vector register type, per-type load/store/broadcast/ops through the
// overloads below; W elements per vector.
struct Avx512d { using V = __m512d; static constexpr int W = 8; };
struct Avx512s { using V = __m512;  static constexpr int W = 16; };
struct Avx2d   { using V = __m256d; static constexpr int W = 4; };
struct Avx2s   { using V = __m256;  static constexpr int W = 8; };

#define SBH_AVX512 __attribute__((target("avx512f")))
SBH_AVX512 inline __m512d vload(const double *p, Avx512d) { return _mm512_load_pd(p); }
SBH_AVX512 inline __m512 vload(const float *p, Avx512s) { return _mm512_load_ps(p); }
SBH_AVX512 inline void vstream(double *p, __m512d v) { _mm512_stream_pd(p, v); }
SBH_AVX512 inline void vstream(float *p, __m512 v) { _mm512_stream_ps(p, v); }
SBH_AVX512 inline __m512d vset(double s, Avx512d) { return _mm512_set1_pd(s); }
SBH_AVX512 inline __m512 vset(float s, Avx512s) { return _mm512_set1_ps(s); }
SBH_AVX512 inline __m512d vmul(__m512d a, __m512d b) { return _mm512_mul_pd(a, b); }
SBH_AVX512 inline __m512 vmul(__m512 a, __m512 b) { return _mm512_mul_ps(a, b); }
SBH_AVX512 inline __m512d vadd(__m512d a, __m512d b) { return _mm512_add_pd(a, b); }
SBH_AVX512 inline __m512 vadd(__m512 a, __m512 b) { return _mm512_add_ps(a, b); }
inline __m256d vload(const double *p, Avx2d) { return _mm256_load_pd(p); }
inline __m256 vload(const float *p, Avx2s) { return _mm256_load_ps(p); }
inline void vstream(double *p, __m256d v) { _mm256_stream_pd(p, v); }
inline void vstream(float *p, __m256 v) { _mm256_stream_ps(p, v); }
inline __m256d vset(double s, Avx2d) { return _mm256_set1_pd(s); }
inline __m256 vset(float s, Avx2s) { return _mm256_set1_ps(s); }
inline __m256d vmul(__m256d a, __m256d b) { return _mm256_mul_pd(a, b); }
inline __m256 vmul(__m256 a, __m256 b) { return _mm256_mul_ps(a, b); }
inline __m256d vadd(__m256d a, __m256d b) { return _mm256_add_pd(a, b); }
inline __m256 vadd(__m256 a, __m256 b) { return _mm256_add_ps(a, b); }

// One kernel over one thread's aligned chunk: K 0 copy, 1 scale, 2 add,
// 3 triad (the STREAM assignments c=a, b=s c, c=a+b, a=b+s c); the scalar
// tail handles the last partial vector.  Multiply and add stay separate
// roundings (no FMA), as in the closed-form check.  The body is compiled
// twice: for AVX2 (the build's baseline) and, in functions that carry the
// avx512f target, for AVX-512, so vector values never cross between code
// built for different vector ABIs.
#define SBH_VCHUNK_BODY                                                                       \
    using V = typename Isa::V;                                                                \
    const V sv = vset(s, Isa{});                                                              \
    int64_t j = lo;                                                                           \
    for (; j + Isa::W <= hi; j += Isa::W) {                                                   \
        if constexpr (K == 0) vstream(x.c + j, vload(x.a + j, Isa{}));                        \
        if constexpr (K == 1) vstream(x.b + j, vmul(sv, vload(x.c + j, Isa{})));              \
        if constexpr (K == 2)                                                                 \
            vstream(x.c + j, vadd(vload(x.a + j, Isa{}), vload(x.b + j, Isa{})));             \
        if constexpr (K == 3)                                                                 \
            vstream(x.a + j, vadd(vload(x.b + j, Isa{}), vmul(sv, vload(x.c + j, Isa{}))));   \
    }                                                                                         \
    for (; j < hi; ++j) {                                                                     \
        if constexpr (K == 0) x.c[j] = x.a[j];                                                \
        if constexpr (K == 1) x.b[j] = s * x.c[j];                                            \
        if constexpr (K == 2) x.c[j] = x.a[j] + x.b[j];                                       \
        if constexpr (K == 3) x.a[j] = x.b[j] + s * x.c[j];                                   \
    }

template <int K, typename Isa, typename T>
inline void vchunk(Arrays<T> &x, T s, int64_t lo, int64_t hi) {
    SBH_VCHUNK_BODY
}

template <int K, typename Isa, typename T>
SBH_AVX512 inline void vchunk512(Arrays<T> &x, T s, int64_t lo, int64_t hi) {
    SBH_VCHUNK_BODY
}

template <int K, typename Isa, typename T>
void vkernel(Arrays<T> &x, T s) {
#pragma omp parallel
    {
        int64_t lo, hi;
        chunk(x.n, omp_get_thread_num(), omp_get_num_threads(), 64, lo, hi);
        vchunk<K, Isa>(x, s, lo, hi);
    }
    _mm_sfence();
}

template <int K, typename Isa, typename T>
SBH_AVX512 void vkernel512(Arrays<T> &x, T s) {
#pragma omp parallel
    {
        int64_t lo, hi;
        chunk(x.n, omp_get_thread_num(), omp_get_num_threads(), 64, lo, hi);
        vchunk512<K, Isa>(x, s, lo, hi);
    }
    _mm_sfence();
}

template <typename T>
void *aligned_array(int64_t n) {
    void *p = nullptr;
    const size_t bytes = size_t(n) * sizeof(T);
    if (posix_memalign(&p, size_t(2) << 20, bytes ? bytes : 1) != 0) return nullptr;
    return p;
}

template <typename T>
int run(int64_t n, int ntimes, int mode, double *times, double *max_rel_err, int *isa) {
    Arrays<T> x;
    x.n = n;
    x.a = static_cast<T *>(aligned_array<T>(n));
    x.b = static_cast<T *>(aligned_array<T>(n));
    x.c = static_cast<T *>(aligned_array<T>(n));
    if (!x.a || !x.b || !x.c) {
        std::free(x.a); std::free(x.b); std::free(x.c);
        g_error = "host STREAM: allocation of 3 x " + std::to_string(n) + " elements failed";
        return 1;
    }
#pragma omp parallel
    {  // first touch by the thread that streams the chunk
        int64_t lo, hi;
        chunk(n, omp_get_thread_num(), omp_get_num_threads(), 64, lo, hi);
        for (int64_t j = lo; j < hi; ++j) {
            x.a[j] = T(1);
            x.b[j] = T(2);
            x.c[j] = T(0);
        }
    }
    const bool avx512 = mode == 1 && __builtin_cpu_supports("avx512f");
    *isa = mode == 0 ? 0 : (avx512 ? 512 : 256);
    const T s = T(3);
    // McCalpin's warm-up: a = 2 a before the timed iterations
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) x.a[j] = T(2) * x.a[j];
    using clock = std::chrono::steady_clock;
    for (int it = 0; it < ntimes; ++it) {
        for (int k = 0; k < 4; ++k) {
            const auto t0 = clock::now();
            if (mode == 0) {
                if (k == 0) copy0(x);
                if (k == 1) scale0(x, s);
                if (k == 2) add0(x);
                if (k == 3) triad0(x, s);
            } else if (avx512) {
                using I = typename std::conditional<sizeof(T) == 8, Avx512d, Avx512s>::type;
                if (k == 0) vkernel512<0, I>(x, s);
                if (k == 1) vkernel512<1, I>(x, s);
                if (k == 2) vkernel512<2, I>(x, s);
                if (k == 3) vkernel512<3, I>(x, s);
            } else {
                using I = typename std::conditional<sizeof(T) == 8, Avx2d, Avx2s>::type;
                if (k == 0) vkernel<0, I>(x, s);
                if (k == 1) vkernel<1, I>(x, s);
                if (k == 2) vkernel<2, I>(x, s);
                if (k == 3) vkernel<3, I>(x, s);
            }
            times[k * ntimes + it] = std::chrono::duration<double>(clock::now() - t0).count();
        }
    }
    // closed form: the same scalar sequence from a = 2, b = 2, c = 0
    T ea = T(2), eb = T(2), ec = T(0);
    for (int it = 0; it < ntimes; ++it) {
        ec = ea;
        eb = s * ec;
        ec = ea + eb;
        ea = eb + s * ec;
    }
    double worst = 0.0;
#pragma omp parallel for schedule(static) reduction(max : worst)
    for (int64_t j = 0; j < n; ++j) {
        const double ra = std::fabs(double(x.a[j]) - double(ea)) / std::fabs(double(ea));
        const double rb = std::fabs(double(x.b[j]) - double(eb)) / std::fabs(double(eb));
        const double rc = std::fabs(double(x.c[j]) - double(ec)) / std::fabs(double(ec));
        const double r = ra > rb ? (ra > rc ? ra : rc) : (rb > rc ? rb : rc);
        if (r > worst) worst = r;
    }
    *max_rel_err = worst;
    std::free(x.a);
    std::free(x.b);
    std::free(x.c);
    return 0;
}

}  // namespace

extern "C" {

// Run the host STREAM: itemsize 8 (double) or 4 (float), n elements per
// array, ntimes iterations of the four kernels; threads > 0 sets the
// OpenMP thread count.  times[k * ntimes + i] = seconds of kernel k
// (copy, scale, add, triad) in iteration i; max_rel_err = largest relative
// deviation of a, b, c from the closed form; isa = 0 (compiler loops),
// 256 or 512 (vector width of the explicit kernels).  Returns 0 or 1 with
// sbh_last_error() set.
int sbh_stream(int itemsize, int64_t n, int ntimes, int threads, int mode, double *times,
               double *max_rel_err, int *isa) {
    if (!times || !max_rel_err || !isa || n <= 0 || ntimes < 1 || (mode != 0 && mode != 1)) {
        g_error = "host STREAM: invalid arguments";
        return 1;
    }
    if (threads > 0) omp_set_num_threads(threads);
    if (itemsize == 8) return run<double>(n, ntimes, mode, times, max_rel_err, isa);
    if (itemsize == 4) return run<float>(n, ntimes, mode, times, max_rel_err, isa);
    g_error = "host STREAM: itemsize must be 4 or 8";
    return 1;
}

int sbh_max_threads(void) { return omp_get_max_threads(); }

const char *sbh_last_error(void) { return g_error.c_str(); }

}  // extern "C"
