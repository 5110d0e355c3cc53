/*
 * sbmi355x.h - C ABI of libsbmi355x.so, the MI355X (gfx950) stencil and
 * STREAM library.
 *
 * Conventions (shared by every entry point):
 *  - Every function returns 0 on success and a non-zero status on failure;
 *    sb_last_error() then returns a message (thread-local).  This replaces
 *    the reference's "return 1 and print to stderr" launcher convention
 *    (stencils/cuda_hip/templates/base.j2:48-55) whose stderr the Python side
 *    captured by redirecting process-wide file descriptors
 *    (tools/compilation.py:155-196); nothing here touches fds.
 *  - Field pointers are DEVICE pointers to the first interior point (the
 *    element at index == halo), exactly like the reference's
 *    data_ptr(device_array, halo) arguments (stencils/cuda_hip/mixin.py:162-175).
 *    The geometry that the reference baked into generated code as literals
 *    (domain, strides) is passed at run time in sb_geom.
 *  - Compute entry points launch on the library's stream of the current
 *    device.  If `seconds` is non-NULL the call runs launch->dry_runs untimed
 *    launches, optionally evicts the caches (launch->flush), brackets one
 *    launch with HIP events on that stream, waits for it and stores the
 *    elapsed time in seconds (the reference's "gpu" timer,
 *    cuda_hip/templates/base.j2:145-160).  If `seconds` is NULL the launch is
 *    asynchronous; call sb_synchronize().
 *  - Precision codes: SB_F64 fields and arithmetic in float64; SB_F32 fields
 *    and arithmetic in float32; SB_F32_F64 fields stored as float32, all
 *    arithmetic in float64, results rounded once to float32.
 *
 * Paths cited below are relative to the reference's stencil_benchmarks/
 * package directory.
 */
#ifndef SBMI355X_H
#define SBMI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SB_ABI_VERSION 2

enum sb_precision { SB_F64 = 0, SB_F32 = 1, SB_F32_F64 = 2 };

enum sb_stream_op { SB_STREAM_COPY = 0, SB_STREAM_SCALE = 1, SB_STREAM_ADD = 2, SB_STREAM_TRIAD = 3 };

/* Field geometry: interior extents, element strides of the logical axes
 * (i, j, k) and the halo width that is addressable on each side.  Replaces
 * the compile-time `domain` / `strides` template arguments
 * (stencils/cuda_hip/mixin.py:113-123, stencils/base.py:115-121). */
typedef struct sb_geom {
    int64_t nx, ny, nz;
    int64_t si, sj, sk;
    int64_t hi, hj, hk;
} sb_geom;

/* Timers of a timed launch (the `timers` choices 'gpu' and 'hip-ext' of
 * stencils/cuda_hip/mixin.py:57; 'wall' is a host clock around the call).
 * SB_TIMER_EVENTS records events on the stream before and after the
 * kernels; SB_TIMER_HIP_EXT has the kernels record them themselves through
 * hipExtLaunchKernelGGL (templates/base.j2:160-161). */
#define SB_TIMER_EVENTS 0
#define SB_TIMER_HIP_EXT 1

/* Launch options; zero means "library default / autotuned".  Replaces the
 * block_size / threads_per_block / unroll_factor / dry_runs / timers
 * parameters (stencils/cuda_hip/basic.py:43-48, horizontal_diffusion.py:40-47,
 * vertical_advection.py:39-57, mixin.py:56-57). */
typedef struct sb_launch {
    int32_t variant;   /* kernel variant id, see the per-function notes */
    int32_t bx, by, bz;
    int32_t unroll;
    int32_t dry_runs;
    int32_t flush;     /* 1: evict L2 and Infinity Cache before the timed launch */
    int32_t timer;     /* SB_TIMER_EVENTS (0) or SB_TIMER_HIP_EXT */
} sb_launch;

typedef struct sb_device_info {
    char name[256];
    char arch[64];
    int32_t compute_units;
    int32_t wave_size;
    int32_t max_threads_per_block;
    int32_t xcds;
    int64_t lds_per_block;
    int64_t lds_per_cu;
    int64_t total_mem;
    int64_t l2_bytes;
    int32_t clock_khz;
    int32_t mem_clock_khz;
    int32_t mem_bus_width;
    int32_t reserved;
} sb_device_info;

/* ---- errors, lifecycle, device ------------------------------------------ */
const char *sb_last_error(void);
int sb_abi_version(void);
/* Select and initialise a device (stream, events).  Replaces the implicit
 * device setup of the generated launcher (cuda_hip/templates/base.j2:68-91). */
int sb_init(int device);
int sb_device_count(int *count);
/* Device properties; replaces hipGetDeviceProperties in base.j2:76-91. */
int sb_get_device_info(int device, sb_device_info *info);
int sb_mem_info(size_t *free_bytes, size_t *total_bytes);

/* ---- memory ---------------------------------------------------------------
 * Replace api.Runtime.malloc / memcpy / device_synchronize
 * (stencils/cuda_hip/api.py:54-95). */
int sb_malloc(void **ptr, size_t bytes);
int sb_free(void *ptr);
int sb_host_malloc(void **ptr, size_t bytes); /* pinned host memory */
int sb_host_free(void *ptr);
int sb_memcpy_h2d(void *dst, const void *src, size_t bytes);
int sb_memcpy_d2h(void *dst, const void *src, size_t bytes);
int sb_memcpy_d2d(void *dst, const void *src, size_t bytes);
int sb_memset(void *dst, int value, size_t bytes);
int sb_synchronize(void);
/* Markers on the library stream (the reference times with events inside the
 * generated launcher, cuda_hip/templates/base.j2:146-183, and has no public
 * event API): sb_event_record enqueues the event behind all work submitted
 * so far; sb_event_synchronize waits until that work has completed, without
 * waiting for work submitted after the record. */
int sb_event_create(void **event);
int sb_event_record(void *event);
int sb_event_synchronize(void *event);
int sb_event_destroy(void *event);
/* Stream a scratch buffer larger than L2 + Infinity Cache through the
 * device; the cold-cache counterpart of cycling `data_sets`
 * (stencils/base.py:47-51, :152). */
int sb_flush_caches(void);

/* ---- basic stencils ------------------------------------------------------
 * Computations of the generated basic kernels
 * (stencils/cuda_hip/basic.py:101-132, templates/basic_1d.j2, basic_3d.j2),
 * interior points only, with the arithmetic order of the verification
 * oracles (stencils/base.py:180-254).  Launch options (averages and
 * Laplacian): ->variant 0 default, 1 row kernel (->by = rows per workgroup
 * 4/8/16), 2 marching kernel (a thread walks ->bz = 8/16/32/64 rows of the
 * second-smallest-stride axis keeping three rows in registers; needs
 * 16-byte vector access along the unit-stride axis and no term along the
 * largest-stride axis), 3 (sb_laplacian only: SB_F64 with terms along x and
 * y only, si == 1, 0 < sj with a row pitch below 16 MiB, nx a multiple of 64 and ny of
 * ->by) LDS-DMA tile: a
 * 64 x ->by tile (8/16/32 rows, 0 = 16 where ny allows, else 8) of one
 * k-plane per workgroup, staged by global-to-LDS loads.  The default is 3
 * where it applies and no ->by/->bz is given, else marches where a term lies
 * along the second axis. */
int sb_copy(int prec, const void *inp, void *out, const sb_geom *g,
            const sb_launch *l, double *seconds);
/* out = (inp[x + e_axis] + inp[x]) / 2 (base.py:191-205). */
int sb_onesided_average(int prec, int axis, const void *inp, void *out,
                        const sb_geom *g, const sb_launch *l, double *seconds);
/* out = (inp[x + e_axis] + inp[x - e_axis]) / 2 (base.py:208-222). */
int sb_symmetric_average(int prec, int axis, const void *inp, void *out,
                         const sb_geom *g, const sb_launch *l, double *seconds);
/* out = sum over enabled axes of ((2 inp - inp[+e]) - inp[-e]) (base.py:225-254). */
int sb_laplacian(int prec, int along_x, int along_y, int along_z,
                 const void *inp, void *out, const sb_geom *g,
                 const sb_launch *l, double *seconds);

/* ---- horizontal diffusion ------------------------------------------------
 * Replaces kernel(double*, inp, coeff, out) of the hdiff variants
 * (stencils/cuda_hip/horizontal_diffusion.py:50-128 and
 * templates/horizontal_diffusion_*.j2); semantics of base.py:276-311.
 * Requires hi, hj >= 2.  Variants: 0 default (2 for SB_F64 domains of
 * whole 64 x 8 tiles and float32 domains of whole 64 x 16 tiles when no
 * tile option is given, else 1 where positive i
 * and j strides and a k-plane below 4 GiB let the tile use 32-bit plane
 * offsets, else 3), 1 LDS-staged tile
 * (launch->by = tile rows 8/16/32, ->bz = k-planes per workgroup 1..16, 0
 * measured default; ->unroll 0 = default (non-temporal coeff loads), else
 * 4 | flags with bit 0 non-temporal out stores, bit 1 non-temporal coeff
 * loads), 2 LDS tile staged by global-to-LDS loads, one k-plane per
 * workgroup (every precision; launch->by = tile rows 8 or 16, 0 = 8 for
 * SB_F64 and 16 for float32 fields; needs nx a multiple of 64, ny of the
 * tile rows, si = 1, 0 < sj with a row pitch below 16 MiB, and for SB_F64 even sj and
 * sk and a 16-byte aligned inp), 3 per-point recomputation through the caches
 * (the reference's OnTheFly schedule). */
int sb_hdiff(int prec, const void *inp, const void *coeff, void *out,
             const sb_geom *g, const sb_launch *l, double *seconds);

/* ---- vertical advection --------------------------------------------------
 * Replaces kernel(double*, ustage, upos, utens, utensstage, wcon, ccol,
 * dcol, datacol) of the vadv variants (stencils/cuda_hip/vertical_advection.py:60-73,
 * templates/vertical_advection_*.j2); semantics of base.py:349-501.
 * `fields` holds the device pointers in the reference's argument order
 * (base.py:329-337): 8 entries, or 16 when all_components != 0.  Only the
 * *tensstage fields are written; ccol, dcol and datacol are never touched
 * (the Thomas coefficients stay on chip, or in a library-owned scratch for
 * variant 2).  Requires hi >= 1 (all components: hj >= 1 too).  Variants:
 * 0 default (5 where nz <= 96 and every 64-column group is in flight at
 * once; else 4 for nz >= 96, or nz >= 48 with float32 storage, with at
 * least two column groups per resident workgroup, else 3; beyond the LDS
 * limit of those (nz > 160 with FP64 arithmetic) 5 up to nz = 201, then 2),
 * 1 one wave per
 * group with the coefficients in LDS, 2 coefficients in a global FP64
 * scratch, 3 single-wave pipeline fusing the forward sweep of one group with
 * the backward of the previous, 4 warp-specialised solver (four waves share
 * one LDS slot buffer; launch->unroll = levels per barrier step, 4 or 8,
 * 0 = 4, when variant 4 is asked for explicitly), 5 one wave per group with
 * the coefficients of the lowest 40 levels (80 with float32 arithmetic) in
 * registers and the rest in LDS (nz <= 201 with FP64 arithmetic).
 * launch->bz: bit 0
 * non-temporal streamed loads, bit 2 (variant 4, SB_F32_F64 only) reciprocal
 * estimate plus Newton steps instead of the correctly rounded division; 0
 * and -1 plain loads. */
int sb_vadv(int prec, int all_components, void *const *fields,
            const sb_geom *g, const sb_launch *l, double *seconds);

/* sb_vadv without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as sb_vadv with the global-scratch solver (variant 2, bit for bit
 * the same arithmetic); its (c', d') scratch is a 64 MiB array of the
 * library's code object (one per device, held until the process ends)
 * instead of an allocation, and a persistent grid of
 * 64 MiB / (64 nz 16 bytes) workgroups (FP64 arithmetic) strides over the
 * columns.  Rejects nz above 65536 (FP64 arithmetic) or 131072 (float32) and
 * launch->flush (cache flushing allocates).  launch->variant 0 or 2,
 * ->unroll prefetch depth 4/8/16, ->bz bit 0 non-temporal loads.  Calls on
 * one device share the array; the library orders its calls on one stream, so
 * calls overlap (and race) only if a caller runs them on other streams or
 * processes at once.  The array keeps the last call's contents; nothing in
 * it needs to start at zero. */
int sb_vadv_noalloc(int prec, int all_components, void *const *fields,
                    const sb_geom *g, const sb_launch *l, double *seconds);

/* ---- STREAM ---------------------------------------------------------------
 * Replaces the generated STREAM program (benchmarks_collection/stream/cuda_hip.j2:132-288).
 * prec is SB_F64 or SB_F32.  sb_stream_kernel runs one of copy (c = a),
 * scale (b = s c), add (c = a + b), triad (a = b + s c) over n elements.
 * launch->bx = threads per block (0: 256), ->by = elements per vector
 * (0: 16 bytes; 1 or 2 for SB_F64, 1, 2 or 4 for SB_F32), ->bz = workgroups
 * per CU of a persistent grid (0: a grid covering the arrays once),
 * ->unroll = vectors per thread (0: 4), ->variant bit 0 = non-temporal
 * loads, bit 1 = non-temporal stores, bits 2-3 = grid axis 0 x, 1 y, 2 z
 * (block_size / vector_size / unroll_factor / streaming_* / axis of
 * stream/cuda_hip.py:41-60). */
int sb_stream_fill(int prec, void *x, size_t n, double value);
/* x[i] = uniform [0,1) from a counter-based hash of (seed, i); returns after
 * the fill completed. */
int sb_stream_fill_random(int prec, void *x, size_t n, uint64_t seed);
/* Replaces the setup-time random_fill of every field (tools/parallel.cpp:51-96,
 * stencils/base.py:66-85): fields[f][i] = the value sb_stream_fill_random
 * gives with seeds[f], for count <= 16 arrays of n elements, in one launch
 * on the library stream (asynchronous: ordered before later launches). */
int sb_fill_random_fields(int prec, void *const *fields, int count, size_t n,
                          const uint64_t *seeds);
int sb_stream_kernel(int prec, int op, void *a, void *b, void *c, size_t n,
                     double scalar, const sb_launch *l, double *seconds);
/* max over i of |x[i] - expected| / |expected| (closed-form check of
 * stream/cuda_hip.j2:290-345). */
int sb_stream_check(int prec, const void *x, size_t n, double expected,
                    double *max_rel_err);

/* ---- batched subdomains ---------------------------------------------------
 * No counterpart in the reference (its stencils run one domain per launch).
 * A subdomain owns seven fields with one shared geometry: the state pair
 * u[0] / u[1] (hdiff input and output, swapped every step), coeff, and the
 * vadv fields upos, utens, utensstage (read-write) and wcon.  One timestep is
 *     u[1-parity] = hdiff(u[parity], coeff)        (interior only)
 *     utensstage  = vadv(ustage = u[1-parity], upos, utens, utensstage, wcon)
 * (u component).  The halos of u[0] and u[1] must hold the same boundary
 * values; they are never written.  Requirements: hi, hj >= 2, nz >= 2,
 * positive strides, every field below 4 GiB, nz small enough that a 64-column
 * Thomas buffer fits in LDS (nz <= 160 with FP64 arithmetic).
 * Launch options: hdiff_launch->by / ->bz / ->unroll = tile height / planes
 * per item / non-temporal flags (as sb_hdiff); vadv_launch->variant = 4
 * (default: warp-specialised solver, items of ->by column groups, 0 = 16,
 * drawn from per-XCD device work queues in order of decreasing cost, ->unroll
 * = levels per barrier step 4 or 8, 0 = 4, ->bx = waves per workgroup 4 or
 * 5, 0 = 4, the fifth loading wcon for the others) or 3
 * (single-wave pipeline, items from a device counter; ->unroll = prefetch
 * depth 4 or 8, ->bz = 1 non-temporal streamed loads); with variant 4,
 * ->bz bit 2 selects the fast reciprocal (SB_F32_F64 only). */
typedef struct sb_subdomain {
    void *u[2];
    void *coeff;
    void *upos, *utens, *utensstage, *wcon;
    sb_geom geom;
    int32_t parity; /* index of the current state in u[] */
    int32_t reserved;
} sb_subdomain;

/* One timestep for `count` subdomains: a counter reset, one persistent hdiff
 * launch (tiles from per-XCD device work queues) and one or two persistent
 * vadv launches (warp-specialised: column-group items from per-XCD device
 * work queues, one launch per LDS class; variant 3: items from one device
 * counter).  Does not flip the parities; the caller does. */
int sb_batch_step(int prec, const sb_subdomain *subs, int count, const sb_launch *hdiff_launch,
                  const sb_launch *vadv_launch, double *seconds);
/* A fixed set of subdomains advanced `steps` steps per run: device tables
 * are built once; with use_graph != 0 the whole step sequence is captured
 * once in a hipGraph and replayed.  After a run subdomain i holds its state
 * in u[(parity_i + steps) % 2]. */
int sb_batch_plan_create(int prec, const sb_subdomain *subs, int count, int steps,
                         const sb_launch *hdiff_launch, const sb_launch *vadv_launch,
                         void **plan);
int sb_batch_plan_run(void *plan, int use_graph, double *seconds);
int sb_batch_plan_destroy(void *plan);

#ifdef __cplusplus
}
#endif

#endif /* SBMI355X_H */
