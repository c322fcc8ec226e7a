/*
 * appfis_hip.h -- C ABI of libappfis_hip.so, the MI355X (gfx950) backend of the
 * APPFIS iterative-stencil runtime.
 *
 * Conventions
 *   - Every entry point returns an int status (APPFIS_OK == 0).  Nothing ever
 *     calls exit(); the message of the last failure on the calling host thread
 *     is available from appfis_last_error().
 *   - Arrays are dense, row-major, x fastest, exactly like APPFIS::Array
 *     (Appfis/Array.hpp:579-590 of the reference):  idx = i + dimX*(j + dimY*k).
 *     "dims" are the padded extents (interior + 2*ghost per axis).
 *   - Device pointers come from appfis_malloc (or any hipMalloc'd memory).
 *   - All kernels are enqueued on the context's single HIP stream and are
 *     asynchronous unless a host result is returned (then the call waits).
 *   - One context per host thread; a context is not thread-safe.
 *   - Faces follow APPFIS::DIRECTION (reference Appfis/Env.hpp:30):
 *     LEFT=0 RIGHT=1 UP=2 DOWN=3 FRONT=4 BACK=5.  In 2D UP/DOWN are the y faces;
 *     in 3D UP/DOWN are the z faces and FRONT/BACK the y faces.
 *   - Reduction ops follow APPFIS::REDUCTION (reference Appfis/Env.hpp:32):
 *     MAX=0 MIN=1 SUM=2.
 *   - Graph capture.  While the context's stream is captured into a graph
 *     (appfis_graph_begin, or the caller's own hipStreamBeginCapture on
 *     appfis_ctx_stream), the entry points that only enqueue kernels can be captured:
 *     appfis_exchange, appfis_set_boundary(_copy), appfis_fill_*, appfis_memcpy_d2d,
 *     appfis_reduce_device_noalloc, appfis_reduce_device (once an earlier uncaptured
 *     call on the context has sized its scratch), and appfis_heat2d / appfis_gol /
 *     appfis_fdtd3d in every variant (the CHOSEN forms once an uncaptured call has tuned
 *     the shape and sized the workspace).  Everything that would allocate, free,
 *     synchronise, memset, return a host result or capture a graph of its own
 *     (appfis_malloc/free/memset/memcpy_h2d/d2h/sync/timer_stop, appfis_kernel_time,
 *     appfis_kernel_timing's first enabling call, appfis_reduce, appfis_reduce_sum_wide,
 *     appfis_himeno, appfis_heat2d_ensemble, the flood entry points, a first CHOSEN
 *     call) returns APPFIS_ERR_UNSUPPORTED under capture and enqueues nothing.
 */
#ifndef APPFIS_HIP_H
#define APPFIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APPFIS_ABI_VERSION 2

/* status codes */
#define APPFIS_OK 0
#define APPFIS_ERR_INVALID 1     /* bad argument (shape, dtype, face, null pointer) */
#define APPFIS_ERR_HIP 2         /* a HIP runtime call failed */
#define APPFIS_ERR_UNSUPPORTED 3 /* valid request that the library does not implement */
#define APPFIS_ERR_NOMEM 4       /* device allocation failed */

/* element types */
#define APPFIS_U8 0
#define APPFIS_I32 1
#define APPFIS_F32 2
#define APPFIS_F64 3
#define APPFIS_F16 4
#define APPFIS_BF16 5
#define APPFIS_I64 6

/* reductions (APPFIS::REDUCTION) */
#define APPFIS_MAX 0
#define APPFIS_MIN 1
#define APPFIS_SUM 2

/* kernel variants of the hand-written config kernels */
#define APPFIS_PLAIN 0   /* one sweep per launch, reference dtype and operation order */
#define APPFIS_CHOSEN 1  /* temporal blocking / packed / mixed-precision form */
/* APPFIS_CANDIDATE0 + i: candidate i of an autotuned kernel's table, run as is (for
 * tests and tuning studies; appfis_*_candidates lists the table).  Every candidate
 * computes the same values as PLAIN. */
#define APPFIS_CANDIDATE0 2

typedef struct appfis_ctx appfis_ctx;

/* ---- library / context (replaces Topology::Initialize device+stream setup,
 *      reference Appfis/Topology.hpp:124-151) ---------------------------------- */
int appfis_abi_version(void);
const char* appfis_status_string(int status);
const char* appfis_last_error(void);
int appfis_device_count(int* count);
int appfis_ctx_create(int device, appfis_ctx** out);
int appfis_ctx_destroy(appfis_ctx* ctx);
int appfis_ctx_device(appfis_ctx* ctx, int* device);
void* appfis_ctx_stream(appfis_ctx* ctx); /* the hipStream_t, for header-side launches */
int appfis_sync(appfis_ctx* ctx);
int appfis_device_name(appfis_ctx* ctx, char* buf, int len);
/* Text of every autotuning decision taken on this context (candidate timings and the
 * chosen variant per kernel / dtype / shape); the CHOSEN variants tune on first use. */
int appfis_tune_report(appfis_ctx* ctx, char* buf, int len);

/* ---- memory (replaces Grid::CreateSubGridsGPU / CopySubGridsGPUToCPU,
 *      reference Appfis/Grid.hpp:403-479) --------------------------------------- */
int appfis_malloc(appfis_ctx* ctx, size_t bytes, void** out); /* zero-filled, like Array::AllocateMemory (Array.hpp:339-342) */
int appfis_free(appfis_ctx* ctx, void* ptr);
int appfis_memcpy_h2d(appfis_ctx* ctx, void* dst, const void* src, size_t bytes);
int appfis_memcpy_d2h(appfis_ctx* ctx, void* dst, const void* src, size_t bytes);
int appfis_memcpy_d2d(appfis_ctx* ctx, void* dst, const void* src, size_t bytes);
int appfis_memset(appfis_ctx* ctx, void* dst, int value, size_t bytes);

/* ---- timing on the context stream (hipEvents) ---------------------------------- */
/* The two events carry no system-scope fence: the interval is the stream's work alone,
 * without a cache writeback at either end (read results through appfis_memcpy_d2h,
 * which synchronises on its own). */
int appfis_timer_start(appfis_ctx* ctx);
int appfis_timer_stop(appfis_ctx* ctx, float* ms); /* waits for the stop event */
/* Timing of a call's generation kernel (a measurement aid; no reference counterpart).
 * With on != 0, each later appfis_gol call records a hipEvent on the context stream
 * before the first and after the last launch of its generation kernel -- the byte
 * kernel of the PLAIN form; for the bit-packed form the temporally blocked (or
 * one-generation) kernel, which advances generations 1 .. gens-1 between the packing
 * and the fused last generation, or, for a full-row variant and gens >= 2, all gens
 * generations in two or more launches, the first packing and the last unpacking the
 * uint8 field -- and appfis_kernel_time gives, for the most recent such call, the time
 * between the two events, the number of launches and the generations they advanced
 * (it waits for the second event).  Nothing is recorded while the stream is being
 * captured. */
int appfis_kernel_timing(appfis_ctx* ctx, int on);
int appfis_kernel_time(appfis_ctx* ctx, double* ms, int* launches, int* generations);

/* ---- hipGraph capture of the context stream -------------------------------------- */
int appfis_graph_begin(appfis_ctx* ctx);
int appfis_graph_end(appfis_ctx* ctx, void** graph_exec);
int appfis_graph_launch(appfis_ctx* ctx, void* graph_exec, int times);
int appfis_graph_destroy(appfis_ctx* ctx, void* graph_exec);

/* ---- seeded synthetic inputs ------------------------------------------------------
 * r(seed, n) = mix64(seed + (n+1)*0x9E3779B97F4A7C15)  (splitmix64, Steele et al. 2014),
 * n = interior linear index (x fastest, interior extents, no ghost).
 * uniform: lo + (hi-lo)*u, u = (r>>11)*2^-53 for F64, (r>>40)*2^-24 for F32/F16/BF16
 *          (then rounded to the storage type); ghost cells are not written.
 * bernoulli: cell = (r < p*2^64) ? 1 : 0 for U8. */
int appfis_fill_uniform(appfis_ctx* ctx, void* data, int dtype, int ndim, const int* dims, int ghost,
                        uint64_t seed, double lo, double hi);
int appfis_fill_bernoulli_u8(appfis_ctx* ctx, uint8_t* data, int ndim, const int* dims, int ghost,
                             uint64_t seed, double p);
int appfis_fill_value(appfis_ctx* ctx, void* data, int dtype, size_t count, double value);

/* ---- boundary fill (replaces Grid::SetBoundary, reference Appfis/Grid.hpp:558-678,
 *      and Grid::SetBoundaryCopy, Grid.hpp:682-830) -------------------------------
 * face = -1 applies every face in the reference order (LEFT, RIGHT, UP, DOWN[, FRONT, BACK]).
 * value points at one host element of dtype. */
int appfis_set_boundary(appfis_ctx* ctx, void* data, int dtype, int ndim, const int* dims, int ghost,
                        int face, const void* value);
int appfis_set_boundary_copy(appfis_ctx* ctx, void* data, int dtype, int ndim, const int* dims,
                             int ghost, int face);

/* ---- halo exchange at MPI world size 1 (replaces Communication::Exchange/ExchangeGPU,
 *      reference Appfis/Communication.hpp:613-748, with the face pack/unpack of
 *      Runtime.hpp:917-1173).  periodic != 0: the ghost layers of every axis are
 *      refreshed from the opposite interior layers, axis X first then Y then Z, each
 *      axis over the full extent of the others, so edges and corners come out as in
 *      the reference CPU path.  periodic == 0: no neighbour exists; nothing moves. */
int appfis_exchange(appfis_ctx* ctx, void* data, int dtype, int ndim, const int* dims, int ghost,
                    int periodic);

/* ---- reductions over a box (replaces Execute{1,2,3}DReduce, reference
 *      Appfis/Runtime.hpp:237-255,598-913 and Runtime_CUDA.hpp:554-740) ---------------
 * box = {startX,endX,startY,endY,startZ,endZ} (half-open, like APPFIS::Config).
 * appfis_reduce: result has the array's dtype (written to host memory `result`);
 *   MAX starts from numeric_limits<T>::min() and MIN from numeric_limits<T>::max(),
 *   as the reference does.  SUM is accumulated in a wide type (int64 for integers,
 *   compensated FP64 for floats) and rounded to T once at the end.
 * appfis_reduce_sum_wide: SUM only, result in int64 (U8/I32) or double (floats).
 * appfis_reduce_device: result written to device memory `dev_result` (dtype of the
 *   array for MAX/MIN; int64 or double for SUM); no host synchronisation. */
int appfis_reduce(appfis_ctx* ctx, const void* data, int dtype, int ndim, const int* dims,
                  const int* box, int op, void* result);
int appfis_reduce_sum_wide(appfis_ctx* ctx, const void* data, int dtype, int ndim, const int* dims,
                           const int* box, void* result);
int appfis_reduce_device(appfis_ctx* ctx, const void* data, int dtype, int ndim, const int* dims,
                         const int* box, int op, void* dev_result);

/* appfis_reduce_device_noalloc: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same reduction and result as
 * appfis_reduce_device, but its scratch (a ticket word and at most 2048 per-block
 * partials, APPFIS_NOALLOC_REDUCE_SCRATCH_BYTES in all) is a fixed device array that
 * each GPU holds from library load until the process ends, and the call allocates
 * nothing, synchronises nothing and records no memset (the ticket is cleared by a
 * kernel), so it can be captured into a graph.  The scratch use does not grow with the
 * box, so any box that appfis_reduce_device accepts is accepted.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (two streams, two contexts,
 * two host threads) race and give wrong results; serialise them. */
#define APPFIS_NOALLOC_REDUCE_SCRATCH_BYTES (256 + 2048 * 16)
int appfis_reduce_device_noalloc(appfis_ctx* ctx, const void* data, int dtype, int ndim, const int* dims,
                                 const int* box, int op, void* dev_result);

/* ---- HeatDiffusion 2D (config form of the reference functor, HeatDiffusion/src/
 *      heat_appfis.hpp:166-174, driven as in heat_appfis.hpp:86-91) -----------------
 * u' = u + c*(uN + uS + uE + uW - 4u) over the nx*ny interior, ghost 1, double
 * buffered.  bufs[0] holds the field at entry; after `iters` sweeps the field is in
 * bufs[iters % 2] (the Grid::Swap parity).  dtype F64 or F32.
 * variant PLAIN: one sweep per launch.  CHOSEN: `tsteps` sweeps per launch with an
 * LDS-resident tile (temporal blocking); tsteps <= 0 picks the tuned depth. */
int appfis_heat2d(appfis_ctx* ctx, int dtype, void* buf0, void* buf1, int nx, int ny, double c,
                  int iters, int variant, int tsteps);

/* Temporal-blocking candidates of appfis_heat2d CHOSEN for `dtype` (variant
 * APPFIS_CANDIDATE0 + i), one name per line; returns their number (< 0: bad dtype). */
int appfis_heat2d_candidates(int dtype, char* names, int len);

/* Heat2D ensemble: `ngrid` independent grids of any sizes (grid g: interior
 * nx[g] x ny[g], buffers buf0[g] / buf1[g], padded as above) advance `iters` sweeps
 * together: one temporal-blocking launch per chunk of sweeps over a tile work-list that
 * spans every grid, the whole launch chain captured and replayed as one hipGraph.
 * Values are identical to appfis_heat2d on each grid; grid g ends in its buffer
 * iters % 2.  Pointer arrays are host arrays.  Waits for completion. */
int appfis_heat2d_ensemble(appfis_ctx* ctx, int dtype, int ngrid, const int* nx, const int* ny,
                           void* const* buf0, void* const* buf1, double c, int iters);

/* ---- GameOfLife 2D, uint8, periodic, rule B3/S23 (config form of GameOfLife/src/
 *      game_appfis.hpp:121-135 plus the periodic exchange of Communication.hpp:613-678).
 * bufs as for heat2d (ghost 1, padded (nx+2)*(ny+2)).  The ghost ring of the
 * written buffer is refreshed every generation.  PLAIN: one generation per launch;
 * CHOSEN (nx % 32 == 0): the generations run on a bit-packed copy of the field (32
 * cells per word), both buffers are left exactly as PLAIN leaves them; other widths
 * run the PLAIN kernel.  `tsteps` is reserved (pass 0).  Cells must hold 0 or 1.
 * APPFIS_CANDIDATE0 + i: the bit-packed form with blocked candidate i forced. */
int appfis_gol(appfis_ctx* ctx, uint8_t* buf0, uint8_t* buf1, int nx, int ny, int gens,
               int variant, int tsteps);

/* Blocked kernels of the bit-packed GoL form (variant APPFIS_CANDIDATE0 + i forces
 * candidate i; it needs nx a multiple of 2048 spanned by its block), one name per
 * line; returns their number. */
int appfis_gol_candidates(char* names, int len);

/* ---- FDTD3d radius-4 star (config form of FDTD3d/src/fdtd3d_appfis.hpp:129-143):
 * u' = c[0]*u + sum_{r=1..4} c[r]*(6 neighbours at distance r); ghost 4 held at 0.
 * dtype F32 / F64 (arithmetic in that type) or F16 / BF16 (storage only, FP32
 * arithmetic).  Field in bufs[steps % 2] at exit.  PLAIN runs a fixed tiling, one
 * launch per sweep: per dtype a preferred one where the shape and buffers admit it (F32:
 * direct-to-LDS, needs (nx + 8) % 4 == 0 and 16-byte aligned buffers; F64: two points
 * per lane, nx even; F16 / BF16: two points per lane in 32-bit pairs, nx even and
 * 4-byte aligned buffers), else one that fits any shape;
 * CHOSEN runs the tiling the autotuner picked for this dtype and shape (timed once on
 * scratch buffers, cached in the context, among the candidates the shape and the
 * buffers' alignment admit); APPFIS_CANDIDATE0 + i runs candidate i of
 * appfis_fdtd3d_candidates -- all compute identical values.  A forced candidate that
 * does not fit fails with APPFIS_ERR_INVALID: "vxN" needs nx % N == 0 (16-bit pairs also
 * 4-byte aligned buffers); "dma ..." (FP32 direct-to-LDS) needs (nx + 8) % 4 == 0 and
 * 16-byte aligned buffers. */
int appfis_fdtd3d(appfis_ctx* ctx, int dtype, void* buf0, void* buf1, int nx, int ny, int nz,
                  const double* coeff5, int steps, int variant);

/* Candidate tilings of appfis_fdtd3d for `dtype`, one name per line, into names[len]
 * (may be NULL); returns their number, < 0 for a bad dtype.  Needs no context. */
int appfis_fdtd3d_candidates(int dtype, char* names, int len);

/* ---- Himeno 19-point Jacobi (Himeno/src/himeno_appfis.hpp:178-198) with the
 * residual gosa = sum ss^2 over the interior after every sweep.
 * arrays (14 device pointers, all of `dtype`, dims (nx,ny,nz) padded incl. the
 * 1-cell boundary shell):  p0, p1, a0, a1, a2, a3, b0, b1, b2, c0, c1, c2, bnd, wrk1.
 * Field in p[iters % 2] at exit.  gosa_host (iters doubles) receives each sweep's
 * gosa.  variant PLAIN: dtype F32 or F64, reference operation order.
 * variant CHOSEN (defect correction): dtype F64 inputs; FP32 inner sweeps on the
 * correction, FP64 outer residual every `outer` sweeps (<=0: every 16).  Both tune
 * the sweep shape (planes per block, coefficient prefetch depth) on first use per
 * shape.  APPFIS_CANDIDATE0 + i runs candidate i of appfis_himeno_candidates: the
 * PLAIN form or the defect correction with that sweep shape forced (same values).
 * A plane of the grid must stay under 2 GiB.  */
int appfis_himeno(appfis_ctx* ctx, int dtype, void* const* arrays, int nx, int ny, int nz,
                  double omega, int iters, int variant, int outer, double* gosa_host);

/* Candidates of appfis_himeno ("plain zc.. d.." then "dc zc.. d.."), one name per
 * line, into names[len] (may be NULL); returns their number.  Needs no context. */
int appfis_himeno_candidates(char* names, int len);

/* ---- Triton shallow-water flood step chain (Triton/src/triton_appfis.hpp:316-386
 * with the functors of triton_appfis.hpp:567-1112) plus uniform rainfall.
 * One scenario: fields (device, padded (nx+2)*(ny+2), dtype F64 or F32):
 *   h, qx, qy (state, updated in place), z (dem, ghost = wall height), n2 (Manning n^2).
 * params: see appfis_flood_params.  Runs `steps` time steps; writes per-step
 * simulated time and FP64 Kahan volume (sum h * dx^2 over the interior) to the host
 * arrays time_host / volume_host when not NULL. */
typedef struct appfis_flood_params {
    double dx;            /* cell size (m) */
    double courant;       /* CFL number */
    double hextra;        /* wet/dry depth threshold (m) */
    double dt_fallback;   /* time step used while every cell is dry (arglist.time_step) */
    double print_interval;/* dt is clamped to the next multiple of this (<=0: never) */
    double rain_rate;     /* uniform rainfall (m/s), added as h += rain*dt after the update */
    double start_time;    /* simulated time at entry (s) */
    int time_increment_fixed; /* != 0: every step takes dt_fallback, no CFL and no print
                               * clamp (arglist.time_increment_fixed, triton_appfis.hpp:322) */
    double end_time;      /* > 0: run while simtime < end_time, as the reference's loop
                           * (triton_appfis.hpp:316), for at most `steps` steps; <= 0: exactly
                           * `steps` steps */
} appfis_flood_params;

int appfis_flood_run(appfis_ctx* ctx, int dtype, void* h, void* qx, void* qy, const void* z,
                     const void* n2, int nx, int ny, const appfis_flood_params* params, int steps,
                     int variant, double* time_host, double* volume_host);

/* Point sources of the reference Triton app (hydrograph inflows, Triton/src/inflow.hpp:
 * 93-136 and the functor of triton_appfis.hpp:1115-1141, applied after the update as at
 * triton_appfis.hpp:358-377).  Every step, the hydrograph interval holding the step's start
 * time is searched from the row the previous step stopped at; each source's flow is
 * interpolated linearly in time (the last row's flow once past it) and flow*dt/dx^2 is
 * added to the depth of its cell. */
typedef struct appfis_flood_sources {
    int count;            /* number of sources (0: none) */
    const int* cell;      /* count distinct interior cells, linear padded index i + (nx+2)*j */
    int rows;             /* hydrograph rows (>= 1 when count > 0) */
    const double* time;   /* rows times (s) */
    const double* flow;   /* rows x count flows (m^3/s), row-major: flow[r*count + k] */
} appfis_flood_sources;

/* appfis_flood_run with point sources (sources may be NULL; host tables; one scenario).
 * steps_done (may be NULL) receives the number of steps run: `steps`, or fewer when
 * params->end_time stopped the loop. */
int appfis_flood_run_src(appfis_ctx* ctx, int dtype, void* h, void* qx, void* qy, const void* z,
                         const void* n2, int nx, int ny, const appfis_flood_params* params,
                         const appfis_flood_sources* sources, int steps, int variant,
                         double* time_host, double* volume_host, int* steps_done);

/* Ensemble: nscen independent scenarios of different sizes advanced together, one
 * batched launch per phase over a tile work-list, the step loop captured in a
 * hipGraph.  Pointer arrays are host arrays of device pointers (one per scenario).
 * time_host / volume_host: nscen*steps doubles (scenario-major) or NULL. */
int appfis_flood_ensemble_run(appfis_ctx* ctx, int dtype, int nscen, const int* nx, const int* ny,
                              void* const* h, void* const* qx, void* const* qy,
                              const void* const* z, const void* const* n2,
                              const appfis_flood_params* params, int steps,
                              double* time_host, double* volume_host);

/* Synthetic DEM of the flood ensemble: sum of `nhills` Gaussian hills plus a planar
 * slope along x, written to the interior of z (F64 or F32); ghost set to `wall`.
 * Hill k (k < nhills): r_j = r(seed, j) for j = 4k..4k+3 ->
 *   amplitude 5 + 45*u0 (m), sigma 200 + 1800*u1 (m), centre (u2*L, u3*L), L = nx*dx. */
int appfis_flood_dem(appfis_ctx* ctx, int dtype, void* z, int nx, int ny, double dx, uint64_t seed,
                     int nhills, double slope, double wall);

#ifdef __cplusplus
}
#endif

#endif /* APPFIS_HIP_H */
