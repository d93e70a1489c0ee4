/*
 * bolt.h -- C ABI of the MI355X lattice update (libbolt_hip.so).
 *
 * This is the boundary a host program binds in place of the reference's OpenCL layer:
 * the device/context/queue/buffer code of src/opencl.rs and the per-step kernel dispatch
 * of src/simulation.rs (djbarker/boltzmann).  Host state -- the host arrays, the
 * iteration counter, the first-iteration equilibrium, checkpoints -- stays with the
 * caller, as it does in the reference's Simulation (simulation.rs:130-192).
 *
 * Conventions
 *  - Every function returns BOLT_OK (0) or a BOLT_E* code; bolt_last_error() then holds a
 *    message for the calling thread.  (The reference panics instead, opencl.rs:109-120.)
 *  - Host arrays are the reference's layout: row-major over the cell counts with x
 *    slowest, per-cell components last (f[..., q], vel[..., d]).  Device layout is private.
 *  - A lattice is used by one thread at a time; distinct lattices may run concurrently.
 *  - No callbacks, no global state besides the per-thread error message.
 */
#ifndef BOLT_H
#define BOLT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define BOLT_OK 0
#define BOLT_EINVAL 1     /* bad argument (odd step count, wrong size, unknown field) */
#define BOLT_ENOTFOUND 2  /* no kernel for this (D, Q, equation, collision, storage) */
#define BOLT_EDEVICE 3    /* HIP runtime error */
#define BOLT_ENOMEM 4     /* device allocation failed */
#define BOLT_ESTATE 5     /* call not valid in the current state */

/* ---- enumerations ---------------------------------------------------------------- */
/* device kinds, as the reference's DeviceType (opencl.rs:20-33) */
#define BOLT_DEVICE_CPU 0 /* not provided by this library: returns BOLT_ENOTFOUND */
#define BOLT_DEVICE_GPU 1

/* collision operators, as Omega::BGK / Omega::TRT (fields.rs:37-41) */
#define BOLT_BGK 0
#define BOLT_TRT 1

/* population storage */
#define BOLT_STORE_F32 0   /* float; equilibrium evaluated as the reference does   */
#define BOLT_STORE_F16S 1  /* binary16 of f_i - w_i, float arithmetic              */
#define BOLT_STORE_BF16S 2 /* bfloat16 of f_i - w_i, float arithmetic              */
#define BOLT_STORE_F64 3   /* double storage and arithmetic                        */
#define BOLT_STORE_F32F 4  /* float storage, plain float arithmetic                */

/* fields (the reference's Data<T,D> buffers, fields.rs:55-214, simulation.rs:22-59) */
#define BOLT_FIELD_FLAGS 0 /* int32 [cells]       cell flags (simulation.rs:80-89)  */
#define BOLT_FIELD_F 1     /* [cells, Q]          fluid populations                 */
#define BOLT_FIELD_RHO 2   /* [cells]             fluid density                     */
#define BOLT_FIELD_VEL 3   /* [cells, D]          fluid velocity                    */
#define BOLT_FIELD_G 4     /* [cells, Qt]         tracer populations                */
#define BOLT_FIELD_C 5     /* [cells]             tracer value                      */

/* host element types for bolt_lattice_write / bolt_lattice_read */
#define BOLT_DT_F32 0
#define BOLT_DT_F64 1
#define BOLT_DT_I32 2

/* reductions, summed in double on the device */
#define BOLT_REDUCE_MASS 0      /* sum of every population slot, walls included: 1 value */
#define BOLT_REDUCE_MOMENTUM 1  /* sum over non-wall cells of rho*u: D values          */
#define BOLT_REDUCE_TRACER 2    /* sum of every tracer population slot: 1 value       */

/* post-processing of the velocity (bolt_lattice_postproc) */
#define BOLT_POST_VMAG 0    /* |u| per cell                                              */
#define BOLT_POST_CURL 1    /* 2D: du/dy - dv/dx per cell; 3D: 3 components per cell      */
#define BOLT_POST_STREAM 2  /* 2D stream function, zero at the origin cell               */

typedef struct bolt_ctx bolt_ctx;
typedef struct bolt_lattice bolt_lattice;

/* Message for the last failed call on this thread ("" if none). */
const char* bolt_last_error(void);

/* Library version string, e.g. "bolt-hip 0.1 gfx950". */
const char* bolt_version(void);

/* ---- context: replaces OpenCLCtx::new / DeviceType (opencl.rs:20-33, 93-171) -------- */
/* device_kind: BOLT_DEVICE_GPU.  device_index: HIP ordinal.  Creates one in-order stream. */
int bolt_ctx_create(int device_kind, int device_index, bolt_ctx** out);
int bolt_ctx_destroy(bolt_ctx* ctx);
/* "<vendor> <GPU|CPU> <name>", the reference's device_info getter (lib.rs:205-215). */
int bolt_ctx_device_info(const bolt_ctx* ctx, char* buf, size_t len);

/* ---- lattice: the device buffers of one Simulation (Cells + Fluid + Scalars) ---------
 * Replaces Cells::new / Fluid::new / Scalar::new device allocation (simulation.rs:32-42,
 * fields.rs:68-82,150-163) and the kernel registry lookup make_kernel_key
 * (fields.rs:24-34, opencl.rs:123-161).  Supported (D, Q): (2, 9), (3, 27) with BGK or TRT
 * (the reference has no D3Q27-TRT kernel; this library adds it).  Device buffers start as
 * the reference's defaults: f = 0, rho = 1, vel = 0, flags = 0. */
int bolt_lattice_create(bolt_ctx* ctx, int ndim, const int64_t* counts, int q, int collision,
                        float omega_pos, float omega_neg, int storage, bolt_lattice** out);
int bolt_lattice_destroy(bolt_lattice* lat);

/* Adds an advection-diffusion tracer, Scalar::new + Simulation::add_tracer
 * (fields.rs:150-163, simulation.rs:215-223).  q: 5 in 2D, 7 in 3D; BGK only, as in the
 * reference (a TRT tracer has no kernel there either).  C = 0, g = 0 initially. */
int bolt_lattice_add_tracer(bolt_lattice* lat, int q, float omega, int* tracer_id);

/* Constant acceleration g[ndim], Simulation::set_gravity (simulation.rs:194-196). */
int bolt_lattice_set_gravity(bolt_lattice* lat, const float* g);

/* Boussinesq coupling acc = g * (alpha * (C_tracer - c0)), add_boussinesq_coupling
 * (simulation.rs:201-213).  As in the reference, couplings overwrite the acceleration each
 * step, the last one added wins, and any coupling overrides gravity (simulation.rs:326-373). */
int bolt_lattice_add_coupling(bolt_lattice* lat, int tracer_id, float alpha, float c0,
                              const float* g);

/* Host -> device copy of a whole field, Data::enqueue_write (opencl.rs:253-266).
 * tracer_id selects the tracer for BOLT_FIELD_G / BOLT_FIELD_C, ignored otherwise.
 * nbytes must equal elements * sizeof(host type). */
int bolt_lattice_write(bolt_lattice* lat, int field, int tracer_id, int host_dtype,
                       const void* host, size_t nbytes);

/* Device -> host copy of a whole field, Data::enqueue_read (opencl.rs:268-281).
 * Synchronous: returns once the host buffer holds the data. */
int bolt_lattice_read(bolt_lattice* lat, int field, int tracer_id, int host_dtype, void* host,
                      size_t nbytes);

/* bolt_lattice_read without any device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory -- tight serving loops that read the
 * fields back after every step call, code that must not call hipMalloc / hipFree (which
 * bolt_lattice_read does for its 64 MiB staging buffer on every call).  The field is
 * staged through a fixed 16 MiB area that each GPU holds for the life of the process;
 * larger fields move through it in pieces.  All calls on one GPU share that area: calls
 * that overlap on one GPU (several threads or contexts) race and must be serialised by
 * the caller.  Same arguments, checks and result as bolt_lattice_read. */
int bolt_lattice_read_noalloc(bolt_lattice* lat, int field, int tracer_id, int host_dtype,
                              void* host, size_t nbytes);

/* Device-side equilibrium initialisation: f = feq(rho, vel) and, for every tracer,
 * g = geq(C, vel), in float with the reference's host formula (velocities.rs:43-52). */
int bolt_lattice_equilibrate(bolt_lattice* lat);

/* Runs n time steps (n even), the per-step dispatch of Simulation::iterate
 * (simulation.rs:297-433): Boussinesq acceleration -> fluid update -> tracer updates.
 * rho/vel/C hold the last step's values when the call returns.  Asynchronous with respect
 * to the host; bolt_lattice_read and bolt_lattice_sync wait. */
int bolt_lattice_step(bolt_lattice* lat, uint64_t n);

/* As bolt_lattice_step, and reports the device time of the n steps in milliseconds,
 * measured with HIP events on the lattice's stream.  Synchronous. */
int bolt_lattice_step_timed(bolt_lattice* lat, uint64_t n, double* ms);

/* Waits for all work queued on the lattice's stream (queue.finish()). */
int bolt_lattice_sync(bolt_lattice* lat);

/* Double-precision reductions (see BOLT_REDUCE_*); out receives 1 or D values. */
int bolt_lattice_reduce(bolt_lattice* lat, int what, int tracer_id, double* out);

/* Steady-state residual in double precision (no counterpart in the reference):
 *   r = max_x |u - u_prev| / (steps_between * max_x |u|)  over non-wall cells,
 * u_prev being the velocity saved by the previous call (which this call replaces).  The
 * first call only saves the snapshot and returns +inf.  Reads the velocity of the last
 * completed bolt_lattice_step. */
int bolt_lattice_residual(bolt_lattice* lat, uint64_t steps_between, double* out);

/* Page-locked host memory for a caller's field arrays, so that bolt_lattice_read and
 * bolt_lattice_write move them by DMA at the link's full rate (the reference's host arrays,
 * lib.rs:57-67, are ordinary memory; this has no counterpart there).  bolt_host_free(NULL)
 * does nothing. */
int bolt_host_alloc(size_t bytes, void** out);
int bolt_host_free(void* p);

/* Number of non-finite values of the device rho (BOLT_FIELD_RHO) or velocity (BOLT_FIELD_VEL)
 * of the last completed bolt_lattice_step.  Replaces, on the device, the host check of
 * python/boltzmann/simulation.py:241-242 (np.any(~np.isfinite(sim.fluid.vel)) after each
 * iterate of run_sim); it has no counterpart in the reference's native code. */
int bolt_lattice_count_nonfinite(bolt_lattice* lat, int field, uint64_t* count);

/* Post-processing of the device velocity into a host float array, with the semantics of
 * python/boltzmann/utils/postproc.py:12-127 on the float32 velocity (np.gradient: central
 * differences inside, one-sided at the ends of each axis; sequential float32 cumsum), in
 * lattice units (the caller applies unit scales):
 *   BOLT_POST_VMAG    calc_vmag: sqrt(sum_a u_a^2), [cells]
 *   BOLT_POST_CURL    calc_curl: 2D du/dy - dv/dx [cells]; 3D [cells][3]
 *   BOLT_POST_STREAM  calc_stream_func (2D): cumsum(u[0, :]) - cumsum(v, axis 0), minus its
 *                     value at origin[2] (NULL: cell (0, 0)), [cells]
 * Reads the velocity of the last completed bolt_lattice_step or bolt_lattice_write. */
int bolt_lattice_postproc(bolt_lattice* lat, int kind, const int* origin, float* host,
                          size_t nbytes);

/* Device bytes held by the lattice's buffers. */
int bolt_lattice_device_bytes(const bolt_lattice* lat, uint64_t* out);

/* Launch tuning: work-group size along the contiguous axis (64, 128, 192 or 256) of the
 * one-cell-per-lane steps; 0 restores the built-in choice. */
int bolt_lattice_set_block(bolt_lattice* lat, int block);
/* Launch tuning: cells of the contiguous axis per lane in the fluid step (1, 2 or 4; 2 at
 * most for double storage; the axis must be a multiple of it), on every step; 0 restores
 * the built-in choice, measured fastest on MI355X: 2 cells per lane on the odd (own-cell)
 * steps of the 2-byte formats, except D2Q9 lattices that fit the Infinity Cache (see
 * bolt_lattice_set_cache_policy), one cell per lane otherwise.  Wider lanes make every
 * population access an aligned 4-, 8- or 16-byte vector and do the streaming shift along
 * the row with lane rotations.  A tracer fused into the fluid step always runs one cell per
 * lane.  Results do not depend on the choice (bit-identical). */
int bolt_lattice_set_vector_width(bolt_lattice* lat, int vw);
/* Launch tuning (no counterpart in the reference, whose kernels take one cell per
 * work-item): whether the even (streaming) step runs row-staged -- one work-group per row of
 * the contiguous axis, every population access an aligned row, the shift along the row
 * taken through LDS -- or addresses each neighbour's slot directly.  The row-staged step can
 * run when one cell per lane is chosen for the even step, the row has at most 1024 cells
 * and its shifted planes' rows (fluid and fused tracer) fit 64 KiB of LDS.
 *   BOLT_ROW_ON    wherever it can run
 *   BOLT_ROW_OFF   never
 *   BOLT_ROW_AUTO  (the default) where it measured faster: rows that fill whole 128-byte
 *                  cache lines, except D3Q27 rows over 512 cells in the 2-byte formats and
 *                  D2Q9 lattices that fit the Infinity Cache
 * Results do not depend on the choice (bit-identical). */
#define BOLT_ROW_OFF 0
#define BOLT_ROW_ON 1
#define BOLT_ROW_AUTO 2
int bolt_lattice_set_row_staging(bolt_lattice* lat, int mode);
/* Launch tuning (no counterpart in the reference): how the steps load the populations.
 *   BOLT_CACHE_STREAM  streaming (non-temporal) loads, which leave the lines out of the GPU's
 *                      caches: the better form when a step's working set exceeds the 256 MB
 *                      Infinity Cache of MI355X
 *   BOLT_CACHE_KEEP    ordinary loads: the lines stay in the Infinity Cache for the next step,
 *                      1.1-1.4x faster where the working set fits it
 *   BOLT_CACHE_AUTO    (the default) KEEP when a step's working set is at most 240 MiB
 *                      (populations, flags, tracer values and any per-step density and
 *                      velocity), STREAM otherwise
 * Results do not depend on the choice (bit-identical). */
#define BOLT_CACHE_STREAM 0
#define BOLT_CACHE_KEEP 1
#define BOLT_CACHE_AUTO 2
int bolt_lattice_set_cache_policy(bolt_lattice* lat, int policy);

/* ---- batches: many small 2D Rayleigh-Benard-type lattices in one launch per step -----
 * No counterpart in the reference (it runs one Simulation per process).  Instance k is a
 * D2Q9 fluid (counts[2k], counts[2k+1]; omega_f[2k] = w+, omega_f[2k+1] = w-, ignored for
 * BGK) with one D2Q5 tracer (omega_g[k]) coupled through acc = g * (alpha * (C - c0))
 * (alpha[k], c0[k], g[2k..2k+1]): the reference's set_boussinesq_acc_2d + update_d2q9_* +
 * update_d2q5_bgk per step (simulation.rs:353-432).  With omega_g == NULL the instances are
 * fluid only (alpha, c0 and g are then ignored).  Fields are the lattice ones
 * (BOLT_FIELD_*; the tracer is implicit).  Every instance evolves exactly as it would alone. */
typedef struct bolt_batch bolt_batch;
int bolt_batch_create(bolt_ctx* ctx, int n, const int64_t* counts, int collision,
                      const float* omega_f, const float* omega_g, const float* alpha,
                      const float* c0, const float* g, bolt_batch** out);
int bolt_batch_destroy(bolt_batch* b);
int bolt_batch_write(bolt_batch* b, int inst, int field, int host_dtype, const void* host,
                     size_t nbytes);
int bolt_batch_read(bolt_batch* b, int inst, int field, int host_dtype, void* host,
                    size_t nbytes);
int bolt_batch_equilibrate(bolt_batch* b);
/* n even; one launch per step over all instances, replayed from a captured HIP graph */
int bolt_batch_step(bolt_batch* b, uint64_t n);
int bolt_batch_step_timed(bolt_batch* b, uint64_t n, double* ms);
int bolt_batch_reduce(bolt_batch* b, int inst, int what, double* out);
int bolt_batch_cells(const bolt_batch* b, uint64_t* out);
/* bolt_lattice_residual of one instance */
int bolt_batch_residual(bolt_batch* b, int inst, uint64_t steps_between, double* out);
/* bolt_lattice_residual of every active instance in one launch (out[n]; an instance's
 * snapshot is shared with bolt_batch_residual); NaN for inactive instances. */
int bolt_batch_residuals(bolt_batch* b, uint64_t steps_between, double* out);
/* launch shape: work-group width along the contiguous axis and steps per captured graph */
int bolt_batch_set_shape(bolt_batch* b, int block, int graph_steps);
/* Even steps row-staged, one work-group per (instance, row), as
 * bolt_lattice_set_row_staging: BOLT_ROW_ON whenever every active row has at most 1024 cells
 * (one launch per group of rows of the same length rounded up to 64), BOLT_ROW_AUTO (the
 * default) when in addition every active row fills whole 128-byte cache lines, BOLT_ROW_OFF
 * never (every step through the tiles of set_shape).  bolt_batch_tune tries on and off.
 * Results do not depend on the choice. */
int bolt_batch_set_row_staging(bolt_batch* b, int mode);
/* Population loads as bolt_lattice_set_cache_policy, for the working set of all active
 * instances together; bolt_batch_tune tries both forms. */
int bolt_batch_set_cache_policy(bolt_batch* b, int policy);
/* instances that later steps advance (active[n], non-zero: advance); an inactive instance
 * keeps its state and can still be read, reduced or checked for its residual.  Change it
 * only between step calls (each of an even count, so every instance stays in its natural
 * population layout). */
int bolt_batch_set_active(bolt_batch* b, const uint8_t* active, int n);
/* autotuner: times the candidate shapes on this batch (state restored afterwards) -- work-
 * group width 64/128/256, 20/100/400 steps per graph, even steps row-staged or not,
 * streaming or ordinary population loads --, keeps the fastest and writes a per-candidate
 * report into report[len]. */
int bolt_batch_tune(bolt_batch* b, int steps, char* report, size_t len);

/* ---- warm-start network (no counterpart in the reference) ----------------------------
 * A U-Net of `depth` 2x2 poolings predicting (rho, u) of a steady channel flow from per-cell
 * features, trained and evaluated with this library's bf16 MFMA kernels (float accumulation,
 * float master weights, Adam).  Host arrays are channels-last: inputs [n][nx][ny][nfeat],
 * outputs [n][nx][ny][3]; nx and ny multiples of 2^depth.  Parameters are flattened per
 * layer (encoder pairs by level, bottleneck pair, decoder from the deepest level, 1x1
 * output) as W[cout][taps*cin] (tap = kx*3 + ky for 3x3 layers) followed by b[cout]. */
typedef struct bolt_nn bolt_nn;
int bolt_nn_create(bolt_ctx* ctx, int nx, int ny, int batch, int width, int depth, uint64_t seed,
                   bolt_nn** out);
int bolt_nn_destroy(bolt_nn* net);
/* Feature channels per cell of the host inputs (1..16, default 8); set before
 * bolt_nn_set_data, bolt_nn_loss_grad and bolt_nn_predict, whose x arrays are then
 * [n][nx][ny][nfeat]. */
int bolt_nn_set_features(bolt_nn* net, int nfeat);
int bolt_nn_features(const bolt_nn* net, int* out);
int bolt_nn_param_count(const bolt_nn* net, uint64_t* out);
int bolt_nn_get_params(bolt_nn* net, float* out, uint64_t n);
int bolt_nn_set_params(bolt_nn* net, const float* in, uint64_t n);
/* one batch (exactly the batch size): mean squared error and its parameter gradients */
int bolt_nn_loss_grad(bolt_nn* net, const float* x, const float* y, double* loss,
                      float* grads, uint64_t n);
/* uploads n training pairs to the device */
int bolt_nn_set_data(bolt_nn* net, int n, const float* x, const float* y);
/* minibatch Adam with cosine learning-rate decay; losses[epochs] receives the epoch means */
int bolt_nn_train(bolt_nn* net, int epochs, float lr, uint64_t seed, float* losses);
/* epochs [first, first+epochs) of a run of `total` epochs: the same shuffles and cosine
 * learning rates as those epochs of one bolt_nn_train(total) call (progress reporting) */
int bolt_nn_train_part(bolt_nn* net, int first, int epochs, int total, float lr, uint64_t seed,
                       float* losses);
int bolt_nn_predict(bolt_nn* net, int n, const float* x, float* y);
/* test hook: C[M][N] = A.B^T through the network's GEMM (a_form 0: A[m][k], 1: A[k][m];
 * b_form 0: B[n][k], 1: B[k][n]); k_split > 0 splits K with atomic accumulation */
int bolt_nn_gemm_check(bolt_ctx* ctx, int a_form, int b_form, int M, int N, int K,
                       const float* a, const float* b, float* c, int k_split);

#ifdef __cplusplus
}
#endif

#endif /* BOLT_H */
