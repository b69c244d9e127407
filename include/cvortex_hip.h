/*
 * cvortex_hip.h -- HIP (gfx950) accelerator seam and device-resident API of
 * libcvortex.so.
 *
 * 1. The seam.  In the reference library every public M2M function first
 *    tries an int-returning OpenCL accelerator function and runs its own
 *    OpenMP CPU path when that returns -1 (reference src/P3D.cpp:352-364,
 *    src/F3D.cpp:169-178, src/P2D.cpp:150-160).  The cvtx_hip_* functions
 *    below have the same arguments as the public functions, run on the
 *    first enabled HIP device, and return 0 on success or -1 on failure
 *    (no device enabled, unsupported regularisation name, invalid argument,
 *    HIP error); cvtx_hip_last_error() then describes the failure.  They
 *    never run a CPU fallback themselves; the public void functions do
 *    (see cvtx_hip_set_dispatch).
 *
 * 2. Fused host-array entry points (no reference counterpart): one pass
 *    over the pairs produces several outputs.
 *
 * 3. The device-resident API: structure-of-arrays device buffers and a
 *    caller stream, for callers that keep particles in GPU memory (time
 *    stepping, batched clouds).  All pointers are device pointers; `stream`
 *    is a hipStream_t passed as void* (NULL = the library's stream for the
 *    device).  Calls are asynchronous.  Calls that need workspace (source
 *    partials, target ordering) share one grow-only workspace per device:
 *    a call on one stream waits (through an event) for the last such call
 *    on any other stream, so calls on different streams do not overlap.
 *    That ordering is not recorded into a stream being captured: captured
 *    calls that need workspace must not run concurrently with each other
 *    or with eager calls on the same device.
 *    A call captured into a graph keeps the workspace address it saw, and
 *    a later call that needs more workspace replaces that buffer; graphs
 *    should therefore be captured around a size that an earlier eager call
 *    has already reached, and replayed while no larger call intervenes.
 *    Capturing the first call of a size fails (the growth synchronises).
 *    The library's own RK4 graphs (cvtx_hip_batch_rk4_*) use buffers of
 *    their batch only.
 *
 * Regularisation ids follow the reference's kernel names
 * (cl_kernel_name_ext, reference src/VortFunc.cpp:210,223,236,249).
 */
#ifndef CVTX_HIP_H
#define CVTX_HIP_H

#include <stddef.h>

#include <cvortex/libcvtx.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CVTX_HIP_REG_SINGULAR 0
#define CVTX_HIP_REG_WINCKELMANS 1
#define CVTX_HIP_REG_PLANETARY 2
#define CVTX_HIP_REG_GAUSSIAN 3

#define CVTX_HIP_OUT_VEL 1
#define CVTX_HIP_OUT_DVORT 2
#define CVTX_HIP_OUT_VISC 4

#define CVTX_HIP_ACC_PLAIN 0 /* one FP32 running sum per output component */
/* FP32 tile sums + TwoSum across tiles, FP64 across partitions; P3D also sums
 * the stretching (and PSE) terms of near pairs in FP64 per pair: coherent
 * fields (a vortex ring) need it, their near terms being large and nearly
 * parallel to the target's strength; P2D sums its viscous exchange
 * (A_t a_s - a_t A_s) eta in FP64 per pair, the two products nearly
 * cancelling for many targets */
#define CVTX_HIP_ACC_COMP 1
/* FP32 tile sums + TwoSum across tiles for every sum, no per-pair FP64:
 * enough for fields without that coherence (independent random strengths);
 * the same as CVTX_HIP_ACC_COMP for P2D velocity, F3D and the vorticity
 * field */
#define CVTX_HIP_ACC_TILES 2

/* ---- 1. seam: replaces opencl_brute_force_* (reference src/ocl_P3D.h:33-68,
 *         src/ocl_F3D.h, src/ocl_P2D.h) ---- */

/* replaces opencl_brute_force_P3D_M2M_vel, reference src/ocl_P3D.cpp:37-62 */
int cvtx_hip_P3D_M2M_vel(const cvtx_P3D** array_start, int num_particles,
                         const bsv_V3f* mes_start, int num_mes, bsv_V3f* result_array,
                         const cvtx_VortFunc* kernel, float regularisation_radius);
/* replaces opencl_brute_force_P3D_M2M_dvort, reference src/ocl_P3D.cpp:64-91 */
int cvtx_hip_P3D_M2M_dvort(const cvtx_P3D** array_start, int num_particles,
                           const cvtx_P3D** induced_start, int num_induced,
                           bsv_V3f* result_array, const cvtx_VortFunc* kernel,
                           float regularisation_radius);
/* replaces opencl_brute_force_P3D_M2M_visc_dvort, reference src/ocl_P3D.cpp:92-119.
 * Singular and planetary have no PSE kernel: like the reference
 * (src/VortFunc.cpp:52-62) the result is zero, with a one-time warning. */
int cvtx_hip_P3D_M2M_visc_dvort(const cvtx_P3D** array_start, int num_particles,
                                const cvtx_P3D** induced_start, int num_induced,
                                bsv_V3f* result_array, const cvtx_VortFunc* kernel,
                                float regularisation_radius, float kinematic_visc);
/* replaces opencl_brute_force_F3D_M2M_vel, reference src/ocl_F3D.cpp:36-70 */
int cvtx_hip_F3D_M2M_vel(const cvtx_F3D** array_start, int num_filaments,
                         const bsv_V3f* mes_start, int num_mes, bsv_V3f* result_array);
/* replaces opencl_brute_force_F3D_M2M_dvort, reference src/ocl_F3D.cpp:422-445 */
int cvtx_hip_F3D_M2M_dvort(const cvtx_F3D** array_start, int num_filaments,
                           const cvtx_P3D** induced_start, int num_induced,
                           bsv_V3f* result_array);
/* replaces opencl_brute_force_P2D_M2M_vel, reference src/ocl_P2D.cpp:37-76 */
int cvtx_hip_P2D_M2M_vel(const cvtx_P2D** array_start, int num_particles,
                         const bsv_V2f* mes_start, int num_mes, bsv_V2f* result_array,
                         const cvtx_VortFunc* kernel, float regularisation_radius);
/* replaces opencl_brute_force_P2D_M2M_visc_dvort, reference src/ocl_P2D.cpp:78-105 */
int cvtx_hip_P2D_M2M_visc_dvort(const cvtx_P2D** array_start, int num_particles,
                                const cvtx_P2D** induced_start, int num_induced,
                                float* result_array, const cvtx_VortFunc* kernel,
                                float regularisation_radius, float kinematic_visc);

/* replaces opencl_brute_force_P3D_M2M_vort, reference src/ocl_P3D.cpp:121-147,728-899,
 * with the reference CPU semantics (src/P3D.cpp:298-322): 5 sigma box
 * cut-off, self-pair included.  (The reference's OpenCL kernel,
 * src/ocl_kernels/nbody.cl:164-197, has no cut-off and drops the self-pair.) */
int cvtx_hip_P3D_M2M_vort(const cvtx_P3D** array_start, int num_particles,
                          const bsv_V3f* mes_start, int num_mes, bsv_V3f* result_array,
                          const cvtx_VortFunc* kernel, float regularisation_radius);
/* the whole of cvtx_P3D_pedrizzetti_relaxation (reference src/P3D.cpp:669-707)
 * on the device: vorticity at the particles, then the strength update */
int cvtx_hip_P3D_pedrizzetti_relaxation(cvtx_P3D** input_array_start, int n_input_particles,
                                        float fdt, const cvtx_VortFunc* kernel,
                                        float regularisation_radius);

/* cvtx_F3D_inf_mtrx on the device (the reference computes it on the CPU
 * only, src/F3D.cpp:204-227) */
int cvtx_hip_F3D_inf_mtrx(const cvtx_F3D** array_start, int num_filaments,
                          const bsv_V3f* mes_start, const bsv_V3f* dir_start, int num_mes,
                          float* result_matrix);

/* ---- 2. fused host-array entry points (no reference counterpart) ----
 * Any of the result pointers may be NULL (that output is not computed). */
int cvtx_hip_P3D_M2M_fused(const cvtx_P3D** array_start, int num_particles,
                           const cvtx_P3D** induced_start, int num_induced,
                           bsv_V3f* vel, bsv_V3f* dvort, bsv_V3f* visc_dvort,
                           const cvtx_VortFunc* kernel, float regularisation_radius,
                           float kinematic_visc);
int cvtx_hip_F3D_M2M_fused(const cvtx_F3D** array_start, int num_filaments,
                           const cvtx_P3D** induced_start, int num_induced,
                           bsv_V3f* vel, bsv_V3f* dvort);
int cvtx_hip_P2D_M2M_fused(const cvtx_P2D** array_start, int num_particles,
                           const cvtx_P2D** induced_start, int num_induced,
                           bsv_V2f* vel, float* visc_dvort, const cvtx_VortFunc* kernel,
                           float regularisation_radius, float kinematic_visc);

/* ---- controls (no reference counterpart) ---- */
const char* cvtx_hip_last_error(void);

/* Which leg evaluates the public void cvtx_* functions.
 * AUTO (default): the device when one is enabled, the regularisation's
 *   cl_kernel_name_ext is one of the four kernel names, and the problem
 *   passes the reference's size guards (both counts >= 256 for P3D, F3D
 *   dvort and P2D visc; the reference's cost models for F3D and P2D vel,
 *   src/ocl_F3D.cpp:49-51, src/ocl_P2D.cpp:54-57); the single-source and
 *   single-target forms go to the device from CVTX_HIP_SMALL_DEVICE_MIN
 *   elements on.  Otherwise, or when the device call fails, the library's
 *   CPU leg: the caller's cvtx_VortFunc function pointers per pair, FP64
 *   sums, host threads (CVTX_NUM_THREADS or OMP_NUM_THREADS).
 * DEVICE: every call the device can take goes to it (the CPU leg remains
 *   for no device, an unknown kernel name, or a failure).
 * HOST: every call runs on the CPU leg.
 * Returns the previous mode, or -1 for an invalid mode.  The int seam
 * (cvtx_hip_*) never uses the CPU leg. */
#define CVTX_HIP_DISPATCH_AUTO 0
#define CVTX_HIP_DISPATCH_DEVICE 1
#define CVTX_HIP_DISPATCH_HOST 2
#define CVTX_HIP_SMALL_DEVICE_MIN 4096
int cvtx_hip_set_dispatch(int mode);
int cvtx_hip_get_dispatch(void);
/* Leg that evaluated the last public void call: CVTX_HIP_LEG_DEVICE,
 * CVTX_HIP_LEG_HOST, or -1 before any call (or after a failed one). */
#define CVTX_HIP_LEG_DEVICE 0
#define CVTX_HIP_LEG_HOST 1
int cvtx_hip_last_leg(void);
/* Host threads the CPU leg uses. */
int cvtx_hip_host_threads(void);
/* Accumulation used by every call that does not name one; returns the
 * previous mode, or -1 for an invalid mode. Default CVTX_HIP_ACC_COMP. */
int cvtx_hip_set_accumulation(int mode);
int cvtx_hip_get_accumulation(void);
/* Spatial (Morton) ordering of targets (on by default) before Gaussian P3D
 * launches, where waves of nearby targets skip the near-field table for
 * distant sources, and before filament (F3D) launches, where compact target
 * blocks let most filament tiles take the far-tile loop.  Results agree to
 * rounding either way.  Returns the previous setting. */
int cvtx_hip_set_target_order(int enable);
/* The library's stream on the active device (hipStream_t as void*), NULL
 * when no device is enabled. */
void* cvtx_hip_stream(void);
/* Block until all work queued by the library on the active device is done. */
int cvtx_hip_synchronize(void);

/* ---- 3. device-resident API (no reference counterpart) ---- */

/* P3D particles: pv = float4 {x, y, z, volume}, a = float4 {ax, ay, az, -}.
 * Targets: tgt_pv (float4) always; tgt_a needed for dvort and visc.
 * Outputs: float[3*n_tgt] each (NULL = not computed). */
typedef struct {
	const float* src_pv;
	const float* src_a;
	int n_src;
	const float* tgt_pv;
	const float* tgt_a;
	int n_tgt;
	float* vel;
	float* dvort;
	float* visc;
	int regularisation; /* CVTX_HIP_REG_* */
	float sigma;        /* regularisation radius */
	float nu;           /* kinematic viscosity (visc only) */
	int accumulate;     /* CVTX_HIP_ACC_* */
} cvtx_hip_p3d_problem;

/* P2D particles and targets: float4 {x, y, vorticity, area}.
 * Outputs: vel float[2*n_tgt], visc float[n_tgt] (NULL = not computed). */
typedef struct {
	const float* src;
	int n_src;
	const float* tgt;
	int n_tgt;
	float* vel;
	float* visc;
	int regularisation;
	float sigma;
	float nu;
	int accumulate;
} cvtx_hip_p2d_problem;

/* F3D filaments: fa = float4 {start x, y, z, strength}, fb = float4 {end x, y, z, -}.
 * Targets as for P3D; tgt_a needed for dvort. */
typedef struct {
	const float* fa;
	const float* fb;
	int n_fil;
	const float* tgt_pv;
	const float* tgt_a;
	int n_tgt;
	float* vel;
	float* dvort;
	int accumulate;
} cvtx_hip_f3d_problem;

/* Vorticity field (cvtx_P3D_M2M_vort semantics): sources src_pv/src_a as
 * above, targets tgt (float4 {x, y, z, -}), vort float[3*n_tgt]. */
typedef struct {
	const float* src_pv;
	const float* src_a;
	int n_src;
	const float* tgt;
	int n_tgt;
	float* vort;
	int regularisation; /* CVTX_HIP_REG_* */
	float sigma;
	int accumulate;     /* CVTX_HIP_ACC_* */
} cvtx_hip_vort_problem;

/* Evaluate with the tuned (or default) launch shape.  A source-partitioned
 * launch needs workspace; the library keeps a per-device workspace that it
 * grows on demand (this may allocate device memory on first use of a size). */
int cvtx_hip_p3d_eval(const cvtx_hip_p3d_problem* p, void* stream);
int cvtx_hip_p2d_eval(const cvtx_hip_p2d_problem* p, void* stream);
int cvtx_hip_f3d_eval(const cvtx_hip_f3d_problem* p, void* stream);

int cvtx_hip_vort_eval(const cvtx_hip_vort_problem* p, void* stream);
int cvtx_hip_vort_eval_shape(const cvtx_hip_vort_problem* p, int targets_per_lane, int n_split,
                             void* stream);
/* Influence matrix of device-resident filaments (fa = float4 {start,
 * strength}, fb = float4 {end, -}) at points mes with directions dir (float4
 * {x, y, z, -} each): out[i * n_fil + j], float[n_mes * n_fil]. */
int cvtx_hip_f3d_inf_eval(const float* fa, const float* fb, int n_fil, const float* mes,
                          const float* dir, int n_mes, float* out, void* stream);
/* Pedrizzetti relaxation of n device-resident particles (pv, a as above):
 * a is updated in place (its 4th component is kept). */
int cvtx_hip_relax_eval(const float* pv, float* a, int n, float fdt, int regularisation,
                        float sigma, int accumulate, void* stream);

/* Same, with an explicit launch shape: targets_per_lane in {1, 2, 4},
 * n_split >= 1 source partitions (rounded to whole 256-source tiles). */
int cvtx_hip_p3d_eval_shape(const cvtx_hip_p3d_problem* p, int targets_per_lane, int n_split,
                            void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the same
 * evaluation as cvtx_hip_p3d_eval_shape, with the FP64 partial sums of a
 * source-partitioned launch in a fixed device array of
 * CVTX_HIP_NOALLOC_SCRATCH_BYTES per GPU (held from library load to process
 * exit) instead of the library's workspace.  Calls no hipMalloc, hipFree or
 * allocating library routine; targets are not reordered (same results).
 * Returns -1 without launching when the shape needs more than the array:
 * n_split x accumulators x n_tgt x 8 bytes, accumulators 3 (vel), 6 (dvort),
 * 4 (visc), added over the requested outputs.  All calls on one GPU share
 * its array: calls outside graph capture are ordered across streams, while
 * concurrent replays of captured calls race. */
#define CVTX_HIP_NOALLOC_SCRATCH_BYTES (64u << 20)
int cvtx_hip_p3d_eval_shape_noalloc(const cvtx_hip_p3d_problem* p, int targets_per_lane,
                                    int n_split, void* stream);
int cvtx_hip_p2d_eval_shape(const cvtx_hip_p2d_problem* p, int targets_per_lane, int n_split,
                            void* stream);
int cvtx_hip_f3d_eval_shape(const cvtx_hip_f3d_problem* p, int targets_per_lane, int n_split,
                            void* stream);

/* ---- batched clouds and RK4 graphs (no reference counterpart) ----
 * A batch holds many independent P3D clouds stored back to back in one pair
 * of device SoA buffers (pv, a as above): cloud c owns particles
 * [offsets[c], offsets[c+1]) and has regularisation radius sigmas[c]; each
 * cloud interacts only with itself.  `outputs` is
 * CVTX_HIP_OUT_VEL|CVTX_HIP_OUT_DVORT (optionally |CVTX_HIP_OUT_VISC);
 * regularisation Winckelmans or Gaussian.  offsets/sigmas are host arrays
 * (copied).  NULL on failure (cvtx_hip_last_error). */
typedef struct cvtx_hip_batch cvtx_hip_batch;
cvtx_hip_batch* cvtx_hip_batch_create(int n_clouds, const int* offsets, const float* sigmas,
                                      int regularisation, int outputs, int accumulate, float nu);
void cvtx_hip_batch_destroy(cvtx_hip_batch* batch);
/* Targets per lane (1, 2, 4) and resident workgroups per CU of the
 * persistent launch (the workgroup size is kept); must precede
 * cvtx_hip_batch_rk4_init. */
int cvtx_hip_batch_set_shape(cvtx_hip_batch* batch, int targets_per_lane, int workgroups_per_cu);
int cvtx_hip_batch_info(const cvtx_hip_batch* batch, int* n_units, int* grid, int* targets_per_lane);
/* The batch's launch shape: targets per lane, workgroup size (128 or 256,
 * also the source tile length) and resident workgroups per CU. */
int cvtx_hip_batch_shape(const cvtx_hip_batch* batch, int* targets_per_lane, int* workgroup_size,
                         int* workgroups_per_cu);
/* Autotuner of the persistent launch: times every candidate shape
 * (targets per lane {1, 2, 4} x workgroup size {128, 256} x resident
 * workgroups per CU {2, 4, 8}; one warm-up and `reps` timed evaluations
 * each, on the library's stream; the outputs are overwritten), keeps the
 * fastest for this batch and records it in the shape cache under the
 * batch's signature (regularisation, outputs, accumulation, log2 of the
 * pair count, log2 of the cloud count): batches created later with the same
 * signature start in that shape, and cvtx_hip_tune_save/load persist it.
 * Must precede cvtx_hip_batch_rk4_init.  Runs on the library's stream and
 * returns when the timing is done: work of the caller's own streams that
 * writes pv or a must be complete before the call.  Returns 0 and the
 * winner, or -1. */
int cvtx_hip_batch_autotune(cvtx_hip_batch* batch, const float* pv, const float* a, float* vel,
                            float* dvort, float* visc, int reps, int* targets_per_lane,
                            int* workgroup_size, int* workgroups_per_cu, float* best_ms);
/* One persistent launch over all clouds: outputs float[3 * n_particles]. */
int cvtx_hip_batch_eval(cvtx_hip_batch* batch, const float* pv, const float* a, float* vel,
                        float* dvort, float* visc, void* stream);
/* The same results with one ordinary launch per cloud (baseline).  plain = 0:
 * each cloud gets its default launch shape (targets per lane, source
 * partitions) and the launches are queued back to back; plain = 1: one
 * target per lane, no source partitions, and the stream is synchronised
 * after each cloud, as a blocking per-cloud API call would be. */
int cvtx_hip_batch_eval_per_cloud(cvtx_hip_batch* batch, const float* pv, const float* a,
                                  float* vel, float* dvort, float* visc, int plain, void* stream);
/* Classical RK4 on (positions, vorticity) with (vel, dvort): init allocates
 * the stage buffers and captures one whole step (4 evaluations, 4 update
 * kernels) as a HIP graph; each step then updates pv and a in place.
 * Requires a vel|dvort batch. */
int cvtx_hip_batch_rk4_init(cvtx_hip_batch* batch, float* pv, float* a, float dt);
int cvtx_hip_batch_rk4_step(cvtx_hip_batch* batch, int n_steps, void* stream);

/* ---- autotuner (no reference counterpart) ----
 * Times the candidate launch shapes of a device-resident problem on the
 * active device (the problem's outputs are overwritten) and records the
 * fastest for the problem's size bucket; later calls of any P3D entry point
 * in that bucket use it.  Returns 0 and the winner, or -1. */
int cvtx_hip_autotune_p3d(const cvtx_hip_p3d_problem* p, int reps, int* best_T, int* best_split,
                          float* best_ms);
/* Save / load the shape cache as text; load returns the number of entries. */
int cvtx_hip_tune_save(const char* path);
int cvtx_hip_tune_load(const char* path);
void cvtx_hip_tune_clear(void);

/* ---- inspection (no reference counterpart) ----
 * Copies the Gaussian near-field table used by the kernels (96 pieces x
 * 8 floats: cubic coefficients of G = g/rho^3 and H/3 = (G - zeta/3)/rho^2
 * on rho in [i/16, (i+1)/16)) into `out`; returns the number of floats. */
int cvtx_hip_gaussian_table(float* out, int capacity);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CVTX_HIP_H */
