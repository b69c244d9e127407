/*
 * libcvtx.h -- public C ABI of the MI355X (gfx950) cvortex direct-sum library.
 *
 * This header declares the subset of the reference cvortex C API that this
 * library implements: the element types, the regularisation plugin struct,
 * the accelerator controls, the seven many-to-many (M2M) direct-sum
 * Biot-Savart entry points, and the vorticity field and relaxation built on
 * the same pairwise machinery.  Types, struct layouts and function signatures
 * are ABI-identical to the reference header include/cvortex/libcvtx.h, so a
 * program (or the CVortex.jl wrapper) built against the reference header can
 * link against libcvortex.so instead.
 *
 * Semantics follow the reference:
 *   - calls are synchronous and return void;
 *   - inputs are arrays of pointers to caller-owned records (AoP);
 *   - result arrays are caller-allocated and are overwritten;
 *   - source and target pointer arrays may alias.
 * Differences (the library has no CPU compute path):
 *   - every M2M call runs on the first enabled HIP device, for any size;
 *   - if no device is enabled, or the regularisation's cl_kernel_name_ext is
 *     not one of "singular", "winckelmans", "planetary", "gaussian", the call
 *     prints one line to stderr and fills the result array with NaN.  The
 *     int-returning seam in cvortex_hip.h reports the same condition as -1
 *     so that a host library can run its own CPU path instead.
 *
 * The library is not thread-safe (nor is the reference: its device state is
 * global, reference src/opencl_acc.cpp:42-49).
 */
#ifndef CVTX_LIBCVTX_H
#define CVTX_LIBCVTX_H

#ifndef CVTX_EXPORT
# ifdef _WIN32
#  define CVTX_EXPORT __declspec(dllimport)
# else
#  define CVTX_EXPORT
# endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

#include <bsv/bsv.h>

/* 3D vortex particle: 28 bytes.  `vorticity` is the particle strength
 * (vorticity x volume).  Reference libcvtx.h:53-57. */
typedef struct {
	bsv_V3f coord;
	bsv_V3f vorticity;
	float volume;
} cvtx_P3D;

/* Straight singular vortex filament: 28 bytes.  Reference libcvtx.h:60-63. */
typedef struct {
	bsv_V3f start, end;
	float strength;
} cvtx_F3D;

/* 2D vortex particle: 16 bytes.  Reference libcvtx.h:66-70. */
typedef struct {
	bsv_V2f coord;
	float vorticity;
	float area;
} cvtx_P2D;

/* Regularisation plugin: 6 function pointers + char[32] = 80 bytes (LP64).
 * The GPU selects its kernel by cl_kernel_name_ext; the function pointers
 * are scalar host functions for callers that evaluate g, zeta or eta
 * directly.  Reference libcvtx.h:86-94. */
typedef struct {
	float(*g_3D)(float rho);
	float(*g_2D)(float rho);
	float(*zeta_3D)(float rho);
	void(*combined_3D)(float rho, float* g, float* zeta);
	float(*eta_3D)(float rho);
	float(*eta_2D)(float rho);
	char cl_kernel_name_ext[32];
} cvtx_VortFunc;

/* Library and accelerator controls.  Reference libcvtx.h:102-110,
 * src/accelerators.cpp:39-118.  An accelerator is a HIP device. */
CVTX_EXPORT void cvtx_initialise();
CVTX_EXPORT void cvtx_finalise();
CVTX_EXPORT const char* cvtx_information();
CVTX_EXPORT int cvtx_num_accelerators();
CVTX_EXPORT int cvtx_num_enabled_accelerators();
CVTX_EXPORT const char* cvtx_accelerator_name(int accelerator_id);
CVTX_EXPORT int cvtx_accelerator_enabled(int accelerator_id);
CVTX_EXPORT void cvtx_accelerator_enable(int accelerator_id);
CVTX_EXPORT void cvtx_accelerator_disable(int accelerator_id);

/* Regularisation factories.  Reference libcvtx.h:113-116,
 * src/VortFunc.cpp:201-251. */
CVTX_EXPORT const cvtx_VortFunc cvtx_VortFunc_singular(void);
CVTX_EXPORT const cvtx_VortFunc cvtx_VortFunc_winckelmans(void);
CVTX_EXPORT const cvtx_VortFunc cvtx_VortFunc_planetary(void);
CVTX_EXPORT const cvtx_VortFunc cvtx_VortFunc_gaussian(void);

/* 3D vortex particles: velocity induced at points.  Reference
 * libcvtx.h:213-220, src/P3D.cpp:343-366. */
CVTX_EXPORT void cvtx_P3D_M2M_vel(
	const cvtx_P3D **array_start,
	const int num_particles,
	const bsv_V3f *mes_start,
	const int num_mes,
	bsv_V3f *result_array,
	const cvtx_VortFunc *kernel,
	float regularisation_radius);

/* 3D vortex particles: inviscid vortex stretching (transpose scheme) on
 * particles.  Reference libcvtx.h:222-229, src/P3D.cpp:387-410. */
CVTX_EXPORT void cvtx_P3D_M2M_dvort(
	const cvtx_P3D **array_start,
	const int num_particles,
	const cvtx_P3D **induced_start,
	const int num_induced,
	bsv_V3f *result_array,
	const cvtx_VortFunc *kernel,
	float regularisation_radius);

/* 3D vortex particles: viscous (PSE) vorticity exchange on particles.
 * Reference libcvtx.h:231-239, src/P3D.cpp:432-456. */
CVTX_EXPORT void cvtx_P3D_M2M_visc_dvort(
	const cvtx_P3D **array_start,
	const int num_particles,
	const cvtx_P3D **induced_start,
	const int num_induced,
	bsv_V3f *result_array,
	const cvtx_VortFunc *kernel,
	float regularisation_radius,
	float kinematic_visc);

/* 3D vortex particles: vorticity field at points.  Reference
 * libcvtx.h:241-248, src/P3D.cpp:476-498 with the CPU semantics of
 * src/P3D.cpp:298-322: only sources inside the box |dx|, |dy|, |dz| <
 * 5 sigma of a point contribute, the self-pair included. */
CVTX_EXPORT void cvtx_P3D_M2M_vort(
	const cvtx_P3D **array_start,
	const int num_particles,
	const bsv_V3f *mes_start,
	const int num_mes,
	bsv_V3f *result_array,
	const cvtx_VortFunc *kernel,
	float regularisation_radius);

/* 3D vortex particles: Pedrizzetti relaxation of the particle strengths
 * towards the local vorticity field, in place.  Reference libcvtx.h:259-264,
 * src/P3D.cpp:669-707. */
CVTX_EXPORT void cvtx_P3D_pedrizzetti_relaxation(
	cvtx_P3D **input_array_start,
	const int n_input_particles,
	float fdt,
	const cvtx_VortFunc *kernel,
	float regularisation_radius);

/* Straight vortex filaments: velocity induced at points.  Reference
 * libcvtx.h:285-290, src/F3D.cpp:162-180. */
CVTX_EXPORT void cvtx_F3D_M2M_vel(
	const cvtx_F3D **array_start,
	const int num_filaments,
	const bsv_V3f *mes_start,
	const int num_mes,
	bsv_V3f *result_array);

/* Straight vortex filaments: vortex stretching on particles.  Reference
 * libcvtx.h:292-297, src/F3D.cpp:182-202. */
CVTX_EXPORT void cvtx_F3D_M2M_dvort(
	const cvtx_F3D **array_start,
	const int num_filaments,
	const cvtx_P3D **induced_start,
	const int num_induced,
	bsv_V3f *result_array);

/* Straight vortex filaments: influence matrix, row-major num_mes x
 * num_filaments: result_matrix[i * num_filaments + j] is the velocity of
 * filament j (with its own strength, as the reference code computes it)
 * at mes_start[i], projected on dir_start[i].  Reference libcvtx.h:299-305,
 * src/F3D.cpp:204-227. */
CVTX_EXPORT void cvtx_F3D_inf_mtrx(
	const cvtx_F3D **array_start,
	const int num_filaments,
	const bsv_V3f *mes_start,
	const bsv_V3f *dir_start,
	const int num_mes,
	float *result_matrix);

/* 2D vortex particles: velocity induced at points.  Reference
 * libcvtx.h:329-336, src/P2D.cpp:141-162. */
CVTX_EXPORT void cvtx_P2D_M2M_vel(
	const cvtx_P2D **array_start,
	const int num_particles,
	const bsv_V2f *mes_start,
	const int num_mes,
	bsv_V2f *result_array,
	const cvtx_VortFunc *kernel,
	float regularisation_radius);

/* 2D vortex particles: viscous (PSE) vorticity exchange on particles.
 * Reference libcvtx.h:362-370, src/P2D.cpp:252-276. */
CVTX_EXPORT void cvtx_P2D_M2M_visc_dvort(
	const cvtx_P2D **array_start,
	const int num_particles,
	const cvtx_P2D **induced_start,
	const int num_induced,
	float *result_array,
	const cvtx_VortFunc *kernel,
	float regularisation_radius,
	float kinematic_visc);

/* ---- redistribution on a grid (remeshing) ----
 * Interpolation kernels, reference libcvtx.h:96-99, 118-123,
 * src/RedistFunc.cpp:36-96.  The device recognises the library's own five
 * functions; a cvtx_RedistFunc around another function is rejected. */
typedef struct {
	float(*func)(float U);
	float radius;
} cvtx_RedistFunc;

CVTX_EXPORT const cvtx_RedistFunc cvtx_RedistFunc_lambda0(void);
CVTX_EXPORT const cvtx_RedistFunc cvtx_RedistFunc_lambda1(void);
CVTX_EXPORT const cvtx_RedistFunc cvtx_RedistFunc_lambda2(void);
CVTX_EXPORT const cvtx_RedistFunc cvtx_RedistFunc_lambda3(void);
CVTX_EXPORT const cvtx_RedistFunc cvtx_RedistFunc_m4p(void);

/* Redistribute particles onto a grid of spacing grid_density; particles
 * weaker than negligible_vort x the mean strength are removed (their
 * strength shared among the rest).  With output_particles NULL returns the
 * number of particles created; otherwise writes at most
 * max_output_particles of them (in grid Morton order) and returns the
 * number written.  -1 on failure (message on stderr).  Reference
 * libcvtx.h:250-257 and 372-380, src/P3D.cpp:509-665, src/P2D.cpp:283-436. */
CVTX_EXPORT int cvtx_P3D_redistribute_on_grid(
	const cvtx_P3D **input_array_start,
	const int n_input_particles,
	cvtx_P3D *output_particles,
	int max_output_particles,
	const cvtx_RedistFunc *redistributor,
	float grid_density,
	float negligible_vort);
CVTX_EXPORT int cvtx_P2D_redistribute_on_grid(
	const cvtx_P2D **input_array_start,
	const int n_input_particles,
	cvtx_P2D *output_particles,
	int max_output_particles,
	const cvtx_RedistFunc *redistributor,
	float grid_density,
	float negligible_vort);

/* ---- single-source (S2S, S2M) and single-target (M2S) forms ----
 * Reference libcvtx.h:125-211 (P3D), 266-283 (F3D), 306-360 (P2D); bodies
 * src/P3D.cpp:74-296, src/F3D.cpp:34-128, src/P2D.cpp:71-231.  Evaluated on
 * the device through the M2M paths (one device round trip per call).  As
 * in the reference, the single-source vorticity forms (S2S_vort, S2M_vort,
 * src/P3D.cpp:146-161, 213-228) have no cut-off and use rho = |r| / sigma,
 * while M2S_vort cuts at the 5 sigma box like M2M_vort. */
CVTX_EXPORT bsv_V3f cvtx_P3D_S2S_vel(const cvtx_P3D *self, const bsv_V3f mes_point,
	const cvtx_VortFunc *kernel, float regularisation_radius);
CVTX_EXPORT bsv_V3f cvtx_P3D_S2S_dvort(const cvtx_P3D *self, const cvtx_P3D *induced_particle,
	const cvtx_VortFunc *kernel, float regularisation_radius);
CVTX_EXPORT bsv_V3f cvtx_P3D_S2S_visc_dvort(const cvtx_P3D *self,
	const cvtx_P3D *induced_particle, const cvtx_VortFunc *kernel,
	float regularisation_radius, float kinematic_visc);
CVTX_EXPORT bsv_V3f cvtx_P3D_S2S_vort(const cvtx_P3D *self, const bsv_V3f mes_point,
	const cvtx_VortFunc *kernel, float regularisation_radius);
CVTX_EXPORT void cvtx_P3D_S2M_vel(const cvtx_P3D *self, const bsv_V3f *mes_start,
	const int num_mes, bsv_V3f *result_array, const cvtx_VortFunc *kernel,
	float regularisation_radius);
CVTX_EXPORT void cvtx_P3D_S2M_dvort(const cvtx_P3D *self, const cvtx_P3D **induced_start,
	const int num_induced, bsv_V3f *result_array, const cvtx_VortFunc *kernel,
	float regularisation_radius);
CVTX_EXPORT void cvtx_P3D_S2M_visc_dvort(const cvtx_P3D *self, const cvtx_P3D **induced_start,
	const int num_induced, bsv_V3f *result_array, const cvtx_VortFunc *kernel,
	float regularisation_radius, float kinematic_visc);
CVTX_EXPORT void cvtx_P3D_S2M_vort(const cvtx_P3D *self, const bsv_V3f *mes_start,
	const int num_mes, bsv_V3f *result_array, const cvtx_VortFunc *kernel,
	float regularisation_radius);
CVTX_EXPORT bsv_V3f cvtx_P3D_M2S_vel(const cvtx_P3D **array_start, const int num_particles,
	const bsv_V3f mes_point, const cvtx_VortFunc *kernel, float regularisation_radius);
CVTX_EXPORT bsv_V3f cvtx_P3D_M2S_dvort(const cvtx_P3D **array_start, const int num_particles,
	const cvtx_P3D *induced_particle, const cvtx_VortFunc *kernel,
	float regularisation_radius);
CVTX_EXPORT bsv_V3f cvtx_P3D_M2S_visc_dvort(const cvtx_P3D **array_start,
	const int num_particles, const cvtx_P3D *induced_particle, const cvtx_VortFunc *kernel,
	float regularisation_radius, float kinematic_visc);
CVTX_EXPORT bsv_V3f cvtx_P3D_M2S_vort(const cvtx_P3D **array_start, const int num_particles,
	const bsv_V3f mes_point, const cvtx_VortFunc *kernel, float regularisation_radius);

CVTX_EXPORT bsv_V3f cvtx_F3D_S2S_vel(const cvtx_F3D *self, const bsv_V3f mes_point);
CVTX_EXPORT bsv_V3f cvtx_F3D_S2S_dvort(const cvtx_F3D *self, const cvtx_P3D *induced_particle);
CVTX_EXPORT bsv_V3f cvtx_F3D_M2S_vel(const cvtx_F3D **array_start, const int num_filaments,
	const bsv_V3f mes_point);
CVTX_EXPORT bsv_V3f cvtx_F3D_M2S_dvort(const cvtx_F3D **array_start, const int num_filaments,
	const cvtx_P3D *induced_particle);

CVTX_EXPORT bsv_V2f cvtx_P2D_S2S_vel(const cvtx_P2D *self, const bsv_V2f mes_point,
	const cvtx_VortFunc *kernel, float regularisation_radius);
CVTX_EXPORT void cvtx_P2D_S2M_vel(const cvtx_P2D *self, const bsv_V2f *mes_start,
	const int num_mes, bsv_V2f *result_array, const cvtx_VortFunc *kernel,
	float regularisation_radius);
CVTX_EXPORT bsv_V2f cvtx_P2D_M2S_vel(const cvtx_P2D **array_start, const int num_particles,
	const bsv_V2f mes_point, const cvtx_VortFunc *kernel, float regularisation_radius);
CVTX_EXPORT float cvtx_P2D_S2S_visc_dvort(const cvtx_P2D *self,
	const cvtx_P2D *induced_particle, const cvtx_VortFunc *kernel,
	float regularisation_radius, float kinematic_visc);
CVTX_EXPORT void cvtx_P2D_S2M_visc_dvort(const cvtx_P2D *self, const cvtx_P2D **induced_start,
	const int num_induced, float *result_array, const cvtx_VortFunc *kernel,
	float regularisation_radius, float kinematic_visc);
CVTX_EXPORT float cvtx_P2D_M2S_visc_dvort(const cvtx_P2D **array_start,
	const int num_particles, const cvtx_P2D *induced_particle, const cvtx_VortFunc *kernel,
	float regularisation_radius, float kinematic_visc);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CVTX_LIBCVTX_H */
