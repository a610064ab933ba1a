/*
 * wrfdc -- C ABI of the MI355X (gfx950) RK3 split-explicit dycore.
 *
 * The library advances the state of one WRF-JAX dycore domain that the
 * caller keeps resident in device memory.  Plain pointers and sizes only; no
 * C++ exceptions cross this boundary.  Every entry point returns 0 on success
 * and a negative WRFDC_E* code on failure (details in wrfdc_last_error()).
 *
 * Reference interfaces replaced (paths relative to the WRF-JAX source tree):
 *   wrfdc_create     build_step_em_core_context  src/wrf_jax/core/step_em.py:40-65
 *                    (+ build_direct_step_plan   src/wrf_jax/core/direct_step.py:2412-2460)
 *   wrfdc_bind       model_state_to_fields       src/wrf_jax/core/conversion.py:62-67
 *                    (caller-owned device arrays instead of a ModelState pytree)
 *   wrfdc_step       step_em_core_once           src/wrf_jax/core/step_em.py:74-93
 *                    = run_direct_step           src/wrf_jax/core/direct_step.py:3425-3531
 *   wrfdc_stage      one StageFn callback of strict_solve_callbacks()
 *                                                src/wrf_jax/runtime.py:16235-16301
 *                    by the stage names solve_em calls
 *                                                src/wrf_jax/dyn_em/solve_em.py:470-790
 *   wrfdc_set_dtbc   _write(grid, "dtbc", dtbc + dt) at the top of solve_em
 *                                                src/wrf_jax/dyn_em/solve_em.py:455
 *   wrfdc_checksums  assert_state_finite         src/wrf_jax/data_io.py:222-228
 *                    (plus FP64 conservation checksums: column mass, theta,
 *                    moisture, dry-air mass and total energy)
 */
#ifndef WRFDC_H
#define WRFDC_H

#ifdef __cplusplus
extern "C" {
#endif

#define WRFDC_OK 0
#define WRFDC_EINVAL -1    /* bad argument / unsupported option (ValueError) */
#define WRFDC_ENOTIMPL -2  /* option not implemented (NotImplementedError) */
#define WRFDC_EUNBOUND -3  /* a required field has no device buffer bound */
#define WRFDC_EHIP -4      /* HIP runtime error */
#define WRFDC_ENONFINITE -5 /* NaN/Inf found by wrfdc_checksums (FloatingPointError) */

/* wrfdc_checksums output, FP64 sums over the mass points:
 * [0] sum(mu_2 + mub), [1] sum(t_2), [2] sum of all moist species,
 * [3] number of non-finite prognostic values, [4] dry-air mass (kg),
 * [5] total energy: sum of dm * (kinetic + cv*T + geopotential) (J). */
#define WRFDC_NCHECKSUMS 6

/* Scalar configuration; field order and types match the Python mirror
 * wrfdc.config.CConfig.  Dimensions are mass-point counts. */
typedef struct wrfdc_config {
  int nx, ny, nz;       /* mass points */
  int ip;               /* i pitch (elements) of 3-D/2-D arrays, >= nx+1 */
  int nmoist, nscalar;  /* moist species (index 0 = QV) and extra scalars */
  int spec_bdy_width, spec_zone, relax_zone;
  int time_step_sound;  /* acoustic steps per dt, 0 = WRF formula */
  int pd_moist, pd_scalar; /* positive-definite advection on the last stage */
  int top_lid;
  int pad_;
  double dx, dy, dt;
  double ptop, p0, t0, g, r_d, r_v, cp, cpovcv, cvpm, rvovrd;
  double epssm, smdiv, emdiv, c_s, prandtl;
  double cf1, cf2, cf3, cfn, cfn1;
  double dtbc;          /* seconds since the last lateral-boundary update */
  double reradius;      /* 1 / earth radius (curvature terms) */
} wrfdc_config;

typedef struct wrfdc_ctx wrfdc_ctx;

/* Field registry: names and kinds of every array the step reads or writes.
 * kind: '3' [jde][kde][ip], '2' [jde][ip], '1' [kde], 'M' [nmoist][jde][kde][ip],
 * 'S' [nscalar][jde][kde][ip], 'X' [bw][kde][jde], 'Y' [bw][kde][ip],
 * 'x' [bw][jde], 'y' [bw][ip]; all float32. */
int wrfdc_field_count(void);
const char *wrfdc_field_name(int index);
int wrfdc_field_kind(int index);

int wrfdc_create(const wrfdc_config *cfg, wrfdc_ctx **out);
void wrfdc_destroy(wrfdc_ctx *ctx);
int wrfdc_bind(wrfdc_ctx *ctx, const char *name, void *device_ptr);
int wrfdc_set_stream(wrfdc_ctx *ctx, void *hip_stream);
int wrfdc_step(wrfdc_ctx *ctx, int nsteps, double dtbc0);
/* One stage of the step on the bound state.  name is a stage name of the
 * reference's solve_em driver (rk_step_prep, rk_phys_bc_dry_1,
 * first_rk_step_part1/2, rk_tendency, relax_bdy_dry, rk_addtend_dry,
 * spec_bdy_dry, small_step_prep, calc_p_rho, calc_coef_w, advance_uv,
 * spec_bdyupdate:u|v|t|mu_2|muts, advance_mu_t, advance_w, sumflux,
 * spec_bdyupdate_ph, zero_grad_bdy, calc_mu_uv_1, small_step_finish,
 * initialize_moist_old, rk_update_scalar_pd:moist_old|scalar_old,
 * moist_scalar_advance, scalar_advance, calc_p_rho_phi[:final],
 * rk_phys_bc_dry_2, set_phys_bc_dry_2, set_physical_bc3d:<field>[:small],
 * set_physical_bc2d:<field>[:small], spec_bdy_final:u|v|t|ph|mu|moist|moist_qv,
 * set_w_surface), one of its parts (advect_u/v/w, advect_scalar[_pd],
 * rhs_ph, horizontal_pressure_gradient, pg_buoy_w, compute_diff_metrics,
 * cal_deform_and_div, smag2d_km, calculate_n2, calculate_km_kh, closure_km,
 * horizontal_diffusion_2, dry_boundary_tendency, acoustic_substep,
 * w_solve:columns | w_solve:levels: the tridiagonal w solve alone, in place
 * on w_2, one column per lane or one column per wavefront) or a
 * whole rk_stage / step.  rk_step 1..3 sets dt_rk, dts_rk and the number of
 * acoustic steps as _rk_step_sizes does; iteration is the acoustic step
 * (calc_p_rho: 0 before the loop).  The lateral boundary time is the one of
 * wrfdc_set_dtbc (or of the last wrfdc_step).  WRFDC_EINVAL for an unknown
 * name or a stage of an unsupported option (spec_bdyupdate:w, nested). */
int wrfdc_stage(wrfdc_ctx *ctx, const char *name, int rk_step, int iteration);
/* Seconds since the lateral boundary record's time, as the next wrfdc_stage
 * calls read it (wrfdc_step sets it to dtbc0 + (n+1) dt itself). */
int wrfdc_set_dtbc(wrfdc_ctx *ctx, double dtbc);
int wrfdc_checksums(wrfdc_ctx *ctx, double out[WRFDC_NCHECKSUMS]);
/* No counterpart in the reference (the batched ensemble's autotuner uses
 * it): a launch-shape setting of the context.  "face_levels": levels per
 * level group in the tile order of the face-sharing stencil kernels
 * (advection, species), 1..32, or 0 for the per-kernel defaults (4, and 8
 * for advect_w and the positive-definite species passes).  Results do not
 * depend on it; it takes effect at the next wrfdc_step / wrfdc_stage (and is
 * baked into a captured ensemble graph).  WRFDC_EINVAL for another name or
 * value. */
int wrfdc_set_tile(wrfdc_ctx *ctx, const char *name, int value);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory or wait on the host (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the checksums of wrfdc_checksums
 * without a context, an allocation or a synchronisation.  fields holds the
 * device pointer of every registry field in wrfdc_field_name order; the ones
 * the sums read must be non-null (u_2 v_2 w_2 t_2 ph_2 p pb phb mu_2 mub c1h
 * c2h dnw msftx msfty moist).  Two kernels on hip_stream write
 * WRFDC_NCHECKSUMS doubles to out_dev (device memory); non-finite values are
 * counted in out_dev[3], not reported by the return code.  The partial sums
 * live in a fixed 24 KiB device array of the library (512 blocks x 6
 * doubles), one per GPU, shared by every call on that GPU: calls that overlap
 * on one GPU race.  Any domain wrfdc_create accepts fits (fewer than 2^30
 * elements per 3-D field); configurations it rejects are rejected here with
 * the same codes, messages via wrfdc_last_error(NULL). */
int wrfdc_checksums_noalloc(const wrfdc_config *cfg, void *const *fields, double *out_dev, void *hip_stream);
int wrfdc_sync(wrfdc_ctx *ctx);
/* ---- batched ensemble (north-star extension; the reference advances one
 * domain per run_direct_step, src/wrf_jax/core/direct_step.py:3425-3531) ----
 * One RK3 step of every member context is captured once into a HIP graph:
 * member q runs on stream branch_of[q] (NULL: q mod nbranches), members of
 * one stream in order, streams concurrently.  wrfdc_ensemble_step replays
 * the graph nsteps times on the ensemble's own stream, after setting each
 * member's boundary time dtbc0 + (n+1)*dt.  The members keep their buffers
 * and stay usable one by one (synchronise with wrfdc_ensemble_sync first);
 * binding other buffers to a member or destroying it invalidates the
 * ensemble. */
typedef struct wrfdc_ensemble wrfdc_ensemble;
int wrfdc_ensemble_create(wrfdc_ctx *const *members, int n, const int *branch_of, int nbranches, wrfdc_ensemble **out);
int wrfdc_ensemble_step(wrfdc_ensemble *ens, int nsteps, double dtbc0);
int wrfdc_ensemble_sync(wrfdc_ensemble *ens);
void *wrfdc_ensemble_stream(wrfdc_ensemble *ens);
long long wrfdc_ensemble_nodes(wrfdc_ensemble *ens);
void wrfdc_ensemble_destroy(wrfdc_ensemble *ens);

/* ---- tangent-linear and adjoint scalar transport (north-star extension: the
 * adjoint of advect_scalar, src/wrf_jax/dyn_em/module_advect_em_fast.py:1437-1769,
 * + rk_update_scalar, src/wrf_jax/dyn_em/module_em.py:1437-1598; no counterpart
 * interface in the reference) ----
 * Checkpoints: every RK stage of later wrfdc_step calls copies the dynamics its
 * species update reads -- ru_m, rv_m, ww_m (3-D) and mu_1, mu_2 (2-D) -- into
 * dst[5*s + 0..4] for stages s = 0 .. nstages-1 in order (nstages 0 stops).
 * A stage of transport with such a checkpoint: q_out = (m_old q_old +
 * dt_rk msfty [inner] A q_in) / m_new, A the advect_scalar operator of the
 * checkpoint's fluxes (linear in q, so this is also its tangent-linear); the
 * adjoint stage adds the transpose: qbar_old += (m_old/m_new) qbar_out,
 * qbar_in += A^T (dt_rk msfty [inner] / m_new) qbar_out (qbar_old and
 * qbar_in may be the same array).  3-D arrays in the registry layout. */
int wrfdc_record_stages(wrfdc_ctx *ctx, float *const *dst, int nstages);
int wrfdc_recorded_stages(wrfdc_ctx *ctx);
int wrfdc_transport_stage(wrfdc_ctx *ctx, int rk_step, const float *const *stage, const float *q_in, const float *q_old, float *q_out);
int wrfdc_transport_stage_adjoint(wrfdc_ctx *ctx, int rk_step, const float *const *stage, const float *qbar_out, float *qbar_old,
                                  float *qbar_in);

/* ---- learned eddy viscosity (north-star extension of smag_km, the km of
 * calculate_km_kh, src/wrf_jax/dyn_em/module_diffusion_em_fast.py:773-1083;
 * no counterpart interface in the reference) ----
 * A 55 -> 64 -> 64 -> 1 network (bf16 MFMA hidden layers, SiLU, softplus
 * output) gives km from the 3x3x3 stencils of log|D| and N^2 and log(dx).
 * Parameters: one FP32 vector theta of WRFDC_MLP_NPARAMS (W1[64][64], b1,
 * W2[64][64], b2, w3, b3) and its bf16 weight copies wbf (W1, W2, W2^T;
 * WRFDC_MLP_WBF_SIZE bf16 values), both in device memory. */
#define WRFDC_MLP_NPARAMS 8385
#define WRFDC_MLP_WBF_SIZE 12288
#define WRFDC_MLP_FEATURES 64 /* bf16 per training row (55 used) */
#define WRFDC_MLP_TARGETS 6   /* floats per training row: t1 t2 s1 s2 divisor km_smag */
typedef struct wrfdc_closure_norm {
  float d_shift, d_scale; /* log|D| features: (log(|D| + 1e-7) - d_shift) / d_scale */
  float n_shift, n_scale; /* N^2 features: (N^2 - n_shift) / n_scale */
  float dx_feature;       /* log(dx / 1000 m) */
  float kref;             /* km = kref * softplus(network output) */
} wrfdc_closure_norm;

/* Use the network for km in the step's first-RK-stage diffusion (the slope
 * divisor and the 10 l cap still apply; a non-finite output falls back to
 * Smagorinsky).  theta NULL switches back to Smagorinsky.  The pointers are
 * kept, not copied. */
int wrfdc_set_closure(wrfdc_ctx *ctx, const float *theta, const void *wbf, const wrfdc_closure_norm *norm);
/* Training rows of the current state, one per point of the smag2d_km range
 * (i 2..nx-1, k 1..nz, j 2..ny-1; row = ((j-2)*nz + k-1)*(nx-2) + i-2):
 * features (bf16 [n][64]) and target columns 2..5 (s1 = (D11-D22)/2,
 * s2 = D12, slope divisor, Smagorinsky km) of rows_t ([n][6]).  Runs the
 * diffusion metrics, deformation, N^2 and Smagorinsky kernels first. */
long long wrfdc_closure_row_count(const wrfdc_config *cfg);
int wrfdc_closure_rows(wrfdc_ctx *ctx, const wrfdc_closure_norm *norm, void *rows_x, float *rows_t);
/* Columns 0..1 of the rows of a coarse grid (nx_coarse x ny_coarse, the fine
 * grid divided by ratio): t1 = (tau11 - tau22)/2 and t2 = tau12, the
 * horizontal sub-grid stresses of the fine domain's winds over each coarse
 * cell. */
int wrfdc_subgrid_stress(wrfdc_ctx *fine, int nx_coarse, int ny_coarse, int ratio, float *rows_t);
/* Block means of the fine domain's dynamic fields (winds on their faces,
 * w, theta, geopotential, pressure, inverse density, column mass, terrain
 * and moisture) into the coarse domain's buffers, on the coarse stream. */
int wrfdc_coarse_grain(wrfdc_ctx *fine, wrfdc_ctx *coarse, int ratio);
/* Training: gradient of the loss sum over rows of
 * ((t1 + K s1)^2 + (t2 + K s2)^2) * lossnorm, K = kref*softplus(z)/divisor
 * (work: wrfdc_mlp_work_floats(nrows) floats; grad: WRFDC_MLP_NPARAMS;
 * loss: 1 float), or that gradient followed by one Adam step on theta and
 * wbf (moments adam_m, adam_v of WRFDC_MLP_NPARAMS, step >= 1).  All device
 * pointers, stream-ordered, no allocation. */
long long wrfdc_mlp_work_floats(long long nrows);
int wrfdc_mlp_pack(const float *theta, void *wbf, void *hip_stream);
int wrfdc_mlp_grad(const float *theta, const void *wbf, const void *rows_x, const float *rows_t, long long nrows, float lossnorm, float kref,
                   float *work, float *grad, float *loss, void *hip_stream);
int wrfdc_mlp_train_step(float *theta, void *wbf, const void *rows_x, const float *rows_t, long long nrows, float lossnorm, float kref,
                         float *work, float *adam_m, float *adam_v, int step, float lr, float *loss, void *hip_stream);
/* Network output z (before softplus) of each row. */
int wrfdc_mlp_forward(const float *theta, const void *wbf, const void *rows_x, long long nrows, float *z, void *hip_stream);

/* Message of the last failure on ctx; with ctx NULL, of the last failure of
 * wrfdc_create (or another entry point without a context) on this thread. */
const char *wrfdc_last_error(wrfdc_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
