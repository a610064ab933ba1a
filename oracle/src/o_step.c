/*
 * Oracle: RK3 step orchestration and the ctypes-facing API
 * (test infrastructure only).  The stage order follows
 * src/wrf_jax/core/direct_step.py:3037-3422 (RK / acoustic loop),
 * src/wrf_jax/dyn_em/solve_em.py:260-800 (final boundary stages) and the
 * moist / scalar stage adapters in src/wrf_jax/runtime.py:13652-14830.
 */
#include <stdio.h>

#include "o_core.h"

/* o_bc.c */
void o_bc3(OCtx *x, real *fl, char var);
void o_bc2(OCtx *x, real *fl, char var);
void o_spec_overlay(OCtx *x, real *tend, char var, int ktop, int is2d, real *bt_xs, real *bt_xe, real *bt_ys, real *bt_ye);
void o_spec_bdyupdate(OCtx *x, real *fl, const real *tend, double dt, char var, int is2d);
void o_spec_bdyupdate_ph(OCtx *x, double dts);
void o_zero_grad_bdy_w(OCtx *x, real *w);
void o_relax_bdytend(OCtx *x, const real *fld, real *tend, char var, int ktop, int is2d, real *b_xs, real *b_xe, real *b_ys, real *b_ye,
                     real *bt_xs, real *bt_xe, real *bt_ys, real *bt_ye);
void o_flow_dep_bdy(OCtx *x, real *q, const real *u, const real *v);
void o_spec_bdy_final(OCtx *x, real *fl, char var, int is2d, const real *mu, const real *msf, const real *c1, const real *c2,
                      real *b_xs, real *b_xe, real *b_ys, real *b_ye, real *bt_xs, real *bt_xe, real *bt_ys, real *bt_ye);
void o_set_w_surface(OCtx *x, real *w, const real *u, const real *v);
/* o_advect.c */
void o_advect_scalar(OCtx *x, const real *q, real *tend, const real *ru, const real *rv, const real *rom);
void o_advect_u(OCtx *x, const real *u, real *tend, const real *ru, const real *rv, const real *rom);
void o_advect_v(OCtx *x, const real *v, real *tend, const real *ru, const real *rv, const real *rom);
void o_advect_w(OCtx *x, const real *w, real *tend, const real *ru, const real *rv, const real *rom);
void o_advect_scalar_pd(OCtx *x, const real *q, const real *q_old, real *tend, const real *ru, const real *rv, const real *rom,
                        const real *mut, const real *mub, const real *mu_old, double dt);
/* o_bigstep.c */
void o_rk_step_prep(OCtx *x);
void o_rhs_ph(OCtx *x);
void o_horizontal_pressure_gradient(OCtx *x);
void o_pg_buoy_w(OCtx *x);
void o_coriolis(OCtx *x);
void o_curvature(OCtx *x);
void o_calc_p_rho_phi(OCtx *x);
void o_rk_addtend_dry(OCtx *x);
/* o_small.c */
void o_small_step_prep(OCtx *x);
void o_calc_p_rho(OCtx *x, int step);
void o_calc_coef_w(OCtx *x);
void o_advance_uv(OCtx *x);
void o_advance_mu_t(OCtx *x);
void o_advance_w(OCtx *x);
void o_sumflux(OCtx *x);
void o_calc_mu_uv_1(OCtx *x);
void o_small_step_finish(OCtx *x, int final_step);
/* o_diff.c */
void o_phy_prep_rho(OCtx *x);
void o_compute_diff_metrics(OCtx *x);
void o_cal_deform_and_div(OCtx *x);
void o_smag2d_km(OCtx *x);
void o_calc_n2(OCtx *x);
void o_horizontal_diffusion_2(OCtx *x);

static const char *o_names[] = {
#define O_NAME(name, kind) #name,
    O_FIELDS(O_NAME)
#undef O_NAME
};
static const char o_kinds[] = {
#define O_KIND(name, kind) kind,
    O_FIELDS(O_KIND)
#undef O_KIND
};

int o_field_count(void) { return O_NFIELDS; }
const char *o_field_name(int i) { return (i >= 0 && i < O_NFIELDS) ? o_names[i] : 0; }
int o_field_kind(int i) { return (i >= 0 && i < O_NFIELDS) ? o_kinds[i] : 0; }
int o_real_size(void) { return (int)sizeof(real); }
int o_cfg_size(void) { return (int)sizeof(OCfg); }

OCtx *o_create(const OCfg *cfg) {
  OCtx *x = (OCtx *)calloc(1, sizeof(OCtx));
  x->c = *cfg;
  x->ide = cfg->nx + 1;
  x->jde = cfg->ny + 1;
  x->kde = cfg->nz + 1;
  x->n3 = (size_t)x->jde * x->kde * cfg->ip;
  x->n2 = (size_t)x->jde * cfg->ip;
  int w = cfg->spec_bdy_width;
  x->fcx = (real *)calloc((size_t)w + 2, sizeof(real));
  x->gcx = (real *)calloc((size_t)w + 2, sizeof(real));
  /* lbc_fcx_gcx (boundary/lbc_fcx_gcx.py:10-56), spec_exp = 0 */
  for (int b = cfg->spec_zone + 1; b <= cfg->spec_zone + cfg->relax_zone && b <= w; b++) {
    double dist = (double)(cfg->spec_zone + cfg->relax_zone) - (double)b;
    double den = (double)(cfg->relax_zone - 1);
    x->fcx[b] = (real)((real)((real)0.1 / (real)cfg->dt) * (real)dist / (real)den);
    x->gcx[b] = (real)((real)((real)1.0 / (real)cfg->dt / (real)50.0) * (real)dist / (real)den);
  }
  return x;
}

void o_destroy(OCtx *x) {
  if (!x) return;
  free(x->fcx);
  free(x->gcx);
  free(x);
}

int o_bind(OCtx *x, const char *name, void *ptr) {
  for (int i = 0; i < O_NFIELDS; i++)
    if (strcmp(o_names[i], name) == 0) {
      x->F[i] = (real *)ptr;
      return 0;
    }
  return -1;
}

void o_set_dtbc(OCtx *x, double dtbc) { x->c.dtbc = dtbc; }

/* WRF's acoustic step count (solve_em.py:220-240). */
int o_num_sound_steps(const OCfg *c) {
  if (c->time_step_sound != 0) return c->time_step_sound;
  double sp = c->dx < c->dy ? c->dx : c->dy;
  int n = 2 * ((int)(300.0 * c->dt / sp - 0.01) + 1);
  return n > 4 ? n : 4;
}

/* _rk_step_sizes (solve_em.py:243-257), rk_order = 3. */
void o_set_rk_stage(OCtx *x, int rk_step) {
  double dt = x->c.dt;
  int ns = o_num_sound_steps(&x->c);
  double dts = dt / (double)ns;
  x->rk_step = rk_step;
  if (rk_step == 1) { x->dt_rk = dt / 3.0; x->dts_rk = x->dt_rk; x->nsmall = 1; }
  else if (rk_step == 2) { x->dt_rk = 0.5 * dt; x->dts_rk = dts; x->nsmall = ns / 2 > 1 ? ns / 2 : 1; }
  else { x->dt_rk = dt; x->dts_rk = dts; x->nsmall = ns; }
}

/* rk_tendency as fused in runtime.py:1528-2131. */
void o_rk_tendency(OCtx *x) {
  memset(FLD(ru_tend), 0, sizeof(real) * x->n3);
  memset(FLD(rv_tend), 0, sizeof(real) * x->n3);
  memset(FLD(rw_tend), 0, sizeof(real) * x->n3);
  memset(FLD(t_tend), 0, sizeof(real) * x->n3);
  memset(FLD(ph_tend), 0, sizeof(real) * x->n3);
  memset(FLD(u_save), 0, sizeof(real) * x->n3);
  memset(FLD(v_save), 0, sizeof(real) * x->n3);
  memset(FLD(w_save), 0, sizeof(real) * x->n3);
  memset(FLD(ph_save), 0, sizeof(real) * x->n3);
  memset(FLD(t_save), 0, sizeof(real) * x->n3);
  memset(FLD(mu_tend), 0, sizeof(real) * x->n2);
  memset(FLD(mu_save), 0, sizeof(real) * x->n2);
  real *ru = FLD(ru), *rv = FLD(rv), *ww = FLD(ww);
  o_advect_u(x, FLD(u_2), FLD(ru_tend), ru, rv, ww);
  o_advect_v(x, FLD(v_2), FLD(rv_tend), ru, rv, ww);
  o_advect_w(x, FLD(w_2), FLD(rw_tend), ru, rv, ww);
  o_advect_scalar(x, FLD(t_2), FLD(t_tend), ru, rv, ww);
  o_rhs_ph(x);
  o_horizontal_pressure_gradient(x);
  o_pg_buoy_w(x);
  o_coriolis(x);
  o_curvature(x);
}

/* first_rk_step_part1 with physics off: init_zero_tendency + phy_prep rho
 * (module_first_rk_step_part1.py:112-417). */
void o_first_rk_step_part1(OCtx *x) {
  memset(FLD(ru_tendf), 0, sizeof(real) * x->n3);
  memset(FLD(rv_tendf), 0, sizeof(real) * x->n3);
  memset(FLD(rw_tendf), 0, sizeof(real) * x->n3);
  memset(FLD(ph_tendf), 0, sizeof(real) * x->n3);
  memset(FLD(t_tendf), 0, sizeof(real) * x->n3);
  memset(FLD(mu_tendf), 0, sizeof(real) * x->n2);
  memset(FLD(moist_tend), 0, sizeof(real) * x->n3 * (size_t)x->c.nmoist);
  if (x->c.nscalar > 0) memset(FLD(scalar_tend), 0, sizeof(real) * x->n3 * (size_t)x->c.nscalar);
  o_phy_prep_rho(x);
}

/* first_rk_step_part2, diff_opt 2 / km_opt 4 (runtime.py:15936-16042). */
void o_first_rk_step_part2(OCtx *x) {
  o_compute_diff_metrics(x);
  o_bc3(x, FLD(rdzw), 'w');
  o_bc3(x, FLD(rdz), 'w');
  o_bc3(x, FLD(z), 'w');
  o_bc3(x, FLD(zx), 'e');
  o_bc3(x, FLD(zy), 'f');
  o_cal_deform_and_div(x);
  /* calculate_km_kh, km_opt 4: N^2 (an output no later stage reads here)
   * and the Smagorinsky diffusivities */
  o_calc_n2(x);
  o_smag2d_km(x);
  /* phy_bc (module_diffusion_em.py:180-363) */
  o_bc3(x, FLD(xkmh), 't');
  o_bc3(x, FLD(xkhh), 't');
  o_bc3(x, FLD(xkmv), 't');
  o_bc3(x, FLD(xkhv), 't');
  o_bc3(x, FLD(div), 't');
  o_bc3(x, FLD(defor11), 't');
  o_bc3(x, FLD(defor22), 't');
  o_bc3(x, FLD(defor33), 't');
  o_bc3(x, FLD(defor12), 'd');
  o_bc3(x, FLD(defor13), 'e');
  o_bc3(x, FLD(defor23), 'f');
  o_bc3(x, FLD(rho), 't');
  o_horizontal_diffusion_2(x);
}

/* relax_bdy_dry (module_bc_em.py:627-760): relaxation-zone tendencies of
 * u, v, ph, theta (the last two mass-weighted) and mu into the physics
 * tendencies. */
void o_relax_bdy_dry(OCtx *x) {
  real *mut = FLD(mut);
  {
    real *tmp = (real *)calloc(x->n3, sizeof(real));
    o_relax_bdytend(x, FLD(ru), FLD(ru_tendf), 'u', 0, 0, FLD(u_bxs), FLD(u_bxe), FLD(u_bys), FLD(u_bye), FLD(u_btxs), FLD(u_btxe), FLD(u_btys), FLD(u_btye));
    o_relax_bdytend(x, FLD(rv), FLD(rv_tendf), 'v', 0, 0, FLD(v_bxs), FLD(v_bxe), FLD(v_bys), FLD(v_bye), FLD(v_btxs), FLD(v_btxe), FLD(v_btys), FLD(v_btye));
    /* mass_weight (module_bc_em.py:561-624) over the tile + 1 */
    real *ph = FLD(ph_2), *c1f = FLD(c1f), *c2f = FLD(c2f);
    for (int j = 1; j <= JDE - 1; j++)
      for (int k = 1; k <= KDE; k++)
        for (int i = 1; i <= IDE - 1; i++) A3(tmp, i, k, j) = A3(ph, i, k, j) * (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k));
    o_relax_bdytend(x, tmp, FLD(ph_tendf), 'h', 1, 0, FLD(ph_bxs), FLD(ph_bxe), FLD(ph_bys), FLD(ph_bye), FLD(ph_btxs), FLD(ph_btxe), FLD(ph_btys), FLD(ph_btye));
    real *t = FLD(t_2), *c1h = FLD(c1h), *c2h = FLD(c2h);
    memset(tmp, 0, sizeof(real) * x->n3);
    for (int j = 1; j <= JDE - 1; j++)
      for (int k = 1; k <= KDE - 1; k++)
        for (int i = 1; i <= IDE - 1; i++) A3(tmp, i, k, j) = A3(t, i, k, j) * (A1(c1h, k) * A2(mut, i, j) + A1(c2h, k));
    o_relax_bdytend(x, tmp, FLD(t_tendf), 't', 0, 0, FLD(t_bxs), FLD(t_bxe), FLD(t_bys), FLD(t_bye), FLD(t_btxs), FLD(t_btxe), FLD(t_btys), FLD(t_btye));
    o_relax_bdytend(x, FLD(mu_2), FLD(mu_tend), 'm', 1, 1, FLD(mu_bxs), FLD(mu_bxe), FLD(mu_bys), FLD(mu_bye), FLD(mu_btxs), FLD(mu_btxe), FLD(mu_btys), FLD(mu_btye));
    free(tmp);
  }
}

/* spec_bdy_dry (module_bc_em.py:1127-1332): the specified zone's tendencies
 * from the boundary tendencies. */
void o_spec_bdy_dry(OCtx *x) {
  o_spec_overlay(x, FLD(ru_tend), 'u', 0, 0, FLD(u_btxs), FLD(u_btxe), FLD(u_btys), FLD(u_btye));
  o_spec_overlay(x, FLD(rv_tend), 'v', 0, 0, FLD(v_btxs), FLD(v_btxe), FLD(v_btys), FLD(v_btye));
  o_spec_overlay(x, FLD(ph_tend), 'h', 1, 0, FLD(ph_btxs), FLD(ph_btxe), FLD(ph_btys), FLD(ph_btye));
  o_spec_overlay(x, FLD(t_tend), 't', 0, 0, FLD(t_btxs), FLD(t_btxe), FLD(t_btys), FLD(t_btye));
  o_spec_overlay(x, FLD(mu_tend), 'm', 1, 1, FLD(mu_btxs), FLD(mu_btxe), FLD(mu_btys), FLD(mu_btye));
}

/* relax_bdy_dry (first stage) + rk_addtend_dry + spec_bdy_dry
 * (direct_step.py:3179-3188). */
void o_dry_boundary_tendency(OCtx *x) {
  if (x->rk_step == 1) o_relax_bdy_dry(x);
  o_rk_addtend_dry(x);
  o_spec_bdy_dry(x);
}

/* One species of moist_scalar_advance / scalar_advance
 * (runtime.py:13900-14330, 14491-14830): advection with the time-averaged
 * mass fluxes (positive definite on the last stage), qv boundary
 * relaxation at rk1, rk_update_scalar (module_em.py:1437-1598), and
 * flow_dep_bdy for non-qv moist species. */
static void o_species_advance(OCtx *x, real *q, real *q_old, real *q_tend, int is_qv, int is_moist, int pd) {
  real *adv = FLD(advect_tend);
  memset(adv, 0, sizeof(real) * x->n3);
  int final = (x->rk_step == 3);
  if (final && pd)
    o_advect_scalar_pd(x, q, q_old, adv, FLD(ru_m), FLD(rv_m), FLD(ww_m), FLD(muts), FLD(mub), FLD(mu_1), x->dt_rk);
  else
    o_advect_scalar(x, q, adv, FLD(ru_m), FLD(rv_m), FLD(ww_m));
  if (is_qv && x->rk_step == 1) {
    /* relax_bdy_scalar + spec_bdy_scalar overlaid on the specified zone */
    real *tmp = (real *)calloc(x->n3, sizeof(real)), *mut = FLD(mut), *c1h = FLD(c1h), *c2h = FLD(c2h);
    for (int j = 1; j <= JDE - 1; j++)
      for (int k = 1; k <= KDE - 1; k++)
        for (int i = 1; i <= IDE - 1; i++) A3(tmp, i, k, j) = A3(q, i, k, j) * (A1(c1h, k) * A2(mut, i, j) + A1(c2h, k));
    o_relax_bdytend(x, tmp, q_tend, 'q', 0, 0, FLD(qv_bxs), FLD(qv_bxe), FLD(qv_bys), FLD(qv_bye), FLD(qv_btxs), FLD(qv_btxe), FLD(qv_btys), FLD(qv_btye));
    o_spec_overlay(x, q_tend, 't', 0, 0, FLD(qv_btxs), FLD(qv_btxe), FLD(qv_btys), FLD(qv_btye));
    free(tmp);
  }
  real *msfty = FLD(msfty), *c1 = FLD(c1h), *c2 = FLD(c2h), *mu_1 = FLD(mu_1), *mu_2 = FLD(mu_2), *mub = FLD(mub);
  real dt = (real)x->dt_rk;
  int sz = x->c.spec_zone;
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        int inner = (i >= 1 + sz && i <= IDE - sz - 1 && j >= 1 + sz && j <= JDE - sz - 1);
        real tend = inner ? A3(adv, i, k, j) * A2(msfty, i, j) : (real)0;
        tend = tend + A3(q_tend, i, k, j);
        real muold = A2(mu_1, i, j) + A2(mub, i, j), munew = A2(mu_2, i, j) + A2(mub, i, j);
        real oldm = A1(c1, k) * muold + A1(c2, k), newm = A1(c1, k) * munew + A1(c2, k);
        real old;
        if (x->rk_step == 1) {
          old = A3(q, i, k, j);
          A3(q_old, i, k, j) = old;
        } else {
          old = A3(q_old, i, k, j);
        }
        A3(q, i, k, j) = (oldm * old + dt * tend) / newm;
      }
  if (is_moist && !is_qv) o_flow_dep_bdy(x, q, FLD(ru_m), FLD(rv_m));
}

/* rk_update_scalar_pd (runtime.py:14422-14489 / module_em.py:1601-1669):
 * apply the accumulated physics tendency to the old scalar on the last stage
 * (mu_old = mu_new = mu_1), then zero it. */
static void o_rk_update_scalar_pd(OCtx *x, real *q_old, real *q_tend) {
  real *c1 = FLD(c1h), *c2 = FLD(c2h), *mu_1 = FLD(mu_1), *mub = FLD(mub);
  real dt = (real)x->dt_rk;
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real m = A1(c1, k) * (A2(mu_1, i, j) + A2(mub, i, j)) + A1(c2, k);
        A3(q_old, i, k, j) = (m * A3(q_old, i, k, j) + dt * A3(q_tend, i, k, j)) / m;
        A3(q_tend, i, k, j) = 0;
      }
}

/* One acoustic substep (direct_step.py:3224-3390). */
void o_acoustic_substep(OCtx *x, int it) {
  x->iteration = it;
  o_advance_uv(x);
  o_spec_bdyupdate(x, FLD(u_2), FLD(ru_tend), x->dts_rk, 'u', 0);
  o_spec_bdyupdate(x, FLD(v_2), FLD(rv_tend), x->dts_rk, 'v', 0);
  o_advance_mu_t(x);
  o_spec_bdyupdate(x, FLD(t_2), FLD(t_tend), x->dts_rk, 't', 0);
  o_spec_bdyupdate(x, FLD(mu_2), FLD(mu_tend), x->dts_rk, 'm', 1);
  o_spec_bdyupdate(x, FLD(muts), FLD(mu_tend), x->dts_rk, 'm', 1);
  o_advance_w(x);
  o_sumflux(x);
  o_spec_bdyupdate_ph(x, x->dts_rk);
  o_zero_grad_bdy_w(x, FLD(w_2));
  o_calc_p_rho(x, it);
  o_bc3(x, FLD(ph_2), 'w'); o_bc3(x, FLD(al), 'p'); o_bc3(x, FLD(p), 'p');
  o_bc2(x, FLD(muts), 't'); o_bc2(x, FLD(mu_2), 't'); o_bc2(x, FLD(mudf), 't');
}

/* rk_phys_bc_dry_1 (module_bc_em.py:109-216) */
static void o_rk_phys_bc_dry_1(OCtx *x) {
  o_bc3(x, FLD(ru), 'u'); o_bc3(x, FLD(rv), 'v'); o_bc3(x, FLD(rw), 'w'); o_bc3(x, FLD(ww), 'w');
  o_bc3(x, FLD(php), 'w'); o_bc3(x, FLD(alt), 't'); o_bc3(x, FLD(p), 'p');
  o_bc2(x, FLD(muu), 'u'); o_bc2(x, FLD(muv), 'v'); o_bc2(x, FLD(mut), 't');
}

/* rk_phys_bc_dry_2 (module_bc_em.py:219-318) */
static void o_rk_phys_bc_dry_2(OCtx *x) {
  o_bc3(x, FLD(u_2), 'u'); o_bc3(x, FLD(v_2), 'v'); o_bc3(x, FLD(w_2), 'w'); o_bc3(x, FLD(t_2), 'p');
  o_bc3(x, FLD(ph_2), 'w'); o_bc2(x, FLD(mu_2), 't');
}

/* set_phys_bc_dry_2 (module_bc_em.py:321-445) */
static void o_set_phys_bc_dry_2(OCtx *x) {
  o_bc3(x, FLD(u_1), 'u'); o_bc3(x, FLD(u_2), 'u'); o_bc3(x, FLD(v_1), 'v'); o_bc3(x, FLD(v_2), 'v');
  o_bc3(x, FLD(w_1), 'w'); o_bc3(x, FLD(w_2), 'w'); o_bc3(x, FLD(t_1), 'p'); o_bc3(x, FLD(t_2), 'p');
  o_bc3(x, FLD(ph_1), 'w'); o_bc3(x, FLD(ph_2), 'w'); o_bc2(x, FLD(mu_1), 't'); o_bc2(x, FLD(mu_2), 't');
}

/* spec_bdy_final of one field (solve_em.py:744-760): 'u' 'v' 't' 'h' (ph)
 * 'm' (mu) or 'q' (QV, the only species with boundary data) */
static void o_spec_bdy_final_of(OCtx *x, char w) {
  switch (w) {
    case 'u': o_spec_bdy_final(x, FLD(u_2), 'u', 0, FLD(muus), FLD(msfuy), FLD(c1h), FLD(c2h), FLD(u_bxs), FLD(u_bxe), FLD(u_bys), FLD(u_bye), FLD(u_btxs), FLD(u_btxe), FLD(u_btys), FLD(u_btye)); break;
    case 'v': o_spec_bdy_final(x, FLD(v_2), 'v', 0, FLD(muvs), FLD(msfvx), FLD(c1h), FLD(c2h), FLD(v_bxs), FLD(v_bxe), FLD(v_bys), FLD(v_bye), FLD(v_btxs), FLD(v_btxe), FLD(v_btys), FLD(v_btye)); break;
    case 't': o_spec_bdy_final(x, FLD(t_2), 't', 0, FLD(muts), FLD(msfty), FLD(c1h), FLD(c2h), FLD(t_bxs), FLD(t_bxe), FLD(t_bys), FLD(t_bye), FLD(t_btxs), FLD(t_btxe), FLD(t_btys), FLD(t_btye)); break;
    case 'h': o_spec_bdy_final(x, FLD(ph_2), 'h', 0, FLD(muts), FLD(msfty), FLD(c1f), FLD(c2f), FLD(ph_bxs), FLD(ph_bxe), FLD(ph_bys), FLD(ph_bye), FLD(ph_btxs), FLD(ph_btxe), FLD(ph_btys), FLD(ph_btye)); break;
    case 'm': o_spec_bdy_final(x, FLD(mu_2), 'm', 1, FLD(muts), FLD(msfty), FLD(c1h), FLD(c2h), FLD(mu_bxs), FLD(mu_bxe), FLD(mu_bys), FLD(mu_bye), FLD(mu_btxs), FLD(mu_btxe), FLD(mu_btys), FLD(mu_btye)); break;
    default: o_spec_bdy_final(x, SPC(moist, 0), 't', 0, FLD(muts), FLD(msfty), FLD(c1h), FLD(c2h), FLD(qv_bxs), FLD(qv_bxe), FLD(qv_bys), FLD(qv_bye), FLD(qv_btxs), FLD(qv_btxe), FLD(qv_btys), FLD(qv_btye)); break;
  }
}

/* initialize_moist_old (module_big_step_utilities_em.py:935-975): moist_old
 * <- moist on the mass points */
static void o_initialize_moist_old(OCtx *x) {
  for (int s = 0; s < x->c.nmoist; s++) {
    real *q = SPC(moist, s), *qo = SPC(moist_old, s);
    for (int j = 1; j <= JDE - 1; j++)
      for (int k = 1; k <= KDE - 1; k++)
        for (int i = 1; i <= IDE - 1; i++) A3(qo, i, k, j) = A3(q, i, k, j);
  }
}

static void o_moist_scalar_advance(OCtx *x) {
  for (int s = 0; s < x->c.nmoist; s++) o_species_advance(x, SPC(moist, s), SPC(moist_old, s), SPC(moist_tend, s), s == 0, 1, x->c.pd_moist);
}

static void o_scalar_advance(OCtx *x) {
  for (int s = 0; s < x->c.nscalar; s++) o_species_advance(x, SPC(scalar, s), SPC(scalar_old, s), SPC(scalar_tend, s), 0, 0, x->c.pd_scalar);
}

/* One RK stage (direct_step.py:3120-3422). */
void o_rk_stage(OCtx *x, int rk_step) {
  o_set_rk_stage(x, rk_step);
  o_rk_step_prep(x);
  /* rk_phys_bc_dry_1 + rho/al/ph_2 physical BC (module_bc_em.py:109-216) */
  o_rk_phys_bc_dry_1(x);
  o_bc3(x, FLD(al), 'p'); o_bc3(x, FLD(ph_2), 'w');
  if (rk_step == 1) {
    o_first_rk_step_part1(x);
    o_first_rk_step_part2(x);
  }
  o_rk_tendency(x);
  o_dry_boundary_tendency(x);
  o_small_step_prep(x);
  o_calc_p_rho(x, 0);
  o_calc_coef_w(x);
  /* pre-small physical BC (direct_step.py:3197-3222) */
  o_bc3(x, FLD(ru_tend), 'u'); o_bc3(x, FLD(rv_tend), 'v'); o_bc3(x, FLD(ph_2), 'w'); o_bc3(x, FLD(al), 'p');
  o_bc3(x, FLD(p), 'p'); o_bc3(x, FLD(t_1), 'p'); o_bc3(x, FLD(t_save), 't');
  o_bc2(x, FLD(mu_1), 't'); o_bc2(x, FLD(mu_2), 't'); o_bc2(x, FLD(mudf), 't');
  for (int it = 1; it <= x->nsmall; it++) o_acoustic_substep(x, it);
  o_calc_mu_uv_1(x);
  o_small_step_finish(x, rk_step == 3);
  if (rk_step == 3) {
    o_bc3(x, FLD(ru_m), 'u'); o_bc3(x, FLD(rv_m), 'v'); o_bc3(x, FLD(ww_m), 'w');
    o_bc2(x, FLD(mut), 't'); o_bc2(x, FLD(muts), 't');
  }
  int nm = x->c.nmoist, nsc = x->c.nscalar;
  if (rk_step == 3 && x->c.pd_moist) {
    for (int s = 0; s < nm; s++) {
      o_rk_update_scalar_pd(x, SPC(moist_old, s), SPC(moist_tend, s));
      o_bc3(x, SPC(moist_old, s), 'p');
    }
  }
  if (rk_step == 3 && x->c.pd_scalar) {
    for (int s = 0; s < nsc; s++) {
      o_rk_update_scalar_pd(x, SPC(scalar_old, s), SPC(scalar_tend, s));
      o_bc3(x, SPC(scalar_old, s), 'p');
    }
  }
  o_moist_scalar_advance(x);
  o_scalar_advance(x);
  o_calc_p_rho_phi(x);
  if (rk_step < 3) {
    /* rk_phys_bc_dry_2 + moist / scalar BC */
    o_rk_phys_bc_dry_2(x);
    for (int s = 0; s < nm; s++) o_bc3(x, SPC(moist, s), 'p');
    for (int s = 0; s < nsc; s++) o_bc3(x, SPC(scalar, s), 'p');
  }
}

/* One full model step: initial moist_old copy, three RK stages, then the
 * final physical / specified boundary stages and set_w_surface. */
void o_step(OCtx *x) {
  int nm = x->c.nmoist;
  o_initialize_moist_old(x);
  for (int s = 0; s < nm; s++) o_bc3(x, SPC(moist_old, s), 'p');
  for (int rk = 1; rk <= 3; rk++) o_rk_stage(x, rk);
  o_set_phys_bc_dry_2(x);
  o_bc3(x, SPC(moist, 0), 'p');
  /* spec_bdy_final (solve_em.py:744-760) */
  o_spec_bdy_final_of(x, 'u');
  o_spec_bdy_final_of(x, 'v');
  o_spec_bdy_final_of(x, 't');
  o_spec_bdy_final_of(x, 'h');
  o_spec_bdy_final_of(x, 'm');
  o_spec_bdy_final_of(x, 'q');
  o_set_w_surface(x, FLD(w_2), FLD(u_2), FLD(v_2));
}

/* Several independent domains advanced nsteps steps each, concurrently, one
 * OpenMP thread per domain at a time (the all-cores CPU baseline over the
 * tiles of a workload; built without -fopenmp the loop runs serially). */
int o_step_many(OCtx **ctxs, int n, int nsteps, double dtbc0, int threads) {
  if (n < 0 || nsteps < 0 || threads < 1) return -1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (int q = 0; q < n; q++)
    for (int s = 0; s < nsteps; s++) {
      o_set_dtbc(ctxs[q], dtbc0 + (double)(s + 1) * ctxs[q]->c.dt);
      o_step(ctxs[q]);
    }
  return 0;
}

static int o_field_index(const char *name) {
  for (int i = 0; i < O_NFIELDS; i++)
    if (strcmp(o_names[i], name) == 0) return i;
  return -1;
}

/* set_physical_bc3d:<field>[:small] / set_physical_bc2d:<field>[:small]
 * (solve_em.py:493-676); kinds of runtime.py:7421-7455.  Returns 1 if the
 * name is such a stage and was run, -1 if it names an unknown field, 0 if
 * it is no boundary stage. */
static int o_named_bc(OCtx *x, const char *name) {
  static const char *p3 = "set_physical_bc3d:", *p2 = "set_physical_bc2d:";
  int is3 = strncmp(name, p3, strlen(p3)) == 0, is2 = strncmp(name, p2, strlen(p2)) == 0;
  if (!is3 && !is2) return 0;
  char f[64];
  snprintf(f, sizeof f, "%s", name + strlen(p3));
  char *colon = strchr(f, ':');
  if (colon) {
    if (strcmp(colon, ":small") != 0) return -1;
    *colon = 0;
  }
  if (is2) {
    static const char *ok2[] = {"mu_1", "mu_2", "mudf", "muts", "mut"};
    for (size_t i = 0; i < sizeof ok2 / sizeof ok2[0]; i++)
      if (strcmp(f, ok2[i]) == 0) {
        o_bc2(x, x->F[o_field_index(f)], 't');
        return 1;
      }
    return -1;
  }
  if (!strcmp(f, "tke_1") || !strcmp(f, "tke_2")) return 1;
  if (!strcmp(f, "moist") || !strcmp(f, "moist_old")) {
    for (int s = 0; s < x->c.nmoist; s++) o_bc3(x, !strcmp(f, "moist") ? SPC(moist, s) : SPC(moist_old, s), 'p');
    return 1;
  }
  if (!strcmp(f, "scalar") || !strcmp(f, "scalar_old")) {
    for (int s = 0; s < x->c.nscalar; s++) o_bc3(x, !strcmp(f, "scalar") ? SPC(scalar, s) : SPC(scalar_old, s), 'p');
    return 1;
  }
  if (!strcmp(f, "moist_qv")) {
    o_bc3(x, SPC(moist, 0), 'p');
    return 1;
  }
  static const struct { const char *name; char kind; } k3[] = {
      {"rho", 'p'}, {"al", 'p'}, {"p", 'p'}, {"ph_2", 'w'}, {"t_1", 'p'}, {"t_save", 't'}, {"ru_tend", 'u'},
      {"rv_tend", 'v'}, {"ru_m", 'u'}, {"rv_m", 'v'}, {"ww_m", 'w'}, {"h_diabatic", 't'}};
  for (size_t i = 0; i < sizeof k3 / sizeof k3[0]; i++)
    if (strcmp(f, k3[i].name) == 0) {
      o_bc3(x, x->F[o_field_index(f)], k3[i].kind);
      return 1;
    }
  return -1;
}

/* Individual stage entry points for stage-level parity tests. */
int o_run_stage(OCtx *x, const char *name, int rk_step, int iteration) {
  o_set_rk_stage(x, rk_step);
  x->iteration = iteration;
#define IS(s) (strcmp(name, s) == 0)
  if (IS("rk_step_prep")) o_rk_step_prep(x);
  else if (IS("first_rk_step_part1")) o_first_rk_step_part1(x);
  else if (IS("first_rk_step_part2")) o_first_rk_step_part2(x);
  else if (IS("compute_diff_metrics")) o_compute_diff_metrics(x);
  else if (IS("cal_deform_and_div")) o_cal_deform_and_div(x);
  else if (IS("smag2d_km")) o_smag2d_km(x);
  else if (IS("calculate_n2")) o_calc_n2(x);
  else if (IS("calculate_km_kh")) { o_calc_n2(x); o_smag2d_km(x); }
  else if (IS("horizontal_diffusion_2")) o_horizontal_diffusion_2(x);
  else if (IS("rk_tendency")) o_rk_tendency(x);
  else if (IS("advect_u")) o_advect_u(x, FLD(u_2), FLD(ru_tend), FLD(ru), FLD(rv), FLD(ww));
  else if (IS("advect_v")) o_advect_v(x, FLD(v_2), FLD(rv_tend), FLD(ru), FLD(rv), FLD(ww));
  else if (IS("advect_w")) o_advect_w(x, FLD(w_2), FLD(rw_tend), FLD(ru), FLD(rv), FLD(ww));
  else if (IS("advect_scalar")) o_advect_scalar(x, FLD(t_2), FLD(t_tend), FLD(ru), FLD(rv), FLD(ww));
  else if (IS("advect_scalar_pd"))
    o_advect_scalar_pd(x, SPC(moist, 0), SPC(moist_old, 0), FLD(advect_tend), FLD(ru_m), FLD(rv_m), FLD(ww_m), FLD(muts), FLD(mub), FLD(mu_1), x->dt_rk);
  else if (IS("rhs_ph")) o_rhs_ph(x);
  else if (IS("horizontal_pressure_gradient")) o_horizontal_pressure_gradient(x);
  else if (IS("pg_buoy_w")) o_pg_buoy_w(x);
  else if (IS("dry_boundary_tendency")) o_dry_boundary_tendency(x);
  else if (IS("small_step_prep")) o_small_step_prep(x);
  else if (IS("calc_p_rho")) o_calc_p_rho(x, iteration);
  else if (IS("calc_coef_w")) o_calc_coef_w(x);
  else if (IS("advance_uv")) o_advance_uv(x);
  else if (IS("advance_mu_t")) o_advance_mu_t(x);
  else if (IS("advance_w")) o_advance_w(x);
  else if (IS("sumflux")) o_sumflux(x);
  else if (IS("acoustic_substep")) o_acoustic_substep(x, iteration);
  else if (IS("small_step_finish")) o_small_step_finish(x, rk_step == 3);
  else if (IS("calc_mu_uv_1")) o_calc_mu_uv_1(x);
  /* the remaining stages solve_em calls by name (solve_em.py:470-790) */
  else if (IS("initialize_moist_old")) o_initialize_moist_old(x);
  else if (IS("rk_phys_bc_dry_1")) o_rk_phys_bc_dry_1(x);
  else if (IS("relax_bdy_dry")) o_relax_bdy_dry(x);
  else if (IS("rk_addtend_dry")) o_rk_addtend_dry(x);
  else if (IS("spec_bdy_dry")) o_spec_bdy_dry(x);
  else if (IS("spec_bdyupdate:u")) o_spec_bdyupdate(x, FLD(u_2), FLD(ru_tend), x->dts_rk, 'u', 0);
  else if (IS("spec_bdyupdate:v")) o_spec_bdyupdate(x, FLD(v_2), FLD(rv_tend), x->dts_rk, 'v', 0);
  else if (IS("spec_bdyupdate:t")) o_spec_bdyupdate(x, FLD(t_2), FLD(t_tend), x->dts_rk, 't', 0);
  else if (IS("spec_bdyupdate:mu_2")) o_spec_bdyupdate(x, FLD(mu_2), FLD(mu_tend), x->dts_rk, 'm', 1);
  else if (IS("spec_bdyupdate:muts")) o_spec_bdyupdate(x, FLD(muts), FLD(mu_tend), x->dts_rk, 'm', 1);
  else if (IS("spec_bdyupdate_ph")) o_spec_bdyupdate_ph(x, x->dts_rk);
  else if (IS("zero_grad_bdy")) o_zero_grad_bdy_w(x, FLD(w_2));
  else if (IS("rk_update_scalar_pd:moist_old")) {
    for (int s = 0; s < x->c.nmoist; s++) o_rk_update_scalar_pd(x, SPC(moist_old, s), SPC(moist_tend, s));
  }
  else if (IS("rk_update_scalar_pd:scalar_old")) {
    for (int s = 0; s < x->c.nscalar; s++) o_rk_update_scalar_pd(x, SPC(scalar_old, s), SPC(scalar_tend, s));
  }
  else if (IS("moist_scalar_advance")) o_moist_scalar_advance(x);
  else if (IS("scalar_advance")) o_scalar_advance(x);
  else if (IS("calc_p_rho_phi:final")) o_calc_p_rho_phi(x);
  else if (IS("rk_phys_bc_dry_2")) o_rk_phys_bc_dry_2(x);
  else if (IS("set_phys_bc_dry_2")) o_set_phys_bc_dry_2(x);
  else if (IS("spec_bdy_final:u")) o_spec_bdy_final_of(x, 'u');
  else if (IS("spec_bdy_final:v")) o_spec_bdy_final_of(x, 'v');
  else if (IS("spec_bdy_final:t")) o_spec_bdy_final_of(x, 't');
  else if (IS("spec_bdy_final:ph")) o_spec_bdy_final_of(x, 'h');
  else if (IS("spec_bdy_final:mu")) o_spec_bdy_final_of(x, 'm');
  else if (IS("spec_bdy_final:moist") || IS("spec_bdy_final:moist_qv")) o_spec_bdy_final_of(x, 'q');
  else if (IS("set_w_surface")) o_set_w_surface(x, FLD(w_2), FLD(u_2), FLD(v_2));
  else if (IS("calc_p_rho_phi")) o_calc_p_rho_phi(x);
  else if (IS("rk_stage")) o_rk_stage(x, rk_step);
  else if (IS("step")) o_step(x);
  else return o_named_bc(x, name) == 1 ? 0 : -1;
#undef IS
  return 0;
}
