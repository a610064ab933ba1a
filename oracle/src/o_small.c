/*
 * Oracle: acoustic (small) step (test infrastructure only).
 * Follows src/wrf_jax/dyn_em/module_small_step_em.py and the stage adapters
 * in src/wrf_jax/runtime.py:12127-13530 (argument mapping).
 */
#include "o_core.h"

void o_calc_mu_uv(OCtx *x, const real *mu, const real *mub, real *muu, real *muv);

/* small_step_prep (module_small_step_em.py:272-717). */
void o_small_step_prep(OCtx *x) {
  int rk = x->rk_step;
  real *u_1 = FLD(u_1), *u_2 = FLD(u_2), *v_1 = FLD(v_1), *v_2 = FLD(v_2), *w_1 = FLD(w_1), *w_2 = FLD(w_2);
  real *t_1 = FLD(t_1), *t_2 = FLD(t_2), *ph_1 = FLD(ph_1), *ph_2 = FLD(ph_2);
  real *mub = FLD(mub), *mu_1 = FLD(mu_1), *mu_2 = FLD(mu_2), *muu = FLD(muu), *muus = FLD(muus), *muv = FLD(muv), *muvs = FLD(muvs);
  real *mut = FLD(mut), *muts = FLD(muts), *mudf = FLD(mudf);
  real *c1h = FLD(c1h), *c2h = FLD(c2h), *c1f = FLD(c1f), *c2f = FLD(c2f);
  real *u_save = FLD(u_save), *v_save = FLD(v_save), *w_save = FLD(w_save), *t_save = FLD(t_save), *ph_save = FLD(ph_save), *mu_save = FLD(mu_save);
  real *ww = FLD(ww), *ww_1 = FLD(ww_1), *c2a = FLD(c2a), *pb = FLD(pb), *p = FLD(p), *alt = FLD(alt);
  real *msfuy = FLD(msfuy), *msfvx_inv = FLD(msfvx_inv), *msfty = FLD(msfty);
  real cpovcv = (real)x->c.cpovcv;
  int ide = IDE, jde = JDE, kde = KDE;
  if (rk == 1) {
    for (int j = 1; j <= jde - 1; j++)
      for (int i = 1; i <= ide - 1; i++) {
        A2(mu_1, i, j) = A2(mu_2, i, j);
        A2(mudf, i, j) = 0;
      }
    for (int j = 1; j <= jde - 1; j++)
      for (int k = 1; k <= kde - 1; k++)
        for (int i = 1; i <= ide; i++) A3(u_1, i, k, j) = A3(u_2, i, k, j);
    for (int j = 1; j <= jde; j++)
      for (int k = 1; k <= kde - 1; k++)
        for (int i = 1; i <= ide - 1; i++) A3(v_1, i, k, j) = A3(v_2, i, k, j);
    for (int j = 1; j <= jde - 1; j++)
      for (int k = 1; k <= kde; k++)
        for (int i = 1; i <= ide - 1; i++) {
          if (k <= kde - 1) A3(t_1, i, k, j) = A3(t_2, i, k, j);
          A3(w_1, i, k, j) = A3(w_2, i, k, j);
          A3(ph_1, i, k, j) = A3(ph_2, i, k, j);
        }
    for (int j = 1; j <= jde - 1; j++)
      for (int i = 1; i <= ide - 1; i++) A2(muts, i, j) = A2(mub, i, j) + A2(mu_2, i, j);
    for (int j = 1; j <= jde - 1; j++)
      for (int i = 1; i <= ide; i++) A2(muus, i, j) = A2(muu, i, j);
    for (int j = 1; j <= jde; j++)
      for (int i = 1; i <= ide - 1; i++) A2(muvs, i, j) = A2(muv, i, j);
    for (int j = 1; j <= jde - 1; j++)
      for (int i = 1; i <= ide - 1; i++) {
        A2(mu_save, i, j) = A2(mu_2, i, j);
        A2(mu_2, i, j) = 0;
      }
  } else {
    for (int j = 1; j <= jde - 1; j++)
      for (int i = 1; i <= ide - 1; i++) A2(muts, i, j) = A2(mub, i, j) + A2(mu_1, i, j);
    for (int j = 1; j <= jde - 1; j++)
      for (int i = 1; i <= ide; i++) {
        int iw = (i == 1) ? 1 : i - 1;
        real s = A2(mub, i, j) + A2(mu_1, i, j);
        s = s + A2(mub, iw, j);
        s = s + A2(mu_1, iw, j);
        A2(muus, i, j) = (real)0.5 * s;
      }
    for (int j = 1; j <= jde; j++)
      for (int i = 1; i <= ide - 1; i++) {
        int js = (j == 1) ? 1 : j - 1;
        real s = A2(mub, i, j) + A2(mu_1, i, j);
        s = s + A2(mub, i, js);
        s = s + A2(mu_1, i, js);
        A2(muvs, i, j) = (real)0.5 * s;
      }
    for (int j = 1; j <= jde - 1; j++)
      for (int i = 1; i <= ide - 1; i++) {
        real m2 = A2(mu_2, i, j);
        A2(mu_save, i, j) = m2;
        A2(mu_2, i, j) = A2(mu_1, i, j) - m2;
      }
  }
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 1; k <= kde - 1; k++)
      for (int i = 1; i <= ide - 1; i++)
        A3(c2a, i, k, j) = (cpovcv * (A3(pb, i, k, j) + A3(p, i, k, j))) / A3(alt, i, k, j);
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 1; k <= kde - 1; k++)
      for (int i = 1; i <= ide; i++) {
        real old = A3(u_2, i, k, j);
        A3(u_save, i, k, j) = old;
        A3(u_2, i, k, j) = (rk == 1) ? (real)0 :
          ((A1(c1h, k) * A2(muus, i, j) + A1(c2h, k)) * A3(u_1, i, k, j) - (A1(c1h, k) * A2(muu, i, j) + A1(c2h, k)) * old) / A2(msfuy, i, j);
      }
  for (int j = 1; j <= jde; j++)
    for (int k = 1; k <= kde - 1; k++)
      for (int i = 1; i <= ide - 1; i++) {
        real old = A3(v_2, i, k, j);
        A3(v_save, i, k, j) = old;
        A3(v_2, i, k, j) = (rk == 1) ? (real)0 :
          ((A1(c1h, k) * A2(muvs, i, j) + A1(c2h, k)) * A3(v_1, i, k, j) - (A1(c1h, k) * A2(muv, i, j) + A1(c2h, k)) * old) * A2(msfvx_inv, i, j);
      }
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 1; k <= kde - 1; k++)
      for (int i = 1; i <= ide - 1; i++) {
        real old = A3(t_2, i, k, j);
        A3(t_save, i, k, j) = old;
        A3(t_2, i, k, j) = (rk == 1) ? (real)0 :
          (A1(c1h, k) * A2(muts, i, j) + A1(c2h, k)) * A3(t_1, i, k, j) - (A1(c1h, k) * A2(mut, i, j) + A1(c2h, k)) * old;
      }
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 1; k <= kde; k++)
      for (int i = 1; i <= ide - 1; i++) {
        real old = A3(w_2, i, k, j);
        A3(w_save, i, k, j) = old;
        A3(w_2, i, k, j) = (rk == 1) ? (real)0 :
          ((A1(c1f, k) * A2(muts, i, j) + A1(c2f, k)) * A3(w_1, i, k, j) - (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k)) * old) / A2(msfty, i, j);
        real pold = A3(ph_2, i, k, j);
        A3(ph_save, i, k, j) = pold;
        A3(ph_2, i, k, j) = A3(ph_1, i, k, j) - pold;
        A3(ww_1, i, k, j) = A3(ww, i, k, j);
      }
}

/* calc_p_rho, non-hydrostatic (module_small_step_em.py:720-1034):
 * al from the geopotential, p from the linearised EOS, divergence damping
 * with the previous pressure pm1 for step > 0.
 * Arguments per runtime.py:12257: ph=ph_2, t_1=t_save, mu=mu_2, mut=muts. */
void o_calc_p_rho(OCtx *x, int step) {
  real *al = FLD(al), *p = FLD(p), *ph = FLD(ph_2), *alt = FLD(alt), *t_2 = FLD(t_2), *t_1 = FLD(t_save);
  real *c2a = FLD(c2a), *pm1 = FLD(pm1), *mu = FLD(mu_2), *mut = FLD(muts), *c1h = FLD(c1h), *c2h = FLD(c2h), *rdnw = FLD(rdnw);
  real t0 = (real)x->c.t0, smdiv = (real)x->c.smdiv;
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real mass = A1(c1h, k) * A2(mut, i, j) + A1(c2h, k);
        real dry = A1(c1h, k) * A2(mu, i, j);
        real bracket = A3(alt, i, k, j) * dry + A1(rdnw, k) * (A3(ph, i, k + 1, j) - A3(ph, i, k, j));
        real alv = ((real)-1 / mass) * bracket;
        real tdel = A3(t_2, i, k, j) - dry * A3(t_1, i, k, j);
        real frac = (A3(alt, i, k, j) * tdel) / (mass * (t0 + A3(t_1, i, k, j)));
        real pr = A3(c2a, i, k, j) * (frac - alv);
        A3(al, i, k, j) = alv;
        if (step == 0) {
          A3(p, i, k, j) = pr;
          A3(pm1, i, k, j) = pr;
        } else {
          A3(p, i, k, j) = pr + smdiv * (pr - A3(pm1, i, k, j));
          A3(pm1, i, k, j) = pr;
        }
      }
}

/* calc_coef_w (module_small_step_em.py:1359-1586): LU factors (a, alpha,
 * gamma) of the implicit w-phi tridiagonal system. */
void o_calc_coef_w(OCtx *x) {
  real *a = FLD(a), *alpha = FLD(alpha), *gamma = FLD(gamma), *mut = FLD(mut), *cqw = FLD(cqw), *c2a = FLD(c2a);
  real *c1h = FLD(c1h), *c2h = FLD(c2h), *c1f = FLD(c1f), *c2f = FLD(c2f), *rdn = FLD(rdn), *rdnw = FLD(rdnw);
  real dts = (real)x->dts_rk, g = (real)x->c.g, epssm = (real)x->c.epssm;
  real lid = x->c.top_lid ? (real)0 : (real)1;
  real cof = (real)0.5 * dts * g * ((real)1 + epssm);
  cof = cof * cof;
  int kde = KDE;
  for (int j = 1; j <= JDE - 1; j++)
    for (int i = 1; i <= IDE - 1; i++) {
      real m = A2(mut, i, j);
      A3(a, i, 2, j) = 0;
      A3(a, i, kde, j) = ((real)-2 * cof * (A1(rdnw, kde - 1) * A1(rdnw, kde - 1)) * A3(c2a, i, kde - 1, j) * lid) /
                         ((A1(c1h, kde - 1) * m + A1(c2h, kde - 1)) * (A1(c1f, kde - 1) * m + A1(c2f, kde - 1)));
      A3(gamma, i, 1, j) = 0;
      for (int k = 3; k <= kde - 1; k++)
        A3(a, i, k, j) = (-A3(cqw, i, k, j) * cof * A1(rdn, k) * A1(rdnw, k - 1) * A3(c2a, i, k - 1, j)) /
                         ((A1(c1h, k - 1) * m + A1(c2h, k - 1)) * (A1(c1f, k - 1) * m + A1(c2f, k - 1)));
      real gprev = 0;
      for (int k = 2; k <= kde - 1; k++) {
        real mh = A1(c1h, k) * m + A1(c2h, k), mf = A1(c1f, k) * m + A1(c2f, k), mhb = A1(c1h, k - 1) * m + A1(c2h, k - 1);
        real mfa = A1(c1f, k + 1) * m + A1(c2f, k + 1);
        real bu = (A1(rdnw, k) * A3(c2a, i, k, j)) / (mh * mf);
        real bl = (A1(rdnw, k - 1) * A3(c2a, i, k - 1, j)) / (mhb * mf);
        real b = (real)1 + A3(cqw, i, k, j) * cof * A1(rdn, k) * (bu + bl);
        real c = (-A3(cqw, i, k, j) * cof * A1(rdn, k) * A1(rdnw, k) * A3(c2a, i, k, j)) / (mh * mfa);
        real al_ = (real)1 / (b - A3(a, i, k, j) * gprev);
        real ga = c * al_;
        A3(alpha, i, k, j) = al_;
        A3(gamma, i, k, j) = ga;
        gprev = ga;
      }
      real btop = (real)1 + ((real)2 * cof * (A1(rdnw, kde - 1) * A1(rdnw, kde - 1)) * A3(c2a, i, kde - 1, j)) /
                                ((A1(c1h, kde - 1) * m + A1(c2h, kde - 1)) * (A1(c1f, kde) * m + A1(c2f, kde)));
      A3(alpha, i, kde, j) = (real)1 / (btop - A3(a, i, kde, j) * gprev);
      A3(gamma, i, kde, j) = 0;
    }
}

/* Interface pressure used by the non-hydrostatic correction of advance_uv
 * (module_small_step_em.py:1627-1645): p summed over the two cells of the
 * face, extrapolated at k=1, zero at the top. */
static real uv_iface(OCtx *x, const real *p, int i, int j, int di, int dj, int level) {
  real *fnm = FLD(fnm), *fnp = FLD(fnp);
  real cf1 = (real)x->c.cf1, cf2 = (real)x->c.cf2, cf3 = (real)x->c.cf3;
#define PP(kk) (A3(p, i, kk, j) + A3(p, i - di, kk, j - dj))
  if (level == 1) return (real)0.5 * (cf1 * PP(1) + cf2 * PP(2) + cf3 * PP(3));
  if (level >= 2 && level <= KDE - 1) return (real)0.5 * (A1(fnm, level) * PP(level) + A1(fnp, level) * PP(level - 1));
#undef PP
  return 0;
}

/* advance_uv (module_small_step_em.py:1589-1987), specified boundaries:
 * u on i in [2, ide-1], j in [2, jde-2]; v on i in [2, ide-2], j in [2, jde-1].
 * Arguments per runtime.py:12856: mu=mu_2, muu/muv from rk_step_prep. */
void o_advance_uv(OCtx *x) {
  real *u = FLD(u_2), *v = FLD(v_2), *rut = FLD(ru_tend), *rvt = FLD(rv_tend), *p = FLD(p), *pb = FLD(pb), *ph = FLD(ph_2);
  real *php = FLD(php), *alt = FLD(alt), *al = FLD(al), *mu = FLD(mu_2), *muu = FLD(muu), *muv = FLD(muv);
  real *cqu = FLD(cqu), *cqv = FLD(cqv), *mudf = FLD(mudf), *c1h = FLD(c1h), *c2h = FLD(c2h), *rdnw = FLD(rdnw);
  real *msfux = FLD(msfux), *msfuy = FLD(msfuy), *msfvx = FLD(msfvx), *msfvx_inv = FLD(msfvx_inv), *msfvy = FLD(msfvy);
  real dts = (real)x->dts_rk, emdiv = (real)x->c.emdiv;
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  int sz = x->c.spec_zone;
  for (int axis = 0; axis < 2; axis++) {
    real *vel = axis == 0 ? u : v, *tend = axis == 0 ? rut : rvt, *mass = axis == 0 ? muu : muv, *cq = axis == 0 ? cqu : cqv;
    real spacing = axis == 0 ? rdx : rdy;
    int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
    int i0 = 1 + sz, i1 = axis == 0 ? IDE - sz : IDE - sz - 1;
    int j0 = 1 + sz, j1 = axis == 0 ? JDE - sz - 1 : JDE - sz;
    real grid_length = (real)1 / spacing;
    for (int j = j0; j <= j1; j++)
      for (int k = 1; k <= KDE - 1; k++)
        for (int i = i0; i <= i1; i++) {
          real adv = A3(vel, i, k, j) + dts * A3(tend, i, k, j);
          real damping = (real)-1 * emdiv * grid_length * (A2(mudf, i, j) - A2(mudf, i - di, j - dj));
          damping = axis == 0 ? damping / A2(msfuy, i, j) : damping * A2(msfvx_inv, i, j);
          real phi = (A3(ph, i, k + 1, j) - A3(ph, i - di, k + 1, j - dj)) + (A3(ph, i, k, j) - A3(ph, i - di, k, j - dj));
          real pres = (A3(alt, i, k, j) + A3(alt, i - di, k, j - dj)) * (A3(p, i, k, j) - A3(p, i - di, k, j - dj));
          real base = (A3(al, i, k, j) + A3(al, i - di, k, j - dj)) * (A3(pb, i, k, j) - A3(pb, i - di, k, j - dj));
          real term = (phi + pres) + base;
          real ratio = axis == 0 ? A2(msfux, i, j) / A2(msfuy, i, j) : A2(msfvy, i, j) / A2(msfvx, i, j);
          real m = A1(c1h, k) * A2(mass, i, j) + A1(c2h, k);
          real grad = (ratio * (real)0.5 * spacing * m) * term;
          real vert = A1(rdnw, k) * (uv_iface(x, p, i, j, di, dj, k + 1) - uv_iface(x, p, i, j, di, dj, k));
          real muv_ = (real)0.5 * (A1(c1h, k) * A2(mu, i - di, j - dj) + A1(c1h, k) * A2(mu, i, j));
          grad = grad + (ratio * spacing * (A3(php, i, k, j) - A3(php, i - di, k, j - dj))) * (vert - muv_);
          A3(vel, i, k, j) = (adv - dts * A3(cq, i, k, j) * grad) + A1(c1h, k) * damping;
        }
  }
}

/* advance_mu_t (module_small_step_em.py:2024-2403) on i in [2, ide-2],
 * j in [2, jde-2].  Arguments per runtime.py:13012: u=u_2, u_1=u_save,
 * v=v_2, v_1=v_save, mu=mu_2, t=t_2, t_1=t_save, t_ave=t_2ave, ft=t_tend. */
void o_advance_mu_t(OCtx *x) {
  real *ww = FLD(ww), *ww_1 = FLD(ww_1), *u = FLD(u_2), *u_1 = FLD(u_save), *v = FLD(v_2), *v_1 = FLD(v_save);
  real *mu = FLD(mu_2), *mut = FLD(mut), *muave = FLD(muave), *muts = FLD(muts), *muu = FLD(muu), *muv = FLD(muv), *mudf = FLD(mudf);
  real *c1h = FLD(c1h), *c2h = FLD(c2h), *t = FLD(t_2), *t_1 = FLD(t_save), *t_ave = FLD(t_2ave), *ft = FLD(t_tend), *mu_tend = FLD(mu_tend);
  real *dnw = FLD(dnw), *fnm = FLD(fnm), *fnp = FLD(fnp), *rdnw = FLD(rdnw);
  real *msfuy = FLD(msfuy), *msfvx_inv = FLD(msfvx_inv), *msftx = FLD(msftx), *msfty = FLD(msfty);
  real dts = (real)x->dts_rk, epssm = (real)x->c.epssm;
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  int k_end = KDE - 1;
  real dvdxi[512], wdtn[512];
  for (int j = 2; j <= JDE - 2; j++)
    for (int i = 2; i <= IDE - 2; i++) {
      /* column mass tendency: the FP32 level terms summed in FP64 and
       * rounded once, as the GPU kernels do */
      double dmdt_acc = 0;
      for (int k = 1; k <= k_end; k++) {
        real c1 = A1(c1h, k), c2 = A1(c2h, k);
        real vn = A3(v, i, k, j + 1) + ((c1 * A2(muv, i, j + 1) + c2) * A3(v_1, i, k, j + 1)) * A2(msfvx_inv, i, j + 1);
        real vs = A3(v, i, k, j) + ((c1 * A2(muv, i, j) + c2) * A3(v_1, i, k, j)) * A2(msfvx_inv, i, j);
        real ue = A3(u, i + 1, k, j) + ((c1 * A2(muu, i + 1, j) + c2) * A3(u_1, i + 1, k, j)) / A2(msfuy, i + 1, j);
        real uw = A3(u, i, k, j) + ((c1 * A2(muu, i, j) + c2) * A3(u_1, i, k, j)) / A2(msfuy, i, j);
        real hf = rdy * (vn - vs) + rdx * (ue - uw);
        dvdxi[k] = (A2(msftx, i, j) * A2(msfty, i, j)) * hf;
        dmdt_acc += (double)(A1(dnw, k) * dvdxi[k]);
      }
      real dmdt = (real)dmdt_acc;
      real old = A2(mu, i, j);
      real tend = dmdt + A2(mu_tend, i, j);
      real nmu = old + dts * tend;
      A2(mu, i, j) = nmu;
      A2(mudf, i, j) = tend;
      A2(muts, i, j) = A2(mut, i, j) + nmu;
      A2(muave, i, j) = (real)0.5 * (((real)1 + epssm) * nmu + ((real)1 - epssm) * old);
      /* omega recurrence, then remove the stage-start omega */
      real prev = A3(ww, i, 1, j);
      for (int k = 2; k <= k_end; k++) {
        real integ = A1(c1h, k - 1) * dmdt + dvdxi[k - 1];
        integ = integ + A1(c1h, k - 1) * A2(mu_tend, i, j);
        real cur = prev - (A1(dnw, k - 1) * integ) / A2(msfty, i, j);
        A3(ww, i, k, j) = cur;
        prev = cur;
      }
      for (int k = 1; k <= k_end; k++) A3(ww, i, k, j) = A3(ww, i, k, j) - A3(ww_1, i, k, j);
      /* theta */
      wdtn[1] = 0;
      wdtn[k_end + 1] = 0;
      for (int k = 2; k <= k_end; k++)
        wdtn[k] = A3(ww, i, k, j) * (A1(fnm, k) * A3(t_1, i, k, j) + A1(fnp, k) * A3(t_1, i, k - 1, j));
      for (int k = 1; k <= k_end; k++) {
        A3(t_ave, i, k, j) = A3(t, i, k, j);
        real ta = A3(t, i, k, j) + (A2(msfty, i, j) * dts) * A3(ft, i, k, j);
        real vnorth = A3(v, i, k, j + 1) * (A3(t_1, i, k, j + 1) + A3(t_1, i, k, j));
        real vsouth = A3(v, i, k, j) * (A3(t_1, i, k, j) + A3(t_1, i, k, j - 1));
        real vft = ((real)0.5 * rdy) * (vnorth - vsouth);
        real ueast = A3(u, i + 1, k, j) * (A3(t_1, i + 1, k, j) + A3(t_1, i, k, j));
        real uwest = A3(u, i, k, j) * (A3(t_1, i, k, j) + A3(t_1, i - 1, k, j));
        real uft = ((real)0.5 * rdx) * (ueast - uwest);
        real horiz = A2(msftx, i, j) * (vft + uft);
        real vert = A1(rdnw, k) * (wdtn[k + 1] - wdtn[k]);
        A3(t, i, k, j) = ta - (dts * A2(msfty, i, j)) * (horiz + vert);
      }
    }
}

/* advance_w (module_small_step_em.py:2406-3131) on i in [2, ide-2],
 * j in [2, jde-2]: explicit terms, then the implicit w-phi tridiagonal
 * solve with the calc_coef_w factors.  Arguments per runtime.py:13108:
 * u=u_2, v=v_2, t_1=t_save, ph=ph_2, ph_1=ph_save. */
void o_advance_w(OCtx *x) {
  real *w = FLD(w_2), *rw_tend = FLD(rw_tend), *ww = FLD(ww), *u = FLD(u_2), *v = FLD(v_2);
  real *mut = FLD(mut), *muave = FLD(muave), *muts = FLD(muts);
  real *c1h = FLD(c1h), *c2h = FLD(c2h), *c1f = FLD(c1f), *c2f = FLD(c2f);
  real *t_2ave = FLD(t_2ave), *t_2 = FLD(t_2), *t_1 = FLD(t_save), *ph = FLD(ph_2), *ph_1 = FLD(ph_save), *phb = FLD(phb), *ph_tend = FLD(ph_tend);
  real *ht = FLD(ht), *c2a = FLD(c2a), *cqw = FLD(cqw), *alt = FLD(alt), *a = FLD(a), *alpha = FLD(alpha), *gamma = FLD(gamma);
  real *fnm = FLD(fnm), *fnp = FLD(fnp), *rdnw = FLD(rdnw), *rdn = FLD(rdn), *msftx = FLD(msftx), *msfty = FLD(msfty);
  real dts = (real)x->dts_rk, g = (real)x->c.g, epssm = (real)x->c.epssm, t0 = (real)x->c.t0;
  real cf1 = (real)x->c.cf1, cf2 = (real)x->c.cf2, cf3 = (real)x->c.cf3;
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  real ope = (real)1 + epssm, ome = (real)1 - epssm;
  int k_end = KDE - 1, kde = KDE;
  real rhs[512], wdwn[512], wl[512];
  for (int j = 2; j <= JDE - 2; j++)
    for (int i = 2; i <= IDE - 2; i++) {
      real m = A2(mut, i, j), mts = A2(muts, i, j), mav = A2(muave, i, j), msft_inv = (real)1 / A2(msfty, i, j);
      for (int k = 1; k <= k_end; k++) {
        real ta = (real)0.5 * (ope * A3(t_2, i, k, j) + ome * A3(t_2ave, i, k, j));
        ta = (ta + (A1(c1h, k) * mav) * t0) / ((A1(c1h, k) * mts + A1(c2h, k)) * (t0 + A3(t_1, i, k, j)));
        A3(t_2ave, i, k, j) = ta;
      }
      for (int k = 2; k <= k_end + 1; k++) {
        rhs[k] = dts * (A3(ph_tend, i, k, j) + (((real)0.5 * g) * ome) * A3(w, i, k, j));
        wdwn[k] = (((real)0.5 * (A3(ww, i, k, j) + A3(ww, i, k - 1, j))) * A1(rdnw, k - 1)) *
                  (A3(ph_1, i, k, j) - A3(ph_1, i, k - 1, j) + A3(phb, i, k, j) - A3(phb, i, k - 1, j));
      }
      for (int k = 2; k <= k_end; k++) rhs[k] = rhs[k] - dts * (A1(fnm, k) * wdwn[k + 1] + A1(fnp, k) * wdwn[k]);
      for (int k = 2; k <= k_end + 1; k++)
        rhs[k] = A3(ph, i, k, j) + (A2(msfty, i, j) * rhs[k]) / (A1(c1f, k) * m + A1(c2f, k));
      if (x->c.top_lid) rhs[k_end + 1] = 0;
      rhs[1] = 0;
      /* surface w from terrain slope */
      real lv = (A2(ht, i, j + 1) - A2(ht, i, j)) * (cf1 * A3(v, i, 1, j + 1) + cf2 * A3(v, i, 2, j + 1) + cf3 * A3(v, i, 3, j + 1)) +
                (A2(ht, i, j) - A2(ht, i, j - 1)) * (cf1 * A3(v, i, 1, j) + cf2 * A3(v, i, 2, j) + cf3 * A3(v, i, 3, j));
      real lu = (A2(ht, i + 1, j) - A2(ht, i, j)) * (cf1 * A3(u, i + 1, 1, j) + cf2 * A3(u, i + 1, 2, j) + cf3 * A3(u, i + 1, 3, j)) +
                (A2(ht, i, j) - A2(ht, i - 1, j)) * (cf1 * A3(u, i, 1, j) + cf2 * A3(u, i, 2, j) + cf3 * A3(u, i, 3, j));
      wl[1] = ((A2(msfty, i, j) * (real)0.5) * rdy) * lv + ((A2(msftx, i, j) * (real)0.5) * rdx) * lu;
      for (int k = 2; k <= k_end; k++) {
        real pu = ((A3(c2a, i, k, j) * A1(rdnw, k)) / (A1(c1h, k) * m + A1(c2h, k))) *
                  (ope * (rhs[k + 1] - rhs[k]) + ome * (A3(ph, i, k + 1, j) - A3(ph, i, k, j)));
        real pl = ((A3(c2a, i, k - 1, j) * A1(rdnw, k - 1)) / (A1(c1h, k - 1) * m + A1(c2h, k - 1))) *
                  (ope * (rhs[k] - rhs[k - 1]) + ome * (A3(ph, i, k, j) - A3(ph, i, k - 1, j)));
        real ppart = ((((real)0.5 * dts) * g) * A1(rdn, k)) * (pu - pl);
        real tu = (A3(c2a, i, k, j) * A3(alt, i, k, j)) * A3(t_2ave, i, k, j);
        real tl = (A3(c2a, i, k - 1, j) * A3(alt, i, k - 1, j)) * A3(t_2ave, i, k - 1, j);
        real bpart = A1(rdn, k) * (tu - tl) - A1(c1f, k) * mav;
        wl[k] = ((A3(w, i, k, j) + dts * A3(rw_tend, i, k, j)) + (msft_inv * A3(cqw, i, k, j)) * ppart) + ((dts * g) * msft_inv) * bpart;
      }
      {
        int k = k_end + 1;
        real fac = (((real)-0.5 * dts) * g) / (A1(c1h, k - 1) * m + A1(c2h, k - 1));
        fac = fac * (A1(rdnw, k - 1) * A1(rdnw, k - 1));
        fac = fac * (real)2;
        fac = fac * A3(c2a, i, k - 1, j);
        real ptop = fac * (ope * (rhs[k] - rhs[k - 1]) + ome * (A3(ph, i, k, j) - A3(ph, i, k - 1, j)));
        real btop = (-dts * g) * ((((real)2 * A1(rdnw, k - 1)) * A3(c2a, i, k - 1, j)) * A3(alt, i, k - 1, j) * A3(t_2ave, i, k - 1, j) + A1(c1f, k) * mav);
        wl[k] = A3(w, i, k, j) + dts * A3(rw_tend, i, k, j) + ((real)1 / A2(msfty, i, j)) * (ptop + btop);
        if (x->c.top_lid) wl[k] = 0;
      }
      for (int k = 2; k <= k_end + 1; k++) wl[k] = (wl[k] - A3(a, i, k, j) * wl[k - 1]) * A3(alpha, i, k, j);
      for (int k = k_end; k >= 2; k--) wl[k] = wl[k] - A3(gamma, i, k, j) * wl[k + 1];
      for (int k = 1; k <= kde; k++) A3(w, i, k, j) = wl[k];
      for (int k = 2; k <= k_end + 1; k++)
        A3(ph, i, k, j) = rhs[k] + ((((A2(msfty, i, j) * (real)0.5) * dts) * g) * ope) * wl[k] / (A1(c1f, k) * mts + A1(c2f, k));
    }
}

/* sumflux (module_small_step_em.py:3134-3310): time-averaged mass fluxes.
 * Arguments per runtime.py:13264: ru=u_2, rv=v_2, u_lin=u_save,
 * v_lin=v_save, ww_lin=ww_1. */
void o_sumflux(OCtx *x) {
  real *ru_m = FLD(ru_m), *rv_m = FLD(rv_m), *ww_m = FLD(ww_m), *u = FLD(u_2), *v = FLD(v_2), *ww = FLD(ww);
  real *u_lin = FLD(u_save), *v_lin = FLD(v_save), *ww_lin = FLD(ww_1), *muu = FLD(muu), *muv = FLD(muv);
  real *c1h = FLD(c1h), *c2h = FLD(c2h), *msfuy = FLD(msfuy), *msfvx_inv = FLD(msfvx_inv);
  int first = (x->iteration == 1), fin = (x->iteration == x->nsmall);
  real ns = (real)x->nsmall;
  if (first) {
    memset(ru_m, 0, sizeof(real) * x->n3);
    memset(rv_m, 0, sizeof(real) * x->n3);
    memset(ww_m, 0, sizeof(real) * x->n3);
  }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE; i++) {
        real acc = A3(ru_m, i, k, j) + A3(u, i, k, j);
        if (fin) acc = acc / ns + ((A1(c1h, k) * A2(muu, i, j) + A1(c2h, k)) * A3(u_lin, i, k, j)) / A2(msfuy, i, j);
        A3(ru_m, i, k, j) = acc;
      }
  for (int j = 1; j <= JDE; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real acc = A3(rv_m, i, k, j) + A3(v, i, k, j);
        if (fin) acc = acc / ns + ((A1(c1h, k) * A2(muv, i, j) + A1(c2h, k)) * A3(v_lin, i, k, j)) * A2(msfvx_inv, i, j);
        A3(rv_m, i, k, j) = acc;
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real acc = A3(ww_m, i, k, j) + A3(ww, i, k, j);
        if (fin) acc = acc / ns + A3(ww_lin, i, k, j);
        A3(ww_m, i, k, j) = acc;
      }
}

/* calc_mu_uv_1 (module_big_step_utilities_em.py:269-316) applied to muts. */
void o_calc_mu_uv_1(OCtx *x) {
  real *zero = (real *)calloc(x->n2, sizeof(real));
  o_calc_mu_uv(x, FLD(muts), zero, FLD(muus), FLD(muvs));
  free(zero);
}

/* small_step_finish (module_small_step_em.py:1119-1332): recouple the
 * acoustic perturbations onto the stage-start state. */
void o_small_step_finish(OCtx *x, int final_step) {
  real *u_2 = FLD(u_2), *v_2 = FLD(v_2), *w_2 = FLD(w_2), *t_2 = FLD(t_2), *ph_2 = FLD(ph_2), *ww = FLD(ww), *ww1 = FLD(ww_1);
  real *mu_2 = FLD(mu_2), *mut = FLD(mut), *muts = FLD(muts), *muu = FLD(muu), *muus = FLD(muus), *muv = FLD(muv), *muvs = FLD(muvs);
  real *u_save = FLD(u_save), *v_save = FLD(v_save), *w_save = FLD(w_save), *t_save = FLD(t_save), *ph_save = FLD(ph_save), *mu_save = FLD(mu_save);
  real *c1h = FLD(c1h), *c2h = FLD(c2h), *c1f = FLD(c1f), *c2f = FLD(c2f), *msfuy = FLD(msfuy), *msfvx = FLD(msfvx), *msfty = FLD(msfty);
  real *hd = FLD(h_diabatic);
  real dts = (real)x->dts_rk, nsm = (real)x->nsmall;
  for (int j = 1; j <= JDE; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real old = A1(c1h, k) * A2(muv, i, j) + A1(c2h, k), nw = A1(c1h, k) * A2(muvs, i, j) + A1(c2h, k);
        A3(v_2, i, k, j) = (A2(msfvx, i, j) * A3(v_2, i, k, j) + A3(v_save, i, k, j) * old) / nw;
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE; i++) {
        real old = A1(c1h, k) * A2(muu, i, j) + A1(c2h, k), nw = A1(c1h, k) * A2(muus, i, j) + A1(c2h, k);
        A3(u_2, i, k, j) = (A2(msfuy, i, j) * A3(u_2, i, k, j) + A3(u_save, i, k, j) * old) / nw;
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real old = A1(c1f, k) * A2(mut, i, j) + A1(c2f, k), nw = A1(c1f, k) * A2(muts, i, j) + A1(c2f, k);
        A3(w_2, i, k, j) = (A2(msfty, i, j) * A3(w_2, i, k, j) + A3(w_save, i, k, j) * old) / nw;
        A3(ph_2, i, k, j) = A3(ph_2, i, k, j) + A3(ph_save, i, k, j);
        A3(ww, i, k, j) = A3(ww, i, k, j) + A3(ww1, i, k, j);
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real old = A1(c1h, k) * A2(mut, i, j) + A1(c2h, k), nw = A1(c1h, k) * A2(muts, i, j) + A1(c2h, k);
        real tt = A3(t_2, i, k, j);
        if (final_step) tt = tt - ((dts * nsm) * old) * A3(hd, i, k, j);
        A3(t_2, i, k, j) = (tt + A3(t_save, i, k, j) * old) / nw;
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int i = 1; i <= IDE - 1; i++) A2(mu_2, i, j) = A2(mu_2, i, j) + A2(mu_save, i, j);
}
