/*
 * Oracle: big-step (RK stage) utilities (test infrastructure only).
 * Follows src/wrf_jax/dyn_em/module_big_step_utilities_em.py and
 * src/wrf_jax/dyn_em/module_em.py.
 */
#include "o_core.h"

/* calc_mu_uv (module_big_step_utilities_em.py:174-266), non-periodic:
 * face averages of the full column mass with edge reuse. */
void o_calc_mu_uv(OCtx *x, const real *mu, const real *mub, real *muu, real *muv) {
  int ide = IDE, jde = JDE;
  for (int j = 1; j <= jde - 1; j++)
    for (int i = 1; i <= ide; i++) {
      int ia = (i == ide) ? ide - 1 : i;
      int ib = (i == 1) ? 1 : ((i == ide) ? ide - 1 : i - 1);
      A2(muu, i, j) = (real)0.5 * (A2(mu, ia, j) + A2(mu, ib, j) + A2(mub, ia, j) + A2(mub, ib, j));
    }
  for (int j = 1; j <= jde; j++)
    for (int i = 1; i <= ide - 1; i++) {
      int ja = (j == jde) ? jde - 1 : j;
      int jb = (j == 1) ? 1 : ((j == jde) ? jde - 1 : j - 1);
      A2(muv, i, j) = (real)0.5 * (A2(mu, i, ja) + A2(mu, i, jb) + A2(mub, i, ja) + A2(mub, i, jb));
    }
}

/* couple_momentum (module_big_step_utilities_em.py:1123-1211). */
void o_couple_momentum(OCtx *x) {
  real *ru = FLD(ru), *rv = FLD(rv), *rw = FLD(rw), *u = FLD(u_2), *v = FLD(v_2), *w = FLD(w_2);
  real *muu = FLD(muu), *muv = FLD(muv), *mut = FLD(mut);
  real *msfuy = FLD(msfuy), *msfvx_inv = FLD(msfvx_inv), *msfty = FLD(msfty);
  real *c1h = FLD(c1h), *c2h = FLD(c2h), *c1f = FLD(c1f), *c2f = FLD(c2f);
  memset(ru, 0, sizeof(real) * x->n3);
  memset(rv, 0, sizeof(real) * x->n3);
  memset(rw, 0, sizeof(real) * x->n3);
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE; i++)
        A3(ru, i, k, j) = A3(u, i, k, j) * (A1(c1h, k) * A2(muu, i, j) + A1(c2h, k)) / A2(msfuy, i, j);
  for (int j = 1; j <= JDE; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++)
        A3(rv, i, k, j) = A3(v, i, k, j) * (A1(c1h, k) * A2(muv, i, j) + A1(c2h, k)) * A2(msfvx_inv, i, j);
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE; k++)
      for (int i = 1; i <= IDE - 1; i++)
        A3(rw, i, k, j) = A3(w, i, k, j) * (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k)) / A2(msfty, i, j);
}

/* calc_ww_cp (module_big_step_utilities_em.py:655-803): omega from the
 * bottom-up column integral of the coupled horizontal mass divergence. */
void o_calc_ww_cp(OCtx *x, real *ww) {
  real *u = FLD(u_2), *v = FLD(v_2), *mup = FLD(mu_2), *mub = FLD(mub);
  real *c1h = FLD(c1h), *c2h = FLD(c2h), *msftx = FLD(msftx), *msfuy = FLD(msfuy), *msfvx_inv = FLD(msfvx_inv);
  real *dnw = FLD(dnw);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  int ktf = KDE - 1;
  real divv[512];
  memset(ww, 0, sizeof(real) * x->n3);
  for (int j = 1; j <= JDE - 1; j++)
    for (int i = 1; i <= IDE - 1; i++) {
      int im = (i - 1 < 1) ? i : i - 1, jm = (j - 1 < 1) ? j : j - 1;
      real m0 = A2(mup, i, j) + A2(mub, i, j);
      real muu_i = (real)0.5 * (A2(mup, i, j) + A2(mub, i, j) + A2(mup, im, j) + A2(mub, im, j));
      real muu_ip = (real)0.5 * (A2(mup, i + 1, j) + A2(mub, i + 1, j) + A2(mup, i, j) + A2(mub, i, j));
      real muv_j = (real)0.5 * (A2(mup, i, j) + A2(mub, i, j) + A2(mup, i, jm) + A2(mub, i, jm));
      real muv_jp = (real)0.5 * (A2(mup, i, j + 1) + A2(mub, i, j + 1) + A2(mup, i, j) + A2(mub, i, j));
      (void)m0;
      real dmdt = 0;
      for (int k = 1; k <= ktf; k++) {
        real c1 = A1(c1h, k), c2 = A1(c2h, k);
        real xf = (c1 * muu_ip + c2) * A3(u, i + 1, k, j) / A2(msfuy, i + 1, j) - (c1 * muu_i + c2) * A3(u, i, k, j) / A2(msfuy, i, j);
        real yf = (c1 * muv_jp + c2) * A3(v, i, k, j + 1) * A2(msfvx_inv, i, j + 1) - (c1 * muv_j + c2) * A3(v, i, k, j) * A2(msfvx_inv, i, j);
        divv[k] = A2(msftx, i, j) * A1(dnw, k) * (rdx * xf + rdy * yf);
        dmdt += divv[k];
      }
      real prev = 0;
      for (int k = 2; k <= ktf; k++) {
        real cur = prev - A1(dnw, k - 1) * A1(c1h, k - 1) * dmdt - divv[k - 1];
        A3(ww, i, k, j) = cur;
        prev = cur;
      }
    }
}

/* calc_cq (module_big_step_utilities_em.py:2873-2991): moisture coupling
 * coefficients; summed over all moist species. */
void o_calc_cq(OCtx *x) {
  real *cqu = FLD(cqu), *cqv = FLD(cqv), *cqw = FLD(cqw);
  int nm = x->c.nmoist, ktf = KDE - 1;
  memset(cqu, 0, sizeof(real) * x->n3);
  memset(cqv, 0, sizeof(real) * x->n3);
  memset(cqw, 0, sizeof(real) * x->n3);
  /* neighbour index -1 wraps to the last stored i (JAX negative indexing). */
  for (int j = 1; j <= JDE; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= IDE; i++) {
        int im = (i == 1) ? IDE : i - 1;
        int jm = (j == 1) ? JDE : j - 1;
        real su = 0, sv = 0, sw = 0;
        for (int s = 0; s < nm; s++) {
          const real *q = SPC(moist, s);
          su = su + A3(q, i, k, j) + A3(q, im, k, j);
          sv = sv + A3(q, i, k, j) + A3(q, i, k, jm);
          if (k > 1) sw = sw + A3(q, i, k, j) + A3(q, i, k - 1, j);
        }
        if (j < JDE) A3(cqu, i, k, j) = (real)1 / ((real)1 + (real)0.5 * su);
        if (i < IDE) A3(cqv, i, k, j) = (real)1 / ((real)1 + (real)0.5 * sv);
        if (i < IDE && j < JDE && k > 1) A3(cqw, i, k, j) = (real)0.5 * sw;
      }
}

/* rk_step_prep (module_em.py:231-459 with runtime.py:7237-7360). */
void o_rk_step_prep(OCtx *x) {
  real *mu = FLD(mu_2), *mub = FLD(mub), *mut = FLD(mut);
  for (int j = 1; j <= JDE; j++)
    for (int i = 1; i <= IDE; i++) A2(mut, i, j) = A2(mub, i, j) + A2(mu, i, j);
  o_calc_mu_uv(x, mu, mub, FLD(muu), FLD(muv));
  o_couple_momentum(x);
  o_calc_ww_cp(x, FLD(ww));
  o_calc_cq(x);
  real *alt = FLD(alt), *al = FLD(al), *alb = FLD(alb), *php = FLD(php), *ph = FLD(ph_2), *phb = FLD(phb);
  memset(alt, 0, sizeof(real) * x->n3);
  memset(php, 0, sizeof(real) * x->n3);
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        A3(alt, i, k, j) = A3(al, i, k, j) + A3(alb, i, k, j);
        A3(php, i, k, j) = (real)0.5 * (A3(phb, i, k, j) + A3(phb, i, k + 1, j) + A3(ph, i, k, j) + A3(ph, i, k + 1, j));
      }
}

/* rhs_ph (module_big_step_utilities_em.py:2386-2794), specified,
 * h_sca_adv_order 5 (6th/4th/2nd-order geopotential advection),
 * phi_adv_z = 1, non-hydrostatic. */
void o_rhs_ph(OCtx *x) {
  real *out = FLD(ph_tend), *u = FLD(u_2), *v = FLD(v_2), *ww = FLD(ww), *ph = FLD(ph_2), *phb = FLD(phb), *w = FLD(w_2);
  real *mut = FLD(mut), *muu = FLD(muu), *muv = FLD(muv), *c1f = FLD(c1f), *c2f = FLD(c2f);
  real *fnm = FLD(fnm), *fnp = FLD(fnp), *rdnw = FLD(rdnw), *msfux = FLD(msfux), *msfvy = FLD(msfvy), *msfty = FLD(msfty);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy), g = (real)x->c.g;
  real cfn = (real)x->c.cfn, cfn1 = (real)x->c.cfn1;
  int ide = IDE, jde = JDE, kte = KDE, kde = KDE;
  /* vertical advection term */
  for (int j = 1; j <= jde - 1; j++)
    for (int i = 1; i <= ide - 1; i++) {
      real wdwn[512];
      for (int k = 2; k <= kte; k++)
        wdwn[k] = (real)0.5 * (A3(ww, i, k, j) + A3(ww, i, k - 1, j)) * A1(rdnw, k - 1) *
                  (A3(ph, i, k, j) - A3(ph, i, k - 1, j) + A3(phb, i, k, j) - A3(phb, i, k - 1, j));
      for (int k = 2; k <= kte - 1; k++) A3(out, i, k, j) += -(A1(fnm, k) * wdwn[k + 1] + A1(fnp, k) * wdwn[k]);
    }
  /* non-hydrostatic w*g term */
  for (int j = 1; j <= jde - 1; j++)
    for (int i = 1; i <= ide - 1; i++) {
      A3(out, i, kde, j) = 0;
      for (int k = 2; k <= kte; k++)
        A3(out, i, k, j) += (A1(c1f, k) * A2(mut, i, j) + A1(c2f, k)) * g * A3(w, i, k, j) / A2(msfty, i, j);
    }
  /* horizontal advection helpers */
#define YSECOND(jj)                                                                                                  \
  do {                                                                                                               \
    int j = (jj);                                                                                                    \
    for (int k = 2; k <= kte - 1; k++)                                                                               \
      for (int i = 1; i <= ide - 1; i++) {                                                                           \
        real term = (A1(c1f, k) * A2(muv, i, j + 1) + A1(c2f, k)) * (A3(v, i, k, j + 1) + A3(v, i, k - 1, j + 1)) *   \
                        A2(msfvy, i, j + 1) * (A3(phb, i, k, j + 1) - A3(phb, i, k, j) + A3(ph, i, k, j + 1) - A3(ph, i, k, j)) + \
                    (A1(c1f, k) * A2(muv, i, j) + A1(c2f, k)) * (A3(v, i, k, j) + A3(v, i, k - 1, j)) * A2(msfvy, i, j) *     \
                        (A3(phb, i, k, j) - A3(phb, i, k, j - 1) + A3(ph, i, k, j) - A3(ph, i, k, j - 1));                 \
        A3(out, i, k, j) += -((real)0.25 * rdy / A2(msfty, i, j)) * term;                                              \
      }                                                                                                              \
    {                                                                                                                \
      int k = kte;                                                                                                   \
      for (int i = 1; i <= ide - 1; i++) {                                                                           \
        real term = (A1(c1f, k) * A2(muv, i, j + 1) + A1(c2f, k)) * (cfn * A3(v, i, k - 1, j + 1) + cfn1 * A3(v, i, k - 2, j + 1)) * \
                        A2(msfvy, i, j + 1) * (A3(phb, i, k, j + 1) - A3(phb, i, k, j) + A3(ph, i, k, j + 1) - A3(ph, i, k, j)) + \
                    (A1(c1f, k) * A2(muv, i, j) + A1(c2f, k)) * (cfn * A3(v, i, k - 1, j) + cfn1 * A3(v, i, k - 2, j)) *        \
                        A2(msfvy, i, j) * (A3(phb, i, k, j) - A3(phb, i, k, j - 1) + A3(ph, i, k, j) - A3(ph, i, k, j - 1)); \
        A3(out, i, k, j) += -((real)0.5 * rdy / A2(msfty, i, j)) * term;                                               \
      }                                                                                                              \
    }                                                                                                                \
  } while (0)
#define XSECOND(ii)                                                                                                  \
  do {                                                                                                               \
    int i = (ii);                                                                                                    \
    for (int j = 1; j <= jde - 1; j++) {                                                                             \
      for (int k = 2; k <= kte - 1; k++) {                                                                           \
        real term = (A1(c1f, k) * A2(muu, i + 1, j) + A1(c2f, k)) * (A3(u, i + 1, k, j) + A3(u, i + 1, k - 1, j)) *   \
                        A2(msfux, i + 1, j) * (A3(phb, i + 1, k, j) - A3(phb, i, k, j) + A3(ph, i + 1, k, j) - A3(ph, i, k, j)) + \
                    (A1(c1f, k) * A2(muu, i, j) + A1(c2f, k)) * (A3(u, i, k, j) + A3(u, i, k - 1, j)) * A2(msfux, i, j) *     \
                        (A3(phb, i, k, j) - A3(phb, i - 1, k, j) + A3(ph, i, k, j) - A3(ph, i - 1, k, j));                 \
        A3(out, i, k, j) += -((real)0.25 * rdx / A2(msfty, i, j)) * term;                                              \
      }                                                                                                              \
      int k = kte;                                                                                                   \
      real term = (A1(c1f, k) * A2(muu, i + 1, j) + A1(c2f, k)) * (cfn * A3(u, i + 1, k - 1, j) + cfn1 * A3(u, i + 1, k - 2, j)) * \
                      A2(msfux, i + 1, j) * (A3(phb, i + 1, k, j) - A3(phb, i, k, j) + A3(ph, i + 1, k, j) - A3(ph, i, k, j)) +   \
                  (A1(c1f, k) * A2(muu, i, j) + A1(c2f, k)) * (cfn * A3(u, i, k - 1, j) + cfn1 * A3(u, i, k - 2, j)) *          \
                      A2(msfux, i, j) * (A3(phb, i, k, j) - A3(phb, i - 1, k, j) + A3(ph, i, k, j) - A3(ph, i - 1, k, j));   \
      A3(out, i, k, j) += -((real)0.5 * rdx / A2(msfty, i, j)) * term;                                                 \
    }                                                                                                                \
  } while (0)
#define YHIGH(jj, ca, cb, cc, scale)                                                                                 \
  do {                                                                                                               \
    int j = (jj), mo = ((cc) != 0) ? 3 : 2;                                                                          \
    for (int k = 2; k <= kte; k++)                                                                                   \
      for (int i = 1; i <= ide - 1; i++) {                                                                           \
        real grad = (ca) * (A3(ph, i, k, j + 1) - A3(ph, i, k, j - 1)) + (cb) * (A3(ph, i, k, j + 2) - A3(ph, i, k, j - 2)) + \
                    (cc) * (A3(ph, i, k, j + mo) - A3(ph, i, k, j - mo)) + (ca) * (A3(phb, i, k, j + 1) - A3(phb, i, k, j - 1)) + \
                    (cb) * (A3(phb, i, k, j + 2) - A3(phb, i, k, j - 2)) + (cc) * (A3(phb, i, k, j + mo) - A3(phb, i, k, j - mo)); \
        real vel, fac;                                                                                               \
        if (k < kte) {                                                                                               \
          vel = (A1(c1f, k) * A2(muv, i, j + 1) + A1(c2f, k)) * (A3(v, i, k, j + 1) + A3(v, i, k - 1, j + 1)) * A2(msfvy, i, j + 1) + \
                (A1(c1f, k) * A2(muv, i, j) + A1(c2f, k)) * (A3(v, i, k, j) + A3(v, i, k - 1, j)) * A2(msfvy, i, j);         \
          fac = (real)0.25;                                                                                          \
        } else {                                                                                                     \
          vel = (A1(c1f, k) * A2(muv, i, j + 1) + A1(c2f, k)) * (cfn * A3(v, i, k - 1, j + 1) + cfn1 * A3(v, i, k - 2, j + 1)) * A2(msfvy, i, j + 1) + \
                (A1(c1f, k) * A2(muv, i, j) + A1(c2f, k)) * (cfn * A3(v, i, k - 1, j) + cfn1 * A3(v, i, k - 2, j)) * A2(msfvy, i, j); \
          fac = (real)0.5;                                                                                           \
        }                                                                                                            \
        A3(out, i, k, j) += -(fac * rdy / A2(msfty, i, j)) * vel * (real)(scale) * grad;                              \
      }                                                                                                              \
  } while (0)
#define XHIGH(ii, ca, cb, cc, scale)                                                                                 \
  do {                                                                                                               \
    int i = (ii), mo = ((cc) != 0) ? 3 : 2;                                                                          \
    for (int j = 1; j <= jde - 1; j++)                                                                               \
      for (int k = 2; k <= kte; k++) {                                                                               \
        real grad = (ca) * (A3(ph, i + 1, k, j) - A3(ph, i - 1, k, j)) + (cb) * (A3(ph, i + 2, k, j) - A3(ph, i - 2, k, j)) + \
                    (cc) * (A3(ph, i + mo, k, j) - A3(ph, i - mo, k, j)) + (ca) * (A3(phb, i + 1, k, j) - A3(phb, i - 1, k, j)) + \
                    (cb) * (A3(phb, i + 2, k, j) - A3(phb, i - 2, k, j)) + (cc) * (A3(phb, i + mo, k, j) - A3(phb, i - mo, k, j)); \
        real vel, fac;                                                                                               \
        if (k < kte) {                                                                                               \
          vel = (A1(c1f, k) * A2(muu, i + 1, j) + A1(c2f, k)) * (A3(u, i + 1, k, j) + A3(u, i + 1, k - 1, j)) * A2(msfux, i + 1, j) + \
                (A1(c1f, k) * A2(muu, i, j) + A1(c2f, k)) * (A3(u, i, k, j) + A3(u, i, k - 1, j)) * A2(msfux, i, j);         \
          fac = (real)0.25;                                                                                          \
        } else {                                                                                                     \
          vel = (A1(c1f, k) * A2(muu, i + 1, j) + A1(c2f, k)) * (cfn * A3(u, i + 1, k - 1, j) + cfn1 * A3(u, i + 1, k - 2, j)) * A2(msfux, i + 1, j) + \
                (A1(c1f, k) * A2(muu, i, j) + A1(c2f, k)) * (cfn * A3(u, i, k - 1, j) + cfn1 * A3(u, i, k - 2, j)) * A2(msfux, i, j); \
          fac = (real)0.5;                                                                                           \
        }                                                                                                            \
        A3(out, i, k, j) += -(fac * rdx / A2(msfty, i, j)) * vel * (real)(scale) * grad;                              \
      }                                                                                                              \
  } while (0)
  /* advective_order <= 6 branch, y direction */
  for (int jj = 4; jj <= jde - 4; jj++) YHIGH(jj, (real)45, (real)-9, (real)1, 1.0 / 60.0);
  YHIGH(3, (real)8, (real)-1, (real)0, 1.0 / 12.0);
  YHIGH(jde - 3, (real)8, (real)-1, (real)0, 1.0 / 12.0);
  YSECOND(2);
  YSECOND(jde - 2);
  /* x direction: the 4th-order columns ids+2 / ide-3 are applied only for
   * open boundaries in the reference (rhs_ph:2718-2721), so the specified
   * case leaves them without horizontal advection. */
  for (int ii = 4; ii <= ide - 4; ii++) XHIGH(ii, (real)45, (real)-9, (real)1, 1.0 / 60.0);
  XSECOND(2);
  XSECOND(ide - 2);
#undef YSECOND
#undef XSECOND
#undef YHIGH
#undef XHIGH
}

/* horizontal_pressure_gradient (module_big_step_utilities_em.py:1452-1671),
 * specified, non-hydrostatic, no lid. */
void o_horizontal_pressure_gradient(OCtx *x) {
  real *ru_t = FLD(ru_tend), *rv_t = FLD(rv_tend), *ph = FLD(ph_2), *alt = FLD(alt), *p = FLD(p), *pb = FLD(pb), *al = FLD(al);
  real *php = FLD(php), *cqu = FLD(cqu), *cqv = FLD(cqv), *muu = FLD(muu), *muv = FLD(muv), *mu = FLD(mu_2);
  real *c1h = FLD(c1h), *c2h = FLD(c2h), *fnm = FLD(fnm), *fnp = FLD(fnp), *rdnw = FLD(rdnw);
  real *msfux = FLD(msfux), *msfuy = FLD(msfuy), *msfvx = FLD(msfvx), *msfvy = FLD(msfvy);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  real cf1 = (real)x->c.cf1, cf2 = (real)x->c.cf2, cf3 = (real)x->c.cf3;
  int ktf = KDE - 1;
  real dpn[512];
  /* v: rows [2, jde-1], i [1, ide-1] */
  for (int j = 2; j <= JDE - 1; j++)
    for (int i = 1; i <= IDE - 1; i++) {
      dpn[1] = (real)0.5 * (cf1 * (A3(p, i, 1, j - 1) + A3(p, i, 1, j)) + cf2 * (A3(p, i, 2, j - 1) + A3(p, i, 2, j)) + cf3 * (A3(p, i, 3, j - 1) + A3(p, i, 3, j)));
      dpn[KDE] = 0;
      for (int k = 2; k <= ktf; k++)
        dpn[k] = (real)0.5 * (A1(fnm, k) * (A3(p, i, k, j - 1) + A3(p, i, k, j)) + A1(fnp, k) * (A3(p, i, k - 1, j - 1) + A3(p, i, k - 1, j)));
      for (int k = 1; k <= ktf; k++) {
        real r = A2(msfvy, i, j) / A2(msfvx, i, j);
        real dpy = r * (real)0.5 * rdy * (A1(c1h, k) * A2(muv, i, j) + A1(c2h, k)) *
                   ((A3(ph, i, k + 1, j) - A3(ph, i, k + 1, j - 1) + A3(ph, i, k, j) - A3(ph, i, k, j - 1)) +
                    (A3(alt, i, k, j) + A3(alt, i, k, j - 1)) * (A3(p, i, k, j) - A3(p, i, k, j - 1)) +
                    (A3(al, i, k, j) + A3(al, i, k, j - 1)) * (A3(pb, i, k, j) - A3(pb, i, k, j - 1)));
        dpy = dpy + r * rdy * (A3(php, i, k, j) - A3(php, i, k, j - 1)) *
                        (A1(rdnw, k) * (dpn[k + 1] - dpn[k]) - (real)0.5 * (A1(c1h, k) * A2(mu, i, j - 1) + A1(c1h, k) * A2(mu, i, j)));
        A3(rv_t, i, k, j) = A3(rv_t, i, k, j) - A3(cqv, i, k, j) * dpy;
      }
    }
  /* u: columns [2, ide-1], j [1, jde-1] */
  for (int j = 1; j <= JDE - 1; j++)
    for (int i = 2; i <= IDE - 1; i++) {
      dpn[1] = (real)0.5 * (cf1 * (A3(p, i - 1, 1, j) + A3(p, i, 1, j)) + cf2 * (A3(p, i - 1, 2, j) + A3(p, i, 2, j)) + cf3 * (A3(p, i - 1, 3, j) + A3(p, i, 3, j)));
      dpn[KDE] = 0;
      for (int k = 2; k <= ktf; k++)
        dpn[k] = (real)0.5 * (A1(fnm, k) * (A3(p, i - 1, k, j) + A3(p, i, k, j)) + A1(fnp, k) * (A3(p, i - 1, k - 1, j) + A3(p, i, k - 1, j)));
      for (int k = 1; k <= ktf; k++) {
        real r = A2(msfux, i, j) / A2(msfuy, i, j);
        real dpx = r * (real)0.5 * rdx * (A1(c1h, k) * A2(muu, i, j) + A1(c2h, k)) *
                   ((A3(ph, i, k + 1, j) - A3(ph, i - 1, k + 1, j) + A3(ph, i, k, j) - A3(ph, i - 1, k, j)) +
                    (A3(alt, i, k, j) + A3(alt, i - 1, k, j)) * (A3(p, i, k, j) - A3(p, i - 1, k, j)) +
                    (A3(al, i, k, j) + A3(al, i - 1, k, j)) * (A3(pb, i, k, j) - A3(pb, i - 1, k, j)));
        dpx = dpx + r * rdx * (A3(php, i, k, j) - A3(php, i - 1, k, j)) *
                        (A1(rdnw, k) * (dpn[k + 1] - dpn[k]) - (real)0.5 * (A1(c1h, k) * A2(mu, i - 1, j) + A1(c1h, k) * A2(mu, i, j)));
        A3(ru_t, i, k, j) = A3(ru_t, i, k, j) - A3(cqu, i, k, j) * dpx;
      }
    }
}

/* pg_buoy_w (module_big_step_utilities_em.py:1674-1800): vertical pressure
 * gradient + buoyancy; converts cqw to 1/(1+qtot) on levels 2..kde-1. */
void o_pg_buoy_w(OCtx *x) {
  real *rw_t = FLD(rw_tend), *p = FLD(p), *cqw = FLD(cqw), *mu = FLD(mu_2), *mub = FLD(mub);
  real *c1f = FLD(c1f), *c2f = FLD(c2f), *rdnw = FLD(rdnw), *rdn = FLD(rdn), *msfty = FLD(msfty);
  real g = (real)x->c.g;
  int kde = KDE;
  for (int j = 1; j <= JDE - 1; j++)
    for (int i = 1; i <= IDE - 1; i++) {
      real cq = A3(cqw, i, kde - 1, j);
      real cq1 = (real)1 / ((real)1 + cq), cq2 = cq * cq1;
      A3(rw_t, i, kde, j) += ((real)1 / A2(msfty, i, j)) * g *
                             (cq1 * (real)2 * A1(rdnw, kde - 1) * (-A3(p, i, kde - 1, j)) - (A1(c1f, kde) * A2(mu, i, j)) -
                              cq2 * (A1(c1f, kde) * A2(mub, i, j) + A1(c2f, kde)));
      for (int k = 2; k <= kde - 1; k++) {
        cq = A3(cqw, i, k, j);
        cq1 = (real)1 / ((real)1 + cq);
        cq2 = cq * cq1;
        A3(rw_t, i, k, j) += ((real)1 / A2(msfty, i, j)) * g *
                             (cq1 * A1(rdn, k) * (A3(p, i, k, j) - A3(p, i, k - 1, j)) - (A1(c1f, k) * A2(mu, i, j)) -
                              cq2 * (A1(c1f, k) * A2(mub, i, j) + A1(c2f, k)));
        A3(cqw, i, k, j) = cq1;
      }
    }
}

/* coriolis (module_big_step_utilities_em.py:978-1120), specified. */
void o_coriolis(OCtx *x) {
  real *ru = FLD(ru), *rv = FLD(rv), *rw = FLD(rw), *rut = FLD(ru_tend), *rvt = FLD(rv_tend), *rwt = FLD(rw_tend);
  real *msftx = FLD(msftx), *msfty = FLD(msfty), *msfux = FLD(msfux), *msfuy = FLD(msfuy), *msfvx = FLD(msfvx), *msfvy = FLD(msfvy);
  real *f = FLD(f), *e = FLD(e), *sina = FLD(sina), *cosa = FLD(cosa), *fzm = FLD(fnm), *fzp = FLD(fnp);
  int ktf = KDE - 1;
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 2; i <= IDE - 1; i++) {
        real val = A3(rut, i, k, j);
        val = val + (A2(msfux, i, j) / A2(msfuy, i, j)) * (real)0.5 * (A2(f, i, j) + A2(f, i - 1, j)) * (real)0.25 *
                        (A3(rv, i - 1, k, j + 1) + A3(rv, i, k, j + 1) + A3(rv, i - 1, k, j) + A3(rv, i, k, j));
        val = val - (real)0.5 * (A2(e, i, j) + A2(e, i - 1, j)) * (real)0.5 * (A2(cosa, i, j) + A2(cosa, i - 1, j)) * (real)0.25 *
                        (A3(rw, i - 1, k + 1, j) + A3(rw, i - 1, k, j) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
        A3(rut, i, k, j) = val;
      }
  for (int j = 2; j <= JDE - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real val = A3(rvt, i, k, j);
        val = val - (A2(msfvy, i, j) / A2(msfvx, i, j)) * (real)0.5 * (A2(f, i, j) + A2(f, i, j - 1)) * (real)0.25 *
                        (A3(ru, i, k, j) + A3(ru, i + 1, k, j) + A3(ru, i, k, j - 1) + A3(ru, i + 1, k, j - 1));
        val = val + (A2(msfvy, i, j) / A2(msfvx, i, j)) * (real)0.5 * (A2(e, i, j) + A2(e, i, j - 1)) * (real)0.5 *
                        (A2(sina, i, j) + A2(sina, i, j - 1)) * (real)0.25 *
                        (A3(rw, i, k + 1, j - 1) + A3(rw, i, k, j - 1) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
        A3(rvt, i, k, j) = val;
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 2; k <= ktf; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        A3(rwt, i, k, j) = A3(rwt, i, k, j) +
                           A2(e, i, j) * (A2(cosa, i, j) * (real)0.5 * (A1(fzm, k) * (A3(ru, i, k, j) + A3(ru, i + 1, k, j)) + A1(fzp, k) * (A3(ru, i, k - 1, j) + A3(ru, i + 1, k - 1, j))) -
                                          (A2(msftx, i, j) / A2(msfty, i, j)) * A2(sina, i, j) * (real)0.5 *
                                              (A1(fzm, k) * (A3(rv, i, k, j) + A3(rv, i, k, j + 1)) + A1(fzp, k) * (A3(rv, i, k - 1, j) + A3(rv, i, k - 1, j + 1))));
      }
}

/* curvature (module_big_step_utilities_em.py:1214-1449), specified,
 * non-polar, map_proj != Cassini. */
void o_curvature(OCtx *x) {
  real *ru = FLD(ru), *rv = FLD(rv), *rw = FLD(rw), *u = FLD(u_2), *v = FLD(v_2);
  real *rut = FLD(ru_tend), *rvt = FLD(rv_tend), *rwt = FLD(rw_tend);
  real *msfux = FLD(msfux), *msfuy = FLD(msfuy), *msfvx = FLD(msfvx), *msfvy = FLD(msfvy), *msftx = FLD(msftx), *msfty = FLD(msfty);
  real *fzm = FLD(fnm), *fzp = FLD(fnp);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy), rer = (real)x->c.reradius;
  int ktf = KDE - 1;
  real *vxgm = (real *)calloc(x->n3, sizeof(real));
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= IDE - 1; i++)
        A3(vxgm, i, k, j) = (real)0.5 * (A3(u, i, k, j) + A3(u, i + 1, k, j)) * (A2(msfvx, i, j + 1) - A2(msfvx, i, j)) * rdy -
                            (real)0.5 * (A3(v, i, k, j) + A3(v, i, k, j + 1)) * (A2(msfuy, i + 1, j) - A2(msfuy, i, j)) * rdx;
  /* the its-1 copy lands on the unused i = ide slot and is then overwritten */
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= ktf; k++) A3(vxgm, IDE, k, j) = A3(vxgm, IDE - 1, k, j);
  for (int k = 1; k <= ktf; k++)
    for (int i = 1; i <= IDE; i++) A3(vxgm, i, k, JDE) = A3(vxgm, i, k, JDE - 1);
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 2; i <= IDE - 1; i++)
        A3(rut, i, k, j) = A3(rut, i, k, j) +
                           (real)0.5 * (A3(vxgm, i, k, j) + A3(vxgm, i - 1, k, j)) * (real)0.25 *
                               (A3(rv, i - 1, k, j + 1) + A3(rv, i, k, j + 1) + A3(rv, i - 1, k, j) + A3(rv, i, k, j)) -
                           A3(u, i, k, j) * rer * (real)0.25 * (A3(rw, i - 1, k + 1, j) + A3(rw, i - 1, k, j) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
  for (int j = 2; j <= JDE - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= IDE - 1; i++)
        A3(rvt, i, k, j) = A3(rvt, i, k, j) -
                           (real)0.5 * (A3(vxgm, i, k, j) + A3(vxgm, i, k, j - 1)) * (real)0.25 *
                               (A3(ru, i, k, j) + A3(ru, i + 1, k, j) + A3(ru, i, k, j - 1) + A3(ru, i + 1, k, j - 1)) -
                           (A2(msfvy, i, j) / A2(msfvx, i, j)) * A3(v, i, k, j) * rer * (real)0.25 *
                               (A3(rw, i, k + 1, j - 1) + A3(rw, i, k, j - 1) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 2; k <= ktf; k++)
      for (int i = 1; i <= IDE - 1; i++)
        A3(rwt, i, k, j) = A3(rwt, i, k, j) +
                           rer * ((real)0.5 * (A1(fzm, k) * (A3(ru, i, k, j) + A3(ru, i + 1, k, j)) + A1(fzp, k) * (A3(ru, i, k - 1, j) + A3(ru, i + 1, k - 1, j))) *
                                      (real)0.5 * (A1(fzm, k) * (A3(u, i, k, j) + A3(u, i + 1, k, j)) + A1(fzp, k) * (A3(u, i, k - 1, j) + A3(u, i + 1, k - 1, j))) +
                                  (A2(msftx, i, j) / A2(msfty, i, j)) * (real)0.5 *
                                      (A1(fzm, k) * (A3(rv, i, k, j) + A3(rv, i, k, j + 1)) + A1(fzp, k) * (A3(rv, i, k - 1, j) + A3(rv, i, k - 1, j + 1))) *
                                      (real)0.5 * (A1(fzm, k) * (A3(v, i, k, j) + A3(v, i, k, j + 1)) + A1(fzp, k) * (A3(v, i, k - 1, j) + A3(v, i, k - 1, j + 1))));
  free(vxgm);
}

/* _calc_p_rho_phi_nonhydro_opt2 (runtime.py:12369-12537; spec
 * module_big_step_utilities_em.py:364-578): hypsometric al and EOS p. */
void o_calc_p_rho_phi(OCtx *x) {
  real *al = FLD(al), *p = FLD(p), *alb = FLD(alb), *muts = FLD(muts), *ph = FLD(ph_2), *phb = FLD(phb), *pb = FLD(pb), *t = FLD(t_2);
  real *c3h = FLD(c3h), *c4h = FLD(c4h), *c3f = FLD(c3f), *c4f = FLD(c4f);
  const real *qv = SPC(moist, 0);
  real ptop = (real)x->c.ptop, p0 = (real)x->c.p0, t0 = (real)x->c.t0, r_d = (real)x->c.r_d;
  real cpovcv = (real)x->c.cpovcv, rvovrd = (real)x->c.rvovrd;
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real m = A2(muts, i, j);
        real pfu = A1(c3f, k + 1) * m + A1(c4f, k + 1) + ptop;
        real pfd = A1(c3f, k) * m + A1(c4f, k) + ptop;
        real phm = A1(c3h, k) * m + A1(c4h, k) + ptop;
        real alv = (A3(ph, i, k + 1, j) - A3(ph, i, k, j) + A3(phb, i, k + 1, j) - A3(phb, i, k, j)) / phm / (real)log(pfd / pfu) - A3(alb, i, k, j);
        A3(al, i, k, j) = alv;
        real qvf = (real)1 + rvovrd * A3(qv, i, k, j);
        real temp = (r_d * (t0 + A3(t, i, k, j)) * qvf) / (p0 * (alv + A3(alb, i, k, j)));
        A3(p, i, k, j) = (real)pow(temp, cpovcv) * p0 - A3(pb, i, k, j);
      }
}

/* rk_addtend_dry (module_em.py:959-1148). */
void o_rk_addtend_dry(OCtx *x) {
  int rk = x->rk_step;
  real *rut = FLD(ru_tend), *rvt = FLD(rv_tend), *rwt = FLD(rw_tend), *pht = FLD(ph_tend), *tt = FLD(t_tend), *mut_t = FLD(mu_tend);
  real *ruf = FLD(ru_tendf), *rvf = FLD(rv_tendf), *rwf = FLD(rw_tendf), *phf = FLD(ph_tendf), *tf = FLD(t_tendf), *muf = FLD(mu_tendf);
  real *us = FLD(u_save), *vs = FLD(v_save), *ws = FLD(w_save), *phs = FLD(ph_save), *ts = FLD(t_save);
  real *msfuy = FLD(msfuy), *msfvx = FLD(msfvx), *msfvx_inv = FLD(msfvx_inv), *msfty = FLD(msfty), *mut = FLD(mut);
  real *c1 = FLD(c1h), *c2 = FLD(c2h), *hd = FLD(h_diabatic);
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE; i++) {
        if (rk == 1) A3(ruf, i, k, j) = A3(ruf, i, k, j) + A3(us, i, k, j) * A2(msfuy, i, j);
        A3(rut, i, k, j) = A3(rut, i, k, j) + A3(ruf, i, k, j) / A2(msfuy, i, j);
      }
  for (int j = 1; j <= JDE; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        if (rk == 1) A3(rvf, i, k, j) = A3(rvf, i, k, j) + A3(vs, i, k, j) * A2(msfvx, i, j);
        A3(rvt, i, k, j) = A3(rvt, i, k, j) + A3(rvf, i, k, j) * A2(msfvx_inv, i, j);
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        if (rk == 1) A3(rwf, i, k, j) = A3(rwf, i, k, j) + A3(ws, i, k, j) * A2(msfty, i, j);
        A3(rwt, i, k, j) = A3(rwt, i, k, j) + A3(rwf, i, k, j) / A2(msfty, i, j);
        if (rk == 1) A3(phf, i, k, j) = A3(phf, i, k, j) + A3(phs, i, k, j);
        A3(pht, i, k, j) = A3(pht, i, k, j) + A3(phf, i, k, j) / A2(msfty, i, j);
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real mass = A1(c1, k) * A2(mut, i, j) + A1(c2, k);
        if (rk == 1) A3(tf, i, k, j) = A3(tf, i, k, j) + A3(ts, i, k, j);
        real heating = mass * A3(hd, i, k, j) / A2(msfty, i, j);
        A3(tt, i, k, j) = A3(tt, i, k, j) + A3(tf, i, k, j) / A2(msfty, i, j) + heating;
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int i = 1; i <= IDE - 1; i++) A2(mut_t, i, j) = A2(mut_t, i, j) + A2(muf, i, j);
}
