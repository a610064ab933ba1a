/*
 * Oracle: flux-form advection, 5th-order horizontal / 3rd-order vertical,
 * specified lateral boundaries (test infrastructure only).
 * Follows src/wrf_jax/dyn_em/module_advect_em.py (loop-form spec).
 */
#include "o_core.h"

/* _flux3 / _flux5 (module_advect_em.py:29-55); time_step > 0 so
 * sign(1, time_step) = 1. */
static inline real flux3(real qm2, real qm1, real q0, real qp1, real ua) {
  return ((real)7 * (q0 + qm1) - (qp1 + qm2)) / (real)12 +
         sgn1(ua) * ((qp1 - qm2) - (real)3 * (q0 - qm1)) / (real)12;
}
static inline real flux5(real qm3, real qm2, real qm1, real q0, real qp1, real qp2, real ua) {
  return ((real)37 * (q0 + qm1) - (real)8 * (qp1 + qm2) + (qp2 + qm3)) / (real)60 -
         sgn1(ua) * ((qp2 - qm3) - (real)5 * (qp1 - qm2) + (real)10 * (q0 - qm1)) / (real)60;
}
/* _flux3_pd / _flux5_pd (module_advect_em.py:58-82) */
static inline real flux3pd(real qm2, real qm1, real q0, real qp1, real ua) {
  real f4 = ((real)7 / (real)12) * (q0 + qm1) - ((real)1 / (real)12) * (qp1 + qm2);
  return f4 + sgn1(ua) * ((real)1 / (real)12) * ((qp1 - qm2) - (real)3 * (q0 - qm1));
}
static inline real flux5pd(real qm3, real qm2, real qm1, real q0, real qp1, real qp2, real ua) {
  real f6 = ((real)37 / (real)60) * (q0 + qm1) - ((real)2 / (real)15) * (qp1 + qm2) + ((real)1 / (real)60) * (qp2 + qm3);
  return f6 - sgn1(ua) * ((real)1 / (real)60) * ((qp2 - qm3) - (real)5 * (qp1 - qm2) + (real)10 * (q0 - qm1));
}
/* _flux_upwind (module_advect_em.py:85-87) */
static inline real flux_up(real qm1, real q0, real cr) {
  return (real)0.5 * rmin((real)1, cr + fabs(cr)) * qm1 + (real)0.5 * rmax((real)-1, cr - fabs(cr)) * q0;
}

/* advect_scalar (module_advect_em.py:90-369), h5/v3, specified boundaries.
 * tend -= div(flux); fluxes use mass fluxes ru, rv, rom directly. */
void o_advect_scalar(OCtx *x, const real *q, real *tend, const real *ru, const real *rv, const real *rom) {
  real *msftx = FLD(msftx), *fzm = FLD(fnm), *fzp = FLD(fnp), *rdzw = FLD(rdnw);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  int ktf = KDE - 1, jde = JDE, ide = IDE;
  real *fq = (real *)malloc(sizeof(real) * (size_t)(IDE + 2) * (size_t)(KDE + 2));
#define FQ(i, k) fq[(size_t)(k) * (IDE + 2) + (i)]
  /* y fluxes: j in [2, jde-1], tendency rows [2, jde-2] (degrade_ys/ye). */
  {
    int j_start = 2, j_end = jde - 2, jsf = 4, jef = jde - 3;
    real *prev = (real *)malloc(sizeof(real) * (size_t)(IDE + 2) * (size_t)(KDE + 2));
    for (int j = j_start; j <= j_end + 1; j++) {
      for (int k = 1; k <= ktf; k++)
        for (int i = 1; i <= ide - 1; i++) {
          real vel = A3(rv, i, k, j), val;
          if (j >= jsf && j <= jef)
            val = vel * flux5(A3(q, i, k, j - 3), A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), A3(q, i, k, j + 2), vel);
          else if (j == 2 || j == jde - 1)
            val = (real)0.5 * vel * (A3(q, i, k, j) + A3(q, i, k, j - 1));
          else
            val = vel * flux3(A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), vel);
          FQ(i, k) = val;
        }
      if (j > j_start)
        for (int k = 1; k <= ktf; k++)
          for (int i = 1; i <= ide - 1; i++)
            A3(tend, i, k, j - 1) -= A2(msftx, i, j - 1) * rdy * (FQ(i, k) - prev[(size_t)k * (IDE + 2) + i]);
      memcpy(prev, fq, sizeof(real) * (size_t)(IDE + 2) * (size_t)(KDE + 2));
    }
    free(prev);
  }
  /* x fluxes: faces [2, ide-1], tendency columns [2, ide-2]. */
  for (int j = 1; j <= jde - 1; j++) {
    for (int k = 1; k <= ktf; k++) {
      for (int i = 2; i <= ide - 1; i++) {
        real vel = A3(ru, i, k, j), val;
        if (i >= 4 && i <= ide - 3)
          val = vel * flux5(A3(q, i - 3, k, j), A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), A3(q, i + 2, k, j), vel);
        else if (i == 2 || i == ide - 1)
          val = (real)0.5 * vel * (A3(q, i, k, j) + A3(q, i - 1, k, j));
        else
          val = vel * flux3(A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), vel);
        FQ(i, k) = val;
      }
      for (int i = 2; i <= ide - 2; i++) A3(tend, i, k, j) -= A2(msftx, i, j) * rdx * (FQ(i + 1, k) - FQ(i, k));
    }
  }
  /* vertical: 3rd order interior, 2nd order at kts+1 and ktf, zero at 1/kde */
  for (int j = 1; j <= jde - 1; j++)
    for (int i = 1; i <= ide - 1; i++) {
      real vf[512];
      vf[1] = 0;
      vf[KDE] = 0;
      for (int k = 3; k <= ktf - 1; k++) {
        real vel = A3(rom, i, k, j);
        vf[k] = vel * flux3(A3(q, i, k - 2, j), A3(q, i, k - 1, j), A3(q, i, k, j), A3(q, i, k + 1, j), -vel);
      }
      vf[2] = A3(rom, i, 2, j) * (A1(fzm, 2) * A3(q, i, 2, j) + A1(fzp, 2) * A3(q, i, 1, j));
      vf[ktf] = A3(rom, i, ktf, j) * (A1(fzm, ktf) * A3(q, i, ktf, j) + A1(fzp, ktf) * A3(q, i, ktf - 1, j));
      for (int k = 1; k <= ktf; k++) A3(tend, i, k, j) -= A1(rdzw, k) * (vf[k + 1] - vf[k]);
    }
#undef FQ
  free(fq);
}

/* advect_u (module_advect_em.py:762-1034), specified boundaries:
 * u points i in [2, ide-1], rows j in [1, jde-1]. */
void o_advect_u(OCtx *x, const real *u, real *tend, const real *ru, const real *rv, const real *rom) {
  real *msfux = FLD(msfux), *fzm = FLD(fnm), *fzp = FLD(fnp), *rdzw = FLD(rdnw);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  int ktf = KDE - 1, jde = JDE, ide = IDE;
  size_t w = (size_t)(IDE + 2);
  real *cur = (real *)malloc(sizeof(real) * w * (KDE + 2));
  real *prev = (real *)malloc(sizeof(real) * w * (KDE + 2));
  /* y: j in [2, jde-1] fluxes, tendency rows [2, jde-2] */
  for (int j = 2; j <= jde - 1; j++) {
    for (int k = 1; k <= ktf; k++)
      for (int i = 2; i <= ide - 1; i++) {
        real val;
        if (j >= 4 && j <= jde - 3) {
          real vel = (real)0.5 * (A3(rv, i, k, j) + A3(rv, i - 1, k, j));
          val = vel * flux5(A3(u, i, k, j - 3), A3(u, i, k, j - 2), A3(u, i, k, j - 1), A3(u, i, k, j), A3(u, i, k, j + 1), A3(u, i, k, j + 2), vel);
        } else if (j == 2 || j == jde - 1) {
          val = (real)0.25 * (A3(rv, i, k, j) + A3(rv, i - 1, k, j)) * (A3(u, i, k, j) + A3(u, i, k, j - 1));
        } else {
          real vel = (real)0.5 * (A3(rv, i, k, j) + A3(rv, i - 1, k, j));
          val = vel * flux3(A3(u, i, k, j - 2), A3(u, i, k, j - 1), A3(u, i, k, j), A3(u, i, k, j + 1), vel);
        }
        cur[k * w + i] = val;
      }
    if (j > 2)
      for (int k = 1; k <= ktf; k++)
        for (int i = 2; i <= ide - 1; i++)
          A3(tend, i, k, j - 1) -= A2(msfux, i, j - 1) * rdy * (cur[k * w + i] - prev[k * w + i]);
    real *t = prev; prev = cur; cur = t;
  }
  /* x: faces [2, ide], tendency at u points [2, ide-1] */
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 1; k <= ktf; k++) {
      real *fqx = cur;
      for (int i = 4; i <= ide - 2; i++) {
        real vel = (real)0.5 * (A3(ru, i, k, j) + A3(ru, i - 1, k, j));
        fqx[i] = vel * flux5(A3(u, i - 3, k, j), A3(u, i - 2, k, j), A3(u, i - 1, k, j), A3(u, i, k, j), A3(u, i + 1, k, j), A3(u, i + 2, k, j), vel);
      }
      {
        int i = 2;
        real ub = A3(u, i - 1, k, j);
        if (A3(u, i, k, j) < (real)0) ub = A3(u, i, k, j);
        fqx[i] = (real)0.25 * (A3(ru, i, k, j) + A3(ru, i - 1, k, j)) * (A3(u, i, k, j) + ub);
        i = 3;
        real vel = (real)0.5 * (A3(ru, i, k, j) + A3(ru, i - 1, k, j));
        fqx[i] = vel * flux3(A3(u, i - 2, k, j), A3(u, i - 1, k, j), A3(u, i, k, j), A3(u, i + 1, k, j), vel);
      }
      {
        int i = ide;
        real ub = A3(u, i, k, j);
        if (A3(u, i - 1, k, j) > (real)0) ub = A3(u, i - 1, k, j);
        fqx[i] = (real)0.25 * (A3(ru, i, k, j) + A3(ru, i - 1, k, j)) * (A3(u, i - 1, k, j) + ub);
        i = ide - 1;
        real vel = (real)0.5 * (A3(ru, i, k, j) + A3(ru, i - 1, k, j));
        fqx[i] = vel * flux3(A3(u, i - 2, k, j), A3(u, i - 1, k, j), A3(u, i, k, j), A3(u, i + 1, k, j), vel);
      }
      for (int i = 2; i <= ide - 1; i++) A3(tend, i, k, j) -= A2(msfux, i, j) * rdx * (fqx[i + 1] - fqx[i]);
    }
  /* vertical at u points i in [2, ide-1] */
  for (int j = 1; j <= jde - 1; j++)
    for (int i = 2; i <= ide - 1; i++) {
      real vf[512];
      vf[1] = 0; vf[KDE] = 0;
      for (int k = 3; k <= ktf - 1; k++) {
        real vel = (real)0.5 * (A3(rom, i - 1, k, j) + A3(rom, i, k, j));
        vf[k] = vel * flux3(A3(u, i, k - 2, j), A3(u, i, k - 1, j), A3(u, i, k, j), A3(u, i, k + 1, j), -vel);
      }
      int kk[2] = {2, ktf};
      for (int n = 0; n < 2; n++) {
        int k = kk[n];
        vf[k] = (real)0.5 * (A3(rom, i, k, j) + A3(rom, i - 1, k, j)) * (A1(fzm, k) * A3(u, i, k, j) + A1(fzp, k) * A3(u, i, k - 1, j));
      }
      for (int k = 1; k <= ktf; k++) A3(tend, i, k, j) -= A1(rdzw, k) * (vf[k + 1] - vf[k]);
    }
  free(cur); free(prev);
}

/* advect_v (module_advect_em.py:1037-1354), specified boundaries:
 * v rows j in [2, jde-1], columns i in [1, ide-1]. */
void o_advect_v(OCtx *x, const real *v, real *tend, const real *ru, const real *rv, const real *rom) {
  real *msfvy = FLD(msfvy), *msfvx = FLD(msfvx), *fzm = FLD(fnm), *fzp = FLD(fnp), *rdzw = FLD(rdnw);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  int ktf = KDE - 1, jde = JDE, ide = IDE;
  size_t w = (size_t)(IDE + 2);
  real *cur = (real *)malloc(sizeof(real) * w * (KDE + 2));
  real *prev = (real *)malloc(sizeof(real) * w * (KDE + 2));
  /* y: fluxes at j in [2, jde], tendency rows [2, jde-1] */
  for (int j = 2; j <= jde; j++) {
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= ide - 1; i++) {
        real val;
        if (j >= 4 && j <= jde - 2) {
          real vel = (real)0.5 * (A3(rv, i, k, j) + A3(rv, i, k, j - 1));
          val = vel * flux5(A3(v, i, k, j - 3), A3(v, i, k, j - 2), A3(v, i, k, j - 1), A3(v, i, k, j), A3(v, i, k, j + 1), A3(v, i, k, j + 2), vel);
        } else if (j == 2) {
          real vb = A3(v, i, k, j - 1);
          if (A3(v, i, k, j) < (real)0) vb = A3(v, i, k, j);
          val = (real)0.25 * (A3(rv, i, k, j) + A3(rv, i, k, j - 1)) * (A3(v, i, k, j) + vb);
        } else if (j == 3 || j == jde - 1) {
          real vel = (real)0.5 * (A3(rv, i, k, j) + A3(rv, i, k, j - 1));
          val = vel * flux3(A3(v, i, k, j - 2), A3(v, i, k, j - 1), A3(v, i, k, j), A3(v, i, k, j + 1), vel);
        } else { /* j == jde */
          real vb = A3(v, i, k, j);
          if (A3(v, i, k, j - 1) > (real)0) vb = A3(v, i, k, j - 1);
          val = (real)0.25 * (A3(rv, i, k, j) + A3(rv, i, k, j - 1)) * (vb + A3(v, i, k, j - 1));
        }
        cur[k * w + i] = val;
      }
    if (j > 2)
      for (int k = 1; k <= ktf; k++)
        for (int i = 1; i <= ide - 1; i++)
          A3(tend, i, k, j - 1) -= A2(msfvy, i, j - 1) * rdy * (cur[k * w + i] - prev[k * w + i]);
    real *t = prev; prev = cur; cur = t;
  }
  /* x: rows j in [2, jde-1]; faces [2, ide-1], tendency columns [2, ide-2] */
  for (int j = 2; j <= jde - 1; j++)
    for (int k = 1; k <= ktf; k++) {
      real *fqx = cur;
      for (int i = 4; i <= ide - 3; i++) {
        real vel = (real)0.5 * (A3(ru, i, k, j) + A3(ru, i, k, j - 1));
        fqx[i] = vel * flux5(A3(v, i - 3, k, j), A3(v, i - 2, k, j), A3(v, i - 1, k, j), A3(v, i, k, j), A3(v, i + 1, k, j), A3(v, i + 2, k, j), vel);
      }
      fqx[2] = (real)0.25 * (A3(ru, 2, k, j) + A3(ru, 2, k, j - 1)) * (A3(v, 2, k, j) + A3(v, 1, k, j));
      {
        int i = 3;
        real vel = (real)0.5 * (A3(ru, i, k, j) + A3(ru, i, k, j - 1));
        fqx[i] = vel * flux3(A3(v, i - 2, k, j), A3(v, i - 1, k, j), A3(v, i, k, j), A3(v, i + 1, k, j), vel);
      }
      {
        int i = ide - 1;
        fqx[i] = (real)0.25 * (A3(ru, i, k, j) + A3(ru, i, k, j - 1)) * (A3(v, i, k, j) + A3(v, i - 1, k, j));
        i = ide - 2;
        real vel = (real)0.5 * (A3(ru, i, k, j) + A3(ru, i, k, j - 1));
        fqx[i] = vel * flux3(A3(v, i - 2, k, j), A3(v, i - 1, k, j), A3(v, i, k, j), A3(v, i + 1, k, j), vel);
      }
      for (int i = 2; i <= ide - 2; i++) A3(tend, i, k, j) -= A2(msfvy, i, j) * rdx * (fqx[i + 1] - fqx[i]);
    }
  /* vertical: rows [2, jde-1], columns [1, ide-1] */
  for (int j = 2; j <= jde - 1; j++)
    for (int i = 1; i <= ide - 1; i++) {
      real vf[512];
      vf[1] = 0; vf[KDE] = 0;
      for (int k = 3; k <= ktf - 1; k++) {
        real vel = (real)0.5 * (A3(rom, i, k, j) + A3(rom, i, k, j - 1));
        vf[k] = vel * flux3(A3(v, i, k - 2, j), A3(v, i, k - 1, j), A3(v, i, k, j), A3(v, i, k + 1, j), -vel);
      }
      int kk[2] = {2, ktf};
      for (int n = 0; n < 2; n++) {
        int k = kk[n];
        vf[k] = (real)0.5 * (A3(rom, i, k, j) + A3(rom, i, k, j - 1)) * (A1(fzm, k) * A3(v, i, k, j) + A1(fzp, k) * A3(v, i, k - 1, j));
      }
      for (int k = 1; k <= ktf; k++)
        A3(tend, i, k, j) -= (A2(msfvy, i, j) / A2(msfvx, i, j)) * A1(rdzw, k) * (vf[k + 1] - vf[k]);
    }
  free(cur); free(prev);
}

/* advect_w (module_advect_em.py:1357-1640), specified boundaries, levels
 * k in [2, kde]; rdzu = rdn. */
void o_advect_w(OCtx *x, const real *w, real *tend, const real *ru, const real *rv, const real *rom) {
  real *msftx = FLD(msftx), *fzm = FLD(fnm), *fzp = FLD(fnp), *rdzu = FLD(rdn);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  int ktf = KDE - 1, jde = JDE, ide = IDE;
  size_t wd = (size_t)(IDE + 2);
#define RUW(i, k, j) ((k) == ktf + 1 ? ((real)2 - A1(fzm, (k) - 1)) * A3(ru, i, (k) - 1, j) - A1(fzp, (k) - 1) * A3(ru, i, (k) - 2, j) \
                                     : A1(fzm, k) * A3(ru, i, k, j) + A1(fzp, k) * A3(ru, i, (k) - 1, j))
#define RVW(i, k, j) ((k) == ktf + 1 ? ((real)2 - A1(fzm, (k) - 1)) * A3(rv, i, (k) - 1, j) - A1(fzp, (k) - 1) * A3(rv, i, (k) - 2, j) \
                                     : A1(fzm, k) * A3(rv, i, k, j) + A1(fzp, k) * A3(rv, i, (k) - 1, j))
  real *cur = (real *)malloc(sizeof(real) * wd * (KDE + 2));
  real *prev = (real *)malloc(sizeof(real) * wd * (KDE + 2));
  for (int j = 2; j <= jde - 1; j++) {
    for (int k = 2; k <= ktf + 1; k++)
      for (int i = 1; i <= ide - 1; i++) {
        real vel = RVW(i, k, j), val;
        if (j >= 4 && j <= jde - 3)
          val = vel * flux5(A3(w, i, k, j - 3), A3(w, i, k, j - 2), A3(w, i, k, j - 1), A3(w, i, k, j), A3(w, i, k, j + 1), A3(w, i, k, j + 2), vel);
        else if (j == 2 || j == jde - 1)
          val = (real)0.5 * vel * (A3(w, i, k, j) + A3(w, i, k, j - 1));
        else
          val = vel * flux3(A3(w, i, k, j - 2), A3(w, i, k, j - 1), A3(w, i, k, j), A3(w, i, k, j + 1), vel);
        cur[k * wd + i] = val;
      }
    if (j > 2)
      for (int k = 2; k <= ktf + 1; k++)
        for (int i = 1; i <= ide - 1; i++)
          A3(tend, i, k, j - 1) -= A2(msftx, i, j - 1) * rdy * (cur[k * wd + i] - prev[k * wd + i]);
    real *t = prev; prev = cur; cur = t;
  }
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 2; k <= ktf + 1; k++) {
      real *fqx = cur;
      for (int i = 2; i <= ide - 1; i++) {
        real vel = RUW(i, k, j), val;
        if (i >= 4 && i <= ide - 3)
          val = vel * flux5(A3(w, i - 3, k, j), A3(w, i - 2, k, j), A3(w, i - 1, k, j), A3(w, i, k, j), A3(w, i + 1, k, j), A3(w, i + 2, k, j), vel);
        else if (i == 2 || i == ide - 1)
          val = (real)0.5 * vel * (A3(w, i, k, j) + A3(w, i - 1, k, j));
        else
          val = vel * flux3(A3(w, i - 2, k, j), A3(w, i - 1, k, j), A3(w, i, k, j), A3(w, i + 1, k, j), vel);
        fqx[i] = val;
      }
      for (int i = 2; i <= ide - 2; i++) A3(tend, i, k, j) -= A2(msftx, i, j) * rdx * (fqx[i + 1] - fqx[i]);
    }
  for (int j = 1; j <= jde - 1; j++)
    for (int i = 1; i <= ide - 1; i++) {
      real vf[512];
      vf[1] = 0; vf[KDE] = 0;
      for (int k = 3; k <= ktf; k++) {
        real vel = (real)0.5 * (A3(rom, i, k, j) + A3(rom, i, k - 1, j));
        vf[k] = vel * flux3(A3(w, i, k - 2, j), A3(w, i, k - 1, j), A3(w, i, k, j), A3(w, i, k + 1, j), -vel);
      }
      int kk[2] = {2, ktf + 1};
      for (int n = 0; n < 2; n++) {
        int k = kk[n];
        vf[k] = (real)0.25 * (A3(rom, i, k, j) + A3(rom, i, k - 1, j)) * (A3(w, i, k, j) + A3(w, i, k - 1, j));
      }
      for (int k = 2; k <= ktf; k++) A3(tend, i, k, j) -= A1(rdzu, k) * (vf[k + 1] - vf[k]);
      A3(tend, i, ktf + 1, j) += (real)2 * A1(rdzu, ktf) * vf[ktf + 1];
    }
#undef RUW
#undef RVW
  free(cur); free(prev);
}

/* advect_scalar_pd (module_advect_em.py:372-759): Skamarock (2006)
 * positive-definite flux renormalisation, specified boundaries.
 * mut = column mass (muts), mu_old = mu_1, dt = dt_rk. */
void o_advect_scalar_pd(OCtx *x, const real *q, const real *q_old, real *tend,
                        const real *ru, const real *rv, const real *rom,
                        const real *mut, const real *mub, const real *mu_old, double dtd) {
  real *c1 = FLD(c1h), *c2 = FLD(c2h), *msftx = FLD(msftx), *msfty = FLD(msfty);
  real *fzm = FLD(fnm), *fzp = FLD(fnp), *rdzw = FLD(rdnw);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy), dt = (real)dtd;
  int ktf = KDE - 1, jde = JDE, ide = IDE;
  size_t n = x->n3;
  real *fqx = (real *)calloc(n, sizeof(real)), *fqy = (real *)calloc(n, sizeof(real)), *fqz = (real *)calloc(n, sizeof(real));
  real *fqxl = (real *)calloc(n, sizeof(real)), *fqyl = (real *)calloc(n, sizeof(real)), *fqzl = (real *)calloc(n, sizeof(real));
  /* y fluxes: j in [2, jde-1], i in [1, ide-1] */
  for (int j = 2; j <= jde - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= ide - 1; i++) {
        real dy = (real)2 / (A2(msftx, i, j) + A2(msftx, i, j - 1)) / rdy;
        real mu = (real)0.5 * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i, j - 1) + A1(c2, k)));
        real vel = A3(rv, i, k, j);
        real cr = vel * dt / dy / mu;
        real low = mu * (dy / dt) * flux_up(A3(q_old, i, k, j - 1), A3(q_old, i, k, j), cr);
        real high;
        if (j >= 4 && j <= jde - 3)
          high = vel * flux5pd(A3(q, i, k, j - 3), A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), A3(q, i, k, j + 2), vel);
        else if (j == 2 || j == jde - 1)
          high = (real)0.5 * vel * (A3(q, i, k, j) + A3(q, i, k, j - 1));
        else
          high = vel * flux3pd(A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), vel);
        A3(fqyl, i, k, j) = low;
        A3(fqy, i, k, j) = high - low;
      }
  /* x fluxes: faces i in [2, ide-1], rows j in [1, jde-1] */
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 2; i <= ide - 1; i++) {
        real dx = (real)2 / (A2(msfty, i, j) + A2(msfty, i - 1, j)) / rdx;
        real mu = (real)0.5 * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i - 1, j) + A1(c2, k)));
        real vel = A3(ru, i, k, j);
        real cr = vel * dt / dx / mu;
        real low = mu * (dx / dt) * flux_up(A3(q_old, i - 1, k, j), A3(q_old, i, k, j), cr);
        real high;
        if (i >= 4 && i <= ide - 3)
          high = vel * flux5pd(A3(q, i - 3, k, j), A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), A3(q, i + 2, k, j), vel);
        else if (i == 2 || i == ide - 1)
          high = (real)0.5 * vel * (A3(q, i, k, j) + A3(q, i - 1, k, j));
        else
          high = vel * flux3pd(A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), vel);
        A3(fqxl, i, k, j) = low;
        A3(fqx, i, k, j) = high - low;
      }
  /* z fluxes: all columns; zero at k=1 and k=kde */
  for (int j = 1; j <= jde - 1; j++)
    for (int i = 1; i <= ide - 1; i++) {
      for (int k = 2; k <= ktf; k++) {
        real dz = (real)2 / (A1(rdzw, k) + A1(rdzw, k - 1));
        real mu = A1(c1, k) * A2(mut, i, j) + A1(c2, k);
        real vel = A3(rom, i, k, j);
        real cr = vel * dt / dz / mu;
        real low = mu * (dz / dt) * flux_up(A3(q_old, i, k - 1, j), A3(q_old, i, k, j), cr);
        real high;
        if (k == 2 || k == ktf)
          high = vel * (A1(fzm, k) * A3(q, i, k, j) + A1(fzp, k) * A3(q, i, k - 1, j));
        else
          high = vel * flux3pd(A3(q, i, k - 2, j), A3(q, i, k - 1, j), A3(q, i, k, j), A3(q, i, k + 1, j), -vel);
        A3(fqzl, i, k, j) = low;
        A3(fqz, i, k, j) = high - low;
      }
    }
  /* limiter on cells i in [2, ide-2], j in [2, jde-2] (specified) */
  const real eps = (real)1.0e-20;
  for (int j = 2; j <= jde - 2; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 2; i <= ide - 2; i++) {
        real mm = A2(msftx, i, j) * A2(msfty, i, j);
        real ph_low = ((A1(c1, k) * A2(mub, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mu_old, i, j))) * A3(q_old, i, k, j) -
                      dt * (mm * (rdx * (A3(fqxl, i + 1, k, j) - A3(fqxl, i, k, j)) + rdy * (A3(fqyl, i, k, j + 1) - A3(fqyl, i, k, j))) +
                            A2(msfty, i, j) * A1(rdzw, k) * (A3(fqzl, i, k + 1, j) - A3(fqzl, i, k, j)));
        real flux_out = dt * (mm * (rdx * (rmax(0, A3(fqx, i + 1, k, j)) - rmin(0, A3(fqx, i, k, j))) +
                                    rdy * (rmax(0, A3(fqy, i, k, j + 1)) - rmin(0, A3(fqy, i, k, j)))) +
                              A2(msfty, i, j) * A1(rdzw, k) * (rmin(0, A3(fqz, i, k + 1, j)) - rmax(0, A3(fqz, i, k, j))));
        if (flux_out > ph_low) {
          real scale = rmax(0, ph_low / (flux_out + eps));
          if (A3(fqx, i + 1, k, j) > 0) A3(fqx, i + 1, k, j) *= scale;
          if (A3(fqx, i, k, j) < 0) A3(fqx, i, k, j) *= scale;
          if (A3(fqy, i, k, j + 1) > 0) A3(fqy, i, k, j + 1) *= scale;
          if (A3(fqy, i, k, j) < 0) A3(fqy, i, k, j) *= scale;
          if (A3(fqz, i, k + 1, j) < 0) A3(fqz, i, k + 1, j) *= scale;
          if (A3(fqz, i, k, j) > 0) A3(fqz, i, k, j) *= scale;
        }
      }
  /* tendencies: z everywhere; x on [2, ide-2]; y on rows [2, jde-2] */
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= ide - 1; i++)
        A3(tend, i, k, j) += -A1(rdzw, k) * (A3(fqz, i, k + 1, j) - A3(fqz, i, k, j) + A3(fqzl, i, k + 1, j) - A3(fqzl, i, k, j));
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 2; i <= ide - 2; i++)
        A3(tend, i, k, j) += -A2(msftx, i, j) * (rdx * (A3(fqx, i + 1, k, j) - A3(fqx, i, k, j) + A3(fqxl, i + 1, k, j) - A3(fqxl, i, k, j)));
  for (int j = 2; j <= jde - 2; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= ide - 1; i++)
        A3(tend, i, k, j) += -A2(msftx, i, j) * (rdy * (A3(fqy, i, k, j + 1) - A3(fqy, i, k, j) + A3(fqyl, i, k, j + 1) - A3(fqyl, i, k, j)));
  free(fqx); free(fqy); free(fqz); free(fqxl); free(fqyl); free(fqzl);
}
