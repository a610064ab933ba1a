/*
 * Oracle: diff_opt=2 / km_opt=4 horizontal diffusion (test infrastructure).
 * Follows src/wrf_jax/dyn_em/module_diffusion_em.py (loop form), specified
 * boundaries, no SFS/NBA terms, mix_full_fields = false.
 */
#include "o_core.h"

/* rho from phy_prep (module_big_step_utilities_em.py:1823-2081, rho at
 * :1965): rho = (1 + qv) / alt on mass points. */
void o_phy_prep_rho(OCtx *x) {
  real *rho = FLD(rho), *alt = FLD(alt);
  const real *qv = SPC(moist, 0);
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) A3(rho, i, k, j) = ((real)1 / A3(alt, i, k, j)) * ((real)1 + A3(qv, i, k, j));
}

/* compute_diff_metrics (module_diffusion_em.py:1825-2043) from ph_2/phb. */
void o_compute_diff_metrics(OCtx *x) {
  real *ph = FLD(ph_2), *phb = FLD(phb), *z = FLD(z), *rdz = FLD(rdz), *rdzw = FLD(rdzw), *zx = FLD(zx), *zy = FLD(zy);
  real g = (real)x->c.g, rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  int ktf = KDE - 1;
#define ZW(i, k, j) ((A3(ph, i, k, j) + A3(phb, i, k, j)) / g)
  for (int j = 1; j <= JDE; j++)
    for (int i = 1; i <= IDE - 1; i++) {
      for (int k = 1; k <= ktf; k++) A3(rdzw, i, k, j) = (real)1 / (ZW(i, k + 1, j) - ZW(i, k, j));
      for (int k = 2; k <= ktf; k++) A3(rdz, i, k, j) = (real)2 / (ZW(i, k + 1, j) - ZW(i, k - 1, j));
      A3(rdz, i, 1, j) = (real)2 / (ZW(i, 2, j) - ZW(i, 1, j));
    }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE; k++)
      for (int i = 2; i <= IDE - 1; i++) {
        real zb = (rdx * (A3(phb, i, k, j) - A3(phb, i - 1, k, j))) / g;
        real zp = (rdx * (A3(ph, i, k, j) - A3(ph, i - 1, k, j))) / g;
        A3(zx, i, k, j) = zb + zp;
      }
  for (int j = 2; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real zb = (rdy * (A3(phb, i, k, j) - A3(phb, i, k, j - 1))) / g;
        real zp = (rdy * (A3(ph, i, k, j) - A3(ph, i, k, j - 1))) / g;
        A3(zy, i, k, j) = zb + zp;
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= ktf; k++) {
      A3(zx, IDE, k, j) = 0;
      A3(zx, 1, k, j) = 0;
    }
  for (int i = 1; i <= IDE - 1; i++)
    for (int k = 1; k <= ktf; k++) {
      A3(zy, i, k, JDE) = 0;
      A3(zy, i, k, 1) = 0;
    }
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real s = A3(ph, i, k, j) + A3(phb, i, k, j);
        s = s + A3(ph, i, k + 1, j);
        s = s + A3(phb, i, k + 1, j);
        A3(z, i, k, j) = ((real)0.5 * s) / g;
      }
#undef ZW
}

/* cal_deform_and_div (module_diffusion_em.py:366-891): deformation tensor
 * and divergence with terrain-slope corrections (u_base = v_base inputs). */
void o_cal_deform_and_div(OCtx *x) {
  real *u = FLD(u_2), *v = FLD(v_2), *w = FLD(w_2), *div = FLD(div);
  real *d11 = FLD(defor11), *d22 = FLD(defor22), *d33 = FLD(defor33), *d12 = FLD(defor12), *d13 = FLD(defor13), *d23 = FLD(defor23);
  real *msfux = FLD(msfux), *msfuy = FLD(msfuy), *msfvx = FLD(msfvx), *msfvy = FLD(msfvy), *msftx = FLD(msftx), *msfty = FLD(msfty);
  real *fnm = FLD(fnm), *fnp = FLD(fnp), *dn = FLD(dn), *dnw = FLD(dnw), *rdz = FLD(rdz), *rdzw = FLD(rdzw), *zx = FLD(zx), *zy = FLD(zy);
  real *ubase = FLD(u_base), *vbase = FLD(v_base);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  real cf1 = (real)x->c.cf1, cf2 = (real)x->c.cf2, cf3 = (real)x->c.cf3;
  int ide = IDE, jde = JDE, kte = KDE, ktf = KDE - 1, ktes1 = kte - 1, ktes2 = kte - 2;
  real cft2 = (real)-0.5 * A1(dnw, ktes1) / A1(dn, ktes1), cft1 = (real)1 - cft2;
  real *hat = (real *)calloc(x->n3, sizeof(real)), *hatavg = (real *)calloc(x->n3, sizeof(real));
  /* D11 */
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= ide; i++) A3(hat, i, k, j) = A3(u, i, k, j) / A2(msfuy, i, j);
  for (int j = 1; j <= jde - 1; j++)
    for (int i = 1; i <= ide - 1; i++) {
      for (int k = 2; k <= ktf; k++)
        A3(hatavg, i, k, j) = (real)0.5 * (A1(fnm, k) * (A3(hat, i, k, j) + A3(hat, i + 1, k, j)) + A1(fnp, k) * (A3(hat, i, k - 1, j) + A3(hat, i + 1, k - 1, j)));
      A3(hatavg, i, 1, j) = (real)0.5 * (cf1 * A3(hat, i, 1, j) + cf2 * A3(hat, i, 2, j) + cf3 * A3(hat, i, 3, j) + cf1 * A3(hat, i + 1, 1, j) +
                                         cf2 * A3(hat, i + 1, 2, j) + cf3 * A3(hat, i + 1, 3, j));
      A3(hatavg, i, kte, j) = (real)0.5 * (cft1 * (A3(hat, i, ktes1, j) + A3(hat, i + 1, ktes1, j)) + cft2 * (A3(hat, i, ktes2, j) + A3(hat, i + 1, ktes2, j)));
      for (int k = 1; k <= ktf; k++) {
        real mm = A2(msftx, i, j) * A2(msfty, i, j);
        real tmpzx = (real)0.25 * (A3(zx, i, k, j) + A3(zx, i + 1, k, j) + A3(zx, i, k + 1, j) + A3(zx, i + 1, k + 1, j));
        real t1 = (A3(hatavg, i, k + 1, j) - A3(hatavg, i, k, j)) * tmpzx * A3(rdzw, i, k, j);
        t1 = mm * (rdx * (A3(hat, i + 1, k, j) - A3(hat, i, k, j)) - t1);
        A3(d11, i, k, j) = (real)2 * t1;
        A3(div, i, k, j) = t1;
      }
    }
  /* D22 */
  for (int j = 1; j <= jde; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= ide - 1; i++) A3(hat, i, k, j) = A3(v, i, k, j) / A2(msfvx, i, j);
  for (int j = 1; j <= jde - 1; j++)
    for (int i = 1; i <= ide - 1; i++) {
      for (int k = 2; k <= ktf; k++)
        A3(hatavg, i, k, j) = (real)0.5 * (A1(fnm, k) * (A3(hat, i, k, j) + A3(hat, i, k, j + 1)) + A1(fnp, k) * (A3(hat, i, k - 1, j) + A3(hat, i, k - 1, j + 1)));
      A3(hatavg, i, 1, j) = (real)0.5 * (cf1 * A3(hat, i, 1, j) + cf2 * A3(hat, i, 2, j) + cf3 * A3(hat, i, 3, j) + cf1 * A3(hat, i, 1, j + 1) +
                                         cf2 * A3(hat, i, 2, j + 1) + cf3 * A3(hat, i, 3, j + 1));
      A3(hatavg, i, kte, j) = (real)0.5 * (cft1 * (A3(hat, i, ktes1, j) + A3(hat, i, ktes1, j + 1)) + cft2 * (A3(hat, i, ktes2, j) + A3(hat, i, ktes2, j + 1)));
      for (int k = 1; k <= ktf; k++) {
        real mm = A2(msftx, i, j) * A2(msfty, i, j);
        real tmpzy = (real)0.25 * (A3(zy, i, k, j) + A3(zy, i, k, j + 1) + A3(zy, i, k + 1, j) + A3(zy, i, k + 1, j + 1));
        real t1 = (A3(hatavg, i, k + 1, j) - A3(hatavg, i, k, j)) * tmpzy * A3(rdzw, i, k, j);
        t1 = mm * (rdy * (A3(hat, i, k, j + 1) - A3(hat, i, k, j)) - t1);
        A3(d22, i, k, j) = (real)2 * t1;
        A3(div, i, k, j) = A3(div, i, k, j) + t1;
      }
    }
  /* D33 */
  for (int j = 1; j <= jde - 1; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 1; i <= ide - 1; i++) {
        real t1 = (A3(w, i, k + 1, j) - A3(w, i, k, j)) * A3(rdzw, i, k, j);
        A3(d33, i, k, j) = (real)2 * t1;
        A3(div, i, k, j) = A3(div, i, k, j) + t1;
      }
  /* D12 at (u, v) corner points, i in [2, ide-1], j in [2, jde-1] */
  {
    int i0 = 2, i1 = ide - 1, j0 = 2, j1 = jde - 1;
    for (int j = j0 - 1; j <= j1; j++)
      for (int k = 1; k <= ktf; k++)
        for (int i = i0; i <= i1; i++) A3(hat, i, k, j) = A3(u, i, k, j) / A2(msfux, i, j);
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        for (int k = 2; k <= ktf; k++)
          A3(hatavg, i, k, j) = (real)0.5 * (A1(fnm, k) * (A3(hat, i, k, j - 1) + A3(hat, i, k, j)) + A1(fnp, k) * (A3(hat, i, k - 1, j - 1) + A3(hat, i, k - 1, j)));
        A3(hatavg, i, 1, j) = (real)0.5 * (cf1 * A3(hat, i, 1, j - 1) + cf2 * A3(hat, i, 2, j - 1) + cf3 * A3(hat, i, 3, j - 1) + cf1 * A3(hat, i, 1, j) +
                                           cf2 * A3(hat, i, 2, j) + cf3 * A3(hat, i, 3, j));
        A3(hatavg, i, kte, j) = (real)0.5 * (cft1 * (A3(hat, i, ktes1, j - 1) + A3(hat, i, ktes1, j)) + cft2 * (A3(hat, i, ktes2, j - 1) + A3(hat, i, ktes2, j)));
        real mm = (real)0.25 * (A2(msfux, i, j - 1) + A2(msfux, i, j)) * (A2(msfvy, i - 1, j) + A2(msfvy, i, j));
        for (int k = 1; k <= ktf; k++) {
          real tmpzy = (real)0.25 * (A3(zy, i - 1, k, j) + A3(zy, i, k, j) + A3(zy, i - 1, k + 1, j) + A3(zy, i, k + 1, j));
          real t1 = (A3(hatavg, i, k + 1, j) - A3(hatavg, i, k, j)) * (real)0.25 * tmpzy *
                    (A3(rdzw, i, k, j) + A3(rdzw, i - 1, k, j) + A3(rdzw, i - 1, k, j - 1) + A3(rdzw, i, k, j - 1));
          A3(d12, i, k, j) = mm * (rdy * (A3(hat, i, k, j) - A3(hat, i, k, j - 1)) - t1);
        }
      }
    for (int j = j0; j <= j1; j++)
      for (int k = 1; k <= ktf; k++)
        for (int i = i0 - 1; i <= i1; i++) A3(hat, i, k, j) = A3(v, i, k, j) / A2(msfvy, i, j);
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        for (int k = 2; k <= ktf; k++)
          A3(hatavg, i, k, j) = (real)0.5 * (A1(fnm, k) * (A3(hat, i - 1, k, j) + A3(hat, i, k, j)) + A1(fnp, k) * (A3(hat, i - 1, k - 1, j) + A3(hat, i, k - 1, j)));
        A3(hatavg, i, 1, j) = (real)0.5 * (cf1 * A3(hat, i - 1, 1, j) + cf2 * A3(hat, i - 1, 2, j) + cf3 * A3(hat, i - 1, 3, j) + cf1 * A3(hat, i, 1, j) +
                                           cf2 * A3(hat, i, 2, j) + cf3 * A3(hat, i, 3, j));
        A3(hatavg, i, kte, j) = (real)0.5 * (cft1 * (A3(hat, i, ktes1, j) + A3(hat, i - 1, ktes1, j)) + cft2 * (A3(hat, i, ktes2, j) + A3(hat, i - 1, ktes2, j)));
        real mm = (real)0.25 * (A2(msfux, i, j - 1) + A2(msfux, i, j)) * (A2(msfvy, i - 1, j) + A2(msfvy, i, j));
        for (int k = 1; k <= ktf; k++) {
          real tmpzx = (real)0.25 * (A3(zx, i, k, j - 1) + A3(zx, i, k, j) + A3(zx, i, k + 1, j - 1) + A3(zx, i, k + 1, j));
          real t1 = (A3(hatavg, i, k + 1, j) - A3(hatavg, i, k, j)) * (real)0.25 * tmpzx *
                    (A3(rdzw, i, k, j) + A3(rdzw, i, k, j - 1) + A3(rdzw, i - 1, k, j - 1) + A3(rdzw, i - 1, k, j));
          real second = mm * (rdx * (A3(hat, i, k, j) - A3(hat, i - 1, k, j)) - t1);
          A3(d12, i, k, j) = A3(d12, i, k, j) + second;
        }
      }
    /* edge copies (reference order: ids, jds, ide, jde) */
    for (int j = 1; j <= jde; j++)
      for (int k = 1; k <= kte; k++) A3(d12, 1, k, j) = A3(d12, 2, k, j);
    for (int k = 1; k <= kte; k++)
      for (int i = 1; i <= ide; i++) A3(d12, i, k, 1) = A3(d12, i, k, 2);
    for (int j = 1; j <= jde; j++)
      for (int k = 1; k <= kte; k++) A3(d12, ide, k, j) = A3(d12, ide - 1, k, j);
    for (int k = 1; k <= kte; k++)
      for (int i = 1; i <= ide; i++) A3(d12, i, k, jde) = A3(d12, i, k, jde - 1);
  }
  /* D13 at (u, w) points: i in [2, ide-1], j in [2, jde-1] */
  {
    int i0 = 2, i1 = ide - 1, j0 = 2, j1 = jde - 1;
    for (int j = j0; j <= j1; j++)
      for (int k = 1; k <= kte; k++)
        for (int i = i0; i <= i1; i++) A3(hat, i, k, j) = A3(w, i, k, j) / A2(msfty, i, j);
    for (int j = j0; j <= imin(JDE - 1, JDE - 1); j++)
      for (int k = 1; k <= kte; k++) A3(hat, i0 - 1, k, j) = A3(w, i0 - 1, k, j) / A2(msfty, i0 - 1, j);
    for (int k = 1; k <= kte; k++)
      for (int i = i0; i <= ide - 1; i++) A3(hat, i, k, j0 - 1) = A3(w, i, k, j0 - 1) / A2(msfty, i, j0 - 1);
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        real mm = A2(msfux, i, j) * A2(msfuy, i, j);
        for (int k = 1; k <= ktf; k++)
          A3(hatavg, i, k, j) = (real)0.25 * (A3(hat, i, k, j) + A3(hat, i, k + 1, j) + A3(hat, i - 1, k, j) + A3(hat, i - 1, k + 1, j));
        for (int k = 2; k <= ktf; k++) {
          real t1 = (A3(hatavg, i, k, j) - A3(hatavg, i, k - 1, j)) * A3(zx, i, k, j) * (real)0.5 * (A3(rdz, i, k, j) + A3(rdz, i - 1, k, j));
          A3(d13, i, k, j) = mm * (rdx * (A3(hat, i, k, j) - A3(hat, i - 1, k, j)) - t1);
        }
        A3(d13, i, 1, j) = 0;
        A3(d13, i, ktf + 1, j) = 0;
        for (int k = 2; k <= ktf; k++) {
          real val = (A3(u, i, k, j) - A1(ubase, k) - A3(u, i, k - 1, j) + A1(ubase, k - 1)) * (real)0.5 * (A3(rdz, i, k, j) + A3(rdz, i - 1, k, j));
          A3(d13, i, k, j) = A3(d13, i, k, j) + val;
        }
      }
  }
  /* D23 at (v, w) points: i in [2, ide-1], j in [2, jde-1] */
  {
    int i0 = 2, i1 = ide - 1, j0 = 2, j1 = jde - 1;
    for (int j = j0; j <= j1; j++)
      for (int k = 1; k <= kte; k++)
        for (int i = i0; i <= i1; i++) A3(hat, i, k, j) = A3(w, i, k, j) / A2(msftx, i, j);
    for (int j = j0; j <= JDE - 1; j++)
      for (int k = 1; k <= kte; k++) A3(hat, i0 - 1, k, j) = A3(w, i0 - 1, k, j) / A2(msftx, i0 - 1, j);
    for (int k = 1; k <= kte; k++)
      for (int i = i0; i <= ide - 1; i++) A3(hat, i, k, j0 - 1) = A3(w, i, k, j0 - 1) / A2(msftx, i, j0 - 1);
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        real mm = A2(msfvx, i, j) * A2(msfvy, i, j);
        for (int k = 1; k <= ktf; k++)
          A3(hatavg, i, k, j) = (real)0.25 * (A3(hat, i, k, j) + A3(hat, i, k + 1, j) + A3(hat, i, k, j - 1) + A3(hat, i, k + 1, j - 1));
        for (int k = 2; k <= ktf; k++) {
          real t1 = (A3(hatavg, i, k, j) - A3(hatavg, i, k - 1, j)) * A3(zy, i, k, j) * (real)0.5 * (A3(rdz, i, k, j) + A3(rdz, i, k, j - 1));
          A3(d23, i, k, j) = mm * (rdy * (A3(hat, i, k, j) - A3(hat, i, k, j - 1)) - t1);
        }
        A3(d23, i, 1, j) = 0;
        A3(d23, i, ktf + 1, j) = 0;
        for (int k = 2; k <= ktf; k++) {
          real val = (A3(v, i, k, j) - A1(vbase, k) - A3(v, i, k - 1, j) + A1(vbase, k - 1)) * (real)0.5 * (A3(rdz, i, k, j) + A3(rdz, i, k, j - 1));
          A3(d23, i, k, j) = A3(d23, i, k, j) + val;
        }
      }
  }
  /* edge copies for D13/D23 (reference order: ids, jds, ide, jde) */
  for (int j = 1; j <= jde; j++)
    for (int k = 1; k <= kte; k++) {
      A3(d13, 1, k, j) = A3(d13, 2, k, j);
      A3(d23, 1, k, j) = A3(d23, 2, k, j);
    }
  for (int k = 1; k <= kte; k++)
    for (int i = 1; i <= ide; i++) {
      A3(d13, i, k, 1) = A3(d13, i, k, 2);
      A3(d23, i, k, 1) = A3(d23, i, k, 2);
    }
  for (int j = 1; j <= jde; j++)
    for (int k = 1; k <= kte; k++) {
      A3(d13, ide, k, j) = A3(d13, ide - 1, k, j);
      A3(d23, ide, k, j) = A3(d23, ide - 1, k, j);
    }
  for (int k = 1; k <= kte; k++)
    for (int i = 1; i <= ide; i++) {
      A3(d13, i, k, jde) = A3(d13, i, k, jde - 1);
      A3(d23, i, k, jde) = A3(d23, i, k, jde - 1);
    }
  free(hat);
  free(hatavg);
}

/* smag2d_km (module_diffusion_em.py:1502-1646) with the diff_opt=2 slope
 * correction; computed on i in [2, ide-2], j in [2, jde-2]. */
void o_smag2d_km(OCtx *x) {
  real *d11 = FLD(defor11), *d22 = FLD(defor22), *d12 = FLD(defor12), *zx = FLD(zx), *zy = FLD(zy), *rdzw = FLD(rdzw);
  real *msftx = FLD(msftx), *msfty = FLD(msfty), *xkmh = FLD(xkmh), *xkmv = FLD(xkmv), *xkhh = FLD(xkhh), *xkhv = FLD(xkhv);
  real dx = (real)x->c.dx, dy = (real)x->c.dy, c_s = (real)x->c.c_s, pr = (real)x->c.prandtl;
  for (int j = 2; j <= JDE - 2; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 2; i <= IDE - 2; i++) {
        real dd = A3(d11, i, k, j) - A3(d22, i, k, j);
        real def2 = (real)0.25 * dd * dd;
        real tmp = (real)0.25 * (A3(d12, i, k, j) + A3(d12, i, k, j + 1) + A3(d12, i + 1, k, j) + A3(d12, i + 1, k, j + 1));
        def2 = def2 + tmp * tmp;
        real mlen = (real)sqrt(dx / A2(msftx, i, j) * dy / A2(msfty, i, j));
        real sd = (real)sqrt(def2);
        real km = c_s * c_s * mlen * mlen * sd;
        km = rmin(km, (real)10 * mlen);
        real dxm = dx / A2(msftx, i, j), dym = dy / A2(msfty, i, j);
        real tzx = (real)0.25 * (fabs(A3(zx, i, k, j)) + fabs(A3(zx, i + 1, k, j)) + fabs(A3(zx, i, k + 1, j)) + fabs(A3(zx, i + 1, k + 1, j))) *
                   A3(rdzw, i, k, j) * dxm;
        real tzy = (real)0.25 * (fabs(A3(zy, i, k, j)) + fabs(A3(zy, i, k, j + 1)) + fabs(A3(zy, i, k + 1, j)) + fabs(A3(zy, i, k + 1, j + 1))) *
                   A3(rdzw, i, k, j) * dym;
        real alpha = rmax((real)sqrt(tzx * tzx + tzy * tzy), (real)1);
        real lim = rmax((real)10 / mlen, (real)1.0e-3);
        real divisor = sd > lim ? alpha * alpha : alpha;
        km = km / divisor;
        A3(xkmh, i, k, j) = km;
        A3(xkmv, i, k, j) = km;
        A3(xkhh, i, k, j) = km / pr;
        A3(xkhv, i, k, j) = 0;
      }
}

/* phy_prep (module_big_step_utilities_em.py:1823-2081, use_theta_m 0): full
 * theta, pressure and temperature at mass points and the surface values of
 * pressure and temperature on w level 1 (linear in height through mass
 * levels 1 and 2).  Arrays are full 3-D temporaries. */
static void phy_prep_thermo(OCtx *x, real *th, real *pp, real *tt, real *p8w1, real *t8w1) {
  real *t2 = FLD(t_2), *p = FLD(p), *pb = FLD(pb), *ph = FLD(ph_2), *phb = FLD(phb);
  real t0 = (real)x->c.t0, g = (real)x->c.g;
  real rcp = (real)(287.0 / (7.0 * 287.0 / 2.0)), p1000 = (real)100000.0;
  for (int j = 1; j <= JDE - 1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = 1; i <= IDE - 1; i++) {
        real thv = A3(t2, i, k, j) + t0;
        real pv = A3(p, i, k, j) + A3(pb, i, k, j);
        real pi = (real)pow((double)(pv / p1000), (double)rcp);
        A3(th, i, k, j) = thv;
        A3(pp, i, k, j) = pv;
        A3(tt, i, k, j) = thv * pi;
      }
  for (int j = 1; j <= JDE - 1; j++)
    for (int i = 1; i <= IDE - 1; i++) {
      real zw1 = (A3(phb, i, 1, j) + A3(ph, i, 1, j)) / g, zw2 = (A3(phb, i, 2, j) + A3(ph, i, 2, j)) / g;
      real zw3 = (A3(phb, i, 3, j) + A3(ph, i, 3, j)) / g;
      real z1 = (real)0.5 * (zw1 + zw2), z2 = (real)0.5 * (zw2 + zw3);
      real w1 = (zw1 - z2) / (z1 - z2), w2 = (real)1 - w1;
      A2(p8w1, i, j) = w1 * A3(pp, i, 1, j) + w2 * A3(pp, i, 2, j);
      A2(t8w1, i, j) = w1 * A3(tt, i, 1, j) + w2 * A3(tt, i, 2, j);
    }
}

/* Brunt-Vaisala frequency squared of calculate_km_kh
 * (module_diffusion_em_fast.py:933-1005; calculate_n2,
 * module_diffusion_em.py:1319-1499): moist where the air is saturated
 * (qv >= qvs or qc >= 1e-5), dry otherwise; level 1 uses the surface values,
 * level ktf repeats ktf-1.  i [2, ide-2], j [2, jde-2]; qc and qi are moist
 * species 1 and 3 when present. */
void o_calc_n2(OCtx *x) {
  real *bn2 = FLD(bn2), *rdz = FLD(rdz), *rdzw = FLD(rdzw);
  real *th = (real *)calloc(x->n3, sizeof(real)), *pp = (real *)calloc(x->n3, sizeof(real)), *tt = (real *)calloc(x->n3, sizeof(real));
  real *qvs = (real *)calloc(x->n3, sizeof(real)), *qt = (real *)calloc(x->n3, sizeof(real));
  real *p8w1 = (real *)calloc(x->n2, sizeof(real)), *t8w1 = (real *)calloc(x->n2, sizeof(real));
  phy_prep_thermo(x, th, pp, tt, p8w1, t8w1);
  const real *qv = SPC(moist, 0);
  const real *qc = x->c.nmoist > 1 ? SPC(moist, 1) : 0;
  const real *qi = x->c.nmoist > 3 ? SPC(moist, 3) : 0;
  real G = (real)9.81, R_D = (real)287.0, CP = (real)(7.0 * 287.0 / 2.0), XLV = (real)2.5e6, RV = (real)461.6;
  real SVP1 = (real)0.6112, SVP2 = (real)17.67, SVP3 = (real)29.65, SVPT0 = (real)273.15, EP_2 = (real)(287.0 / 461.6);
  real cf1 = (real)x->c.cf1, cf2 = (real)x->c.cf2, cf3 = (real)x->c.cf3, P1000MB = (real)100000.0;
  real qcmin = (real)0.00001, c161 = (real)1.61;
  int ktf = KDE - 1;
  for (int j = 2; j <= JDE - 2; j++)
    for (int k = 1; k <= ktf; k++)
      for (int i = 2; i <= IDE - 2; i++) {
        real tv = A3(tt, i, k, j), tc = tv - SVPT0;
        real es = (real)1000.0 * SVP1 * (real)exp((double)(SVP2 * tc / (tv - SVP3)));
        A3(qvs, i, k, j) = EP_2 * es / (A3(pp, i, k, j) - es);
        real q = A3(qv, i, k, j);
        if (qc) q = q + A3(qc, i, k, j);
        if (qi) q = q + A3(qi, i, k, j);
        A3(qt, i, k, j) = q;
      }
#define QC(i, k, j) (qc ? A3(qc, i, k, j) : (real)0)
  for (int j = 2; j <= JDE - 2; j++)
    for (int i = 2; i <= IDE - 2; i++) {
      for (int k = 2; k <= ktf - 1; k++) {
        real tmpdz = (real)1.0 / A3(rdz, i, k, j) + (real)1.0 / A3(rdz, i, k + 1, j);
        int sat = (A3(qv, i, k, j) >= A3(qvs, i, k, j)) || (QC(i, k, j) >= qcmin);
        real tk = A3(tt, i, k, j), thk = A3(th, i, k, j), xlvqv = XLV * A3(qv, i, k, j);
        real val;
        if (sat) {
          real coefa = ((real)1.0 + xlvqv / R_D / tk) / ((real)1.0 + XLV * xlvqv / CP / RV / tk / tk) / thk;
          real thetaep1 = A3(th, i, k + 1, j) * ((real)1.0 + XLV * A3(qvs, i, k + 1, j) / CP / A3(tt, i, k + 1, j));
          real thetaem1 = A3(th, i, k - 1, j) * ((real)1.0 + XLV * A3(qvs, i, k - 1, j) / CP / A3(tt, i, k - 1, j));
          val = G * (coefa * (thetaep1 - thetaem1) / tmpdz - (A3(qt, i, k + 1, j) - A3(qt, i, k - 1, j)) / tmpdz);
        } else {
          val = G * ((A3(th, i, k + 1, j) - A3(th, i, k - 1, j)) / thk / tmpdz +
                     c161 * (A3(qv, i, k + 1, j) - A3(qv, i, k - 1, j)) / tmpdz - (A3(qt, i, k + 1, j) - A3(qt, i, k - 1, j)) / tmpdz);
        }
        A3(bn2, i, k, j) = val;
      }
      {
        real tmpdz_m = (real)1.0 / A3(rdz, i, 2, j) + (real)0.5 / A3(rdzw, i, 1, j);
        real thsfc = A2(t8w1, i, j) / (real)pow((double)(A2(p8w1, i, j) / P1000MB), (double)(R_D / CP));
        int sat = (A3(qv, i, 1, j) >= A3(qvs, i, 1, j)) || (QC(i, 1, j) >= qcmin);
        real qvsfc_sat = cf1 * A3(qvs, i, 1, j) + cf2 * A3(qvs, i, 2, j) + cf3 * A3(qvs, i, 3, j);
        real t1 = A3(tt, i, 1, j), th1 = A3(th, i, 1, j), xlvqv = XLV * A3(qv, i, 1, j);
        real qtsfc = cf1 * A3(qv, i, 1, j) + cf2 * A3(qv, i, 2, j) + cf3 * A3(qv, i, 3, j);
        if (qc) qtsfc = qtsfc + cf1 * A3(qc, i, 1, j) + cf2 * A3(qc, i, 2, j) + cf3 * A3(qc, i, 3, j);
        if (qi) qtsfc = qtsfc + cf1 * A3(qi, i, 1, j) + cf2 * A3(qi, i, 2, j) + cf3 * A3(qi, i, 3, j);
        real val;
        if (sat) {
          real coefa = ((real)1.0 + xlvqv / R_D / t1) / ((real)1.0 + XLV * xlvqv / CP / RV / t1 / t1) / th1;
          real thetaep1 = A3(th, i, 2, j) * ((real)1.0 + XLV * A3(qvs, i, 2, j) / CP / A3(tt, i, 2, j));
          real thetaesfc = thsfc * ((real)1.0 + XLV * qvsfc_sat / CP / A2(t8w1, i, j));
          val = G * (coefa * (thetaep1 - thetaesfc) / tmpdz_m - (A3(qt, i, 2, j) - qtsfc) / tmpdz_m);
        } else {
          real qvsfc = cf1 * A3(qv, i, 1, j) + cf2 * A3(qv, i, 2, j) + cf3 * A3(qv, i, 3, j);
          real tmpdz_d = (real)1.0 / A3(rdzw, i, 1, j);
          val = G * ((A3(th, i, 2, j) - th1) / th1 / tmpdz_d + c161 * (A3(qv, i, 2, j) - qvsfc) / tmpdz_d - (A3(qt, i, 2, j) - qtsfc) / tmpdz_d);
        }
        A3(bn2, i, 1, j) = val;
      }
      A3(bn2, i, ktf, j) = A3(bn2, i, ktf - 1, j);
    }
#undef QC
  free(th); free(pp); free(tt); free(qvs); free(qt); free(p8w1); free(t8w1);
}

/* titau helpers (module_diffusion_em.py:894-1316): stress components with
 * rho and diffusivity averaged to the component's point.  Output arrays are
 * full 3-D, zero outside the computed range. */
static void titau_11(OCtx *x, real *tt, const real *xk, const real *defor, int is_ext, int ie_ext, int js_ext, int je_ext) {
  real *rho = FLD(rho);
  memset(tt, 0, sizeof(real) * x->n3);
  int i0 = 2 - is_ext, i1 = IDE - 1 + ie_ext, j0 = 2 - js_ext, j1 = JDE - 1 + je_ext;
  for (int j = j0; j <= j1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = i0; i <= i1; i++) A3(tt, i, k, j) = -A3(rho, i, k, j) * A3(xk, i, k, j) * A3(defor, i, k, j);
}
static void titau_12(OCtx *x, real *tt, const real *xk, const real *defor, int is_ext, int ie_ext, int js_ext, int je_ext) {
  real *rho = FLD(rho);
  memset(tt, 0, sizeof(real) * x->n3);
  int i0 = 2 - is_ext, i1 = IDE - 1 + ie_ext, j0 = 2 - js_ext, j1 = JDE - 1 + je_ext;
  for (int j = j0; j <= j1; j++)
    for (int k = 1; k <= KDE - 1; k++)
      for (int i = i0; i <= i1; i++) {
        real ra = (real)0.25 * (A3(rho, i - 1, k, j) + A3(rho, i, k, j) + A3(rho, i - 1, k, j - 1) + A3(rho, i, k, j - 1));
        real xa = ra * (real)0.25 * (A3(xk, i - 1, k, j) + A3(xk, i, k, j) + A3(xk, i - 1, k, j - 1) + A3(xk, i, k, j - 1));
        A3(tt, i, k, j) = -xa * A3(defor, i, k, j);
      }
}
static void titau_13(OCtx *x, real *tt, const real *xk, const real *defor, int is_ext, int ie_ext, int js_ext, int je_ext) {
  real *rho = FLD(rho), *fnm = FLD(fnm), *fnp = FLD(fnp);
  memset(tt, 0, sizeof(real) * x->n3);
  int i0 = 2 - is_ext, i1 = IDE - 1 + ie_ext, j0 = 2 - js_ext, j1 = JDE - 2 + je_ext, ktf = KDE - 1;
  for (int j = j0; j <= j1; j++)
    for (int i = i0; i <= i1; i++) {
      for (int k = 2; k <= ktf; k++) {
        real ra = (real)0.5 * (A1(fnm, k) * (A3(rho, i - 1, k, j) + A3(rho, i, k, j)) + A1(fnp, k) * (A3(rho, i - 1, k - 1, j) + A3(rho, i, k - 1, j)));
        real xa = ra * (real)0.5 * (A1(fnm, k) * (A3(xk, i, k, j) + A3(xk, i - 1, k, j)) + A1(fnp, k) * (A3(xk, i, k - 1, j) + A3(xk, i - 1, k - 1, j)));
        A3(tt, i, k, j) = -xa * A3(defor, i, k, j);
      }
      A3(tt, i, 1, j) = 0;
      A3(tt, i, ktf + 1, j) = 0;
    }
}
static void titau_23(OCtx *x, real *tt, const real *xk, const real *defor, int is_ext, int ie_ext, int js_ext, int je_ext) {
  real *rho = FLD(rho), *fnm = FLD(fnm), *fnp = FLD(fnp);
  memset(tt, 0, sizeof(real) * x->n3);
  int i0 = 2 - is_ext, i1 = IDE - 2 + ie_ext, j0 = 2 - js_ext, j1 = JDE - 1 + je_ext, ktf = KDE - 1;
  for (int j = j0; j <= j1; j++)
    for (int i = i0; i <= i1; i++) {
      for (int k = 2; k <= ktf; k++) {
        real ra = (real)0.5 * (A1(fnm, k) * (A3(rho, i, k, j) + A3(rho, i, k, j - 1)) + A1(fnp, k) * (A3(rho, i, k - 1, j) + A3(rho, i, k - 1, j - 1)));
        real xa = ra * (real)0.5 * (A1(fnm, k) * (A3(xk, i, k, j) + A3(xk, i, k, j - 1)) + A1(fnp, k) * (A3(xk, i, k - 1, j) + A3(xk, i, k - 1, j - 1)));
        A3(tt, i, k, j) = -xa * A3(defor, i, k, j);
      }
      A3(tt, i, 1, j) = 0;
      A3(tt, i, ktf + 1, j) = 0;
    }
}

/* horizontal_diffusion_u_2 (module_diffusion_em.py:2302-2532):
 * u points i in [2, ide-1], j in [2, jde-2]. */
static void hdiff_u(OCtx *x, real *t1, real *t2) {
  real *tend = FLD(ru_tendf), *zx = FLD(zx), *zy = FLD(zy), *rdzw = FLD(rdzw), *fnm = FLD(fnm), *fnp = FLD(fnp), *dnw = FLD(dnw);
  real *msfux = FLD(msfux), *msfuy = FLD(msfuy);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy), g = (real)x->c.g;
  int ktf = KDE - 1;
  titau_11(x, t1, FLD(xkmh), FLD(defor11), 1, 0, 0, 0);
  titau_12(x, t2, FLD(xkmh), FLD(defor12), 0, 0, 0, 1);
  real a1[512], a2[512], zxu[512], zyu[512];
  for (int j = 2; j <= JDE - 2; j++)
    for (int i = 2; i <= IDE - 1; i++) {
      for (int k = 1; k <= ktf; k++) {
        if (k >= 2) {
          a1[k] = (real)0.5 * (A1(fnm, k) * (A3(t1, i - 1, k, j) + A3(t1, i, k, j)) + A1(fnp, k) * (A3(t1, i - 1, k - 1, j) + A3(t1, i, k - 1, j)));
          a2[k] = (real)0.5 * (A1(fnm, k) * (A3(t2, i, k, j + 1) + A3(t2, i, k, j)) + A1(fnp, k) * (A3(t2, i, k - 1, j + 1) + A3(t2, i, k - 1, j)));
        }
        zxu[k] = (real)0.5 * (A3(zx, i, k, j) + A3(zx, i, k + 1, j));
        zyu[k] = (real)0.125 * (A3(zy, i - 1, k, j) + A3(zy, i, k, j) + A3(zy, i - 1, k, j + 1) + A3(zy, i, k, j + 1) + A3(zy, i - 1, k + 1, j) +
                                A3(zy, i, k + 1, j) + A3(zy, i - 1, k + 1, j + 1) + A3(zy, i, k + 1, j + 1));
      }
      a1[1] = 0; a1[ktf + 1] = 0; a2[1] = 0; a2[ktf + 1] = 0;
      for (int k = 1; k <= ktf; k++) {
        real mrdx = A2(msfux, i, j) * rdx, mrdy = A2(msfuy, i, j) * rdy;
        real tmpdz = ((real)1 / A3(rdzw, i, k, j) + (real)1 / A3(rdzw, i - 1, k, j)) / (real)2;
        real inc = g * tmpdz / A1(dnw, k) *
                   (mrdx * (A3(t1, i, k, j) - A3(t1, i - 1, k, j)) + mrdy * (A3(t2, i, k, j + 1) - A3(t2, i, k, j)) -
                    A2(msfux, i, j) * zxu[k] * (a1[k + 1] - a1[k]) / tmpdz - A2(msfuy, i, j) * zyu[k] * (a2[k + 1] - a2[k]) / tmpdz);
        A3(tend, i, k, j) = A3(tend, i, k, j) + inc;
      }
    }
}

/* horizontal_diffusion_v_2 (module_diffusion_em.py:2535-2765):
 * v points i in [2, ide-2], j in [2, jde-1]. */
static void hdiff_v(OCtx *x, real *t1, real *t2) {
  real *tend = FLD(rv_tendf), *zx = FLD(zx), *zy = FLD(zy), *rdzw = FLD(rdzw), *fnm = FLD(fnm), *fnp = FLD(fnp), *dnw = FLD(dnw);
  real *msfvx = FLD(msfvx), *msfvy = FLD(msfvy);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy), g = (real)x->c.g;
  int ktf = KDE - 1;
  titau_12(x, t1, FLD(xkmh), FLD(defor12), 0, 1, 0, 0);
  titau_11(x, t2, FLD(xkmh), FLD(defor22), 0, 0, 1, 0);
  real a1[512], a2[512], zxv[512], zyv[512];
  for (int j = 2; j <= JDE - 1; j++)
    for (int i = 2; i <= IDE - 2; i++) {
      for (int k = 1; k <= ktf; k++) {
        if (k >= 2) {
          a1[k] = (real)0.5 * (A1(fnm, k) * (A3(t1, i + 1, k, j) + A3(t1, i, k, j)) + A1(fnp, k) * (A3(t1, i + 1, k - 1, j) + A3(t1, i, k - 1, j)));
          a2[k] = (real)0.5 * (A1(fnm, k) * (A3(t2, i, k, j - 1) + A3(t2, i, k, j)) + A1(fnp, k) * (A3(t2, i, k - 1, j - 1) + A3(t2, i, k - 1, j)));
        }
        zxv[k] = (real)0.125 * (A3(zx, i, k, j) + A3(zx, i + 1, k, j) + A3(zx, i, k, j - 1) + A3(zx, i + 1, k, j - 1) + A3(zx, i, k + 1, j) +
                                A3(zx, i + 1, k + 1, j) + A3(zx, i, k + 1, j - 1) + A3(zx, i + 1, k + 1, j - 1));
        zyv[k] = (real)0.5 * (A3(zy, i, k, j) + A3(zy, i, k + 1, j));
      }
      a1[1] = 0; a1[ktf + 1] = 0; a2[1] = 0; a2[ktf + 1] = 0;
      for (int k = 1; k <= ktf; k++) {
        real mrdx = A2(msfvx, i, j) * rdx, mrdy = A2(msfvy, i, j) * rdy;
        real tmpdz = ((real)1 / A3(rdzw, i, k, j) + (real)1 / A3(rdzw, i, k, j - 1)) / (real)2;
        real inc = g * tmpdz / A1(dnw, k) *
                   (mrdy * (A3(t2, i, k, j) - A3(t2, i, k, j - 1)) + mrdx * (A3(t1, i + 1, k, j) - A3(t1, i, k, j)) -
                    A2(msfvx, i, j) * zxv[k] * (a1[k + 1] - a1[k]) / tmpdz - A2(msfvy, i, j) * zyv[k] * (a2[k + 1] - a2[k]) / tmpdz);
        A3(tend, i, k, j) = A3(tend, i, k, j) + inc;
      }
    }
}

/* horizontal_diffusion_w_2 (module_diffusion_em.py:2768-2985):
 * w points i in [2, ide-2], j in [2, jde-2], k in [2, kde-1]. */
static void hdiff_w(OCtx *x, real *t1, real *t2) {
  real *tend = FLD(rw_tendf), *zx = FLD(zx), *zy = FLD(zy), *rdz = FLD(rdz), *dn = FLD(dn);
  real *msftx = FLD(msftx), *msfty = FLD(msfty);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy), g = (real)x->c.g;
  int ktf = KDE - 1;
  titau_13(x, t1, FLD(xkmv), FLD(defor13), 0, 1, 0, 0);
  titau_23(x, t2, FLD(xkmv), FLD(defor23), 0, 0, 0, 1);
  real a1[512], a2[512];
  for (int j = 2; j <= JDE - 2; j++)
    for (int i = 2; i <= IDE - 2; i++) {
      for (int k = 1; k <= ktf; k++) {
        a1[k] = (real)0.25 * (A3(t1, i + 1, k + 1, j) + A3(t1, i, k + 1, j) + A3(t1, i + 1, k, j) + A3(t1, i, k, j));
        a2[k] = (real)0.25 * (A3(t2, i, k + 1, j + 1) + A3(t2, i, k + 1, j) + A3(t2, i, k, j + 1) + A3(t2, i, k, j));
      }
      for (int k = 2; k <= ktf; k++) {
        real zxw = (real)0.5 * (A3(zx, i, k, j) + A3(zx, i + 1, k, j));
        real zyw = (real)0.5 * (A3(zy, i, k, j) + A3(zy, i, k, j + 1));
        real mrdx = A2(msftx, i, j) * rdx, mrdy = A2(msfty, i, j) * rdy;
        real inc = g / (A1(dn, k) * A3(rdz, i, k, j)) *
                   (mrdx * (A3(t1, i + 1, k, j) - A3(t1, i, k, j)) + mrdy * (A3(t2, i, k, j + 1) - A3(t2, i, k, j)) -
                    A2(msfty, i, j) * A3(rdz, i, k, j) * (zxw * (a1[k] - a1[k - 1]) + zyw * (a2[k] - a2[k - 1])));
        A3(tend, i, k, j) = A3(tend, i, k, j) + inc;
      }
    }
}

/* horizontal_diffusion_s (module_diffusion_em.py:2046-2291): scalar
 * diffusion with xkhh on mass points i in [2, ide-2], j in [2, jde-2]. */
void o_hdiff_s(OCtx *x, real *tend, const real *var, real *h1, real *h2) {
  real *xkhh = FLD(xkhh), *rho = FLD(rho), *zx = FLD(zx), *zy = FLD(zy), *rdzw = FLD(rdzw), *fnm = FLD(fnm), *fnp = FLD(fnp);
  real *dnw = FLD(dnw), *dn = FLD(dn), *msfux = FLD(msfux), *msfvy = FLD(msfvy), *msftx = FLD(msftx), *msfty = FLD(msfty);
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy), g = (real)x->c.g;
  real cf1 = (real)x->c.cf1, cf2 = (real)x->c.cf2, cf3 = (real)x->c.cf3;
  int kte = KDE, ktf = KDE - 1, ktes1 = kte - 1, ktes2 = kte - 2;
  int i0 = 2, i1 = IDE - 2, j0 = 2, j1 = JDE - 2;
  real xkx[512], hav[512];
  /* H1 on u faces i in [i0, i1+1], rows j in [j0, j1] */
  for (int j = j0; j <= j1; j++)
    for (int i = i0; i <= i1 + 1; i++) {
      for (int k = 2; k <= ktf; k++)
        hav[k] = (real)0.5 * (A1(fnm, k) * (A3(var, i - 1, k, j) + A3(var, i, k, j)) + A1(fnp, k) * (A3(var, i - 1, k - 1, j) + A3(var, i, k - 1, j)));
      hav[1] = (real)0.5 * (cf1 * A3(var, i, 1, j) + cf2 * A3(var, i, 2, j) + cf3 * A3(var, i, 3, j) + cf1 * A3(var, i - 1, 1, j) +
                            cf2 * A3(var, i - 1, 2, j) + cf3 * A3(var, i - 1, 3, j));
      hav[ktf + 1] = (real)0.5 * (A3(var, i, ktes1, j) + (A3(var, i, ktes1, j) - A3(var, i, ktes2, j)) * (real)0.5 * A1(dnw, ktes1) / A1(dn, ktes1) +
                                  A3(var, i - 1, ktes1, j) + (A3(var, i - 1, ktes1, j) - A3(var, i - 1, ktes2, j)) * (real)0.5 * A1(dnw, ktes1) / A1(dn, ktes1));
      for (int k = 1; k <= ktf; k++) {
        xkx[k] = (real)0.5 * (A3(xkhh, i - 1, k, j) + A3(xkhh, i, k, j)) * (real)0.5 * (A3(rho, i - 1, k, j) + A3(rho, i, k, j));
        real tmpzx = (real)0.5 * (A3(zx, i, k, j) + A3(zx, i, k + 1, j));
        real rdzu = (real)2 / ((real)1 / A3(rdzw, i, k, j) + (real)1 / A3(rdzw, i - 1, k, j));
        A3(h1, i, k, j) = -A2(msfux, i, j) * xkx[k] * (rdx * (A3(var, i, k, j) - A3(var, i - 1, k, j)) - tmpzx * (hav[k + 1] - hav[k]) * rdzu);
      }
    }
  /* H2 on v faces j in [j0, j1+1], columns i in [i0, i1] */
  for (int j = j0; j <= j1 + 1; j++)
    for (int i = i0; i <= i1; i++) {
      for (int k = 2; k <= ktf; k++)
        hav[k] = (real)0.5 * (A1(fnm, k) * (A3(var, i, k, j - 1) + A3(var, i, k, j)) + A1(fnp, k) * (A3(var, i, k - 1, j - 1) + A3(var, i, k - 1, j)));
      hav[1] = (real)0.5 * (cf1 * A3(var, i, 1, j) + cf2 * A3(var, i, 2, j) + cf3 * A3(var, i, 3, j) + cf1 * A3(var, i, 1, j - 1) +
                            cf2 * A3(var, i, 2, j - 1) + cf3 * A3(var, i, 3, j - 1));
      hav[ktf + 1] = (real)0.5 * (A3(var, i, ktes1, j) + (A3(var, i, ktes1, j) - A3(var, i, ktes2, j)) * (real)0.5 * A1(dnw, ktes1) / A1(dn, ktes1) +
                                  A3(var, i, ktes1, j - 1) + (A3(var, i, ktes1, j - 1) - A3(var, i, ktes2, j - 1)) * (real)0.5 * A1(dnw, ktes1) / A1(dn, ktes1));
      for (int k = 1; k <= ktf; k++) {
        xkx[k] = (real)0.5 * (A3(xkhh, i, k, j - 1) + A3(xkhh, i, k, j)) * (real)0.5 * (A3(rho, i, k, j - 1) + A3(rho, i, k, j));
        real tmpzy = (real)0.5 * (A3(zy, i, k, j) + A3(zy, i, k + 1, j));
        real rdzv = (real)2 / ((real)1 / A3(rdzw, i, k, j) + (real)1 / A3(rdzw, i, k, j - 1));
        A3(h2, i, k, j) = -A2(msfvy, i, j) * xkx[k] * (rdy * (A3(var, i, k, j) - A3(var, i, k, j - 1)) - tmpzy * (hav[k + 1] - hav[k]) * rdzv);
      }
    }
  real h1a[512], h2a[512];
  for (int j = j0; j <= j1; j++)
    for (int i = i0; i <= i1; i++) {
      for (int k = 2; k <= ktf; k++) {
        h1a[k] = (real)0.5 * (A1(fnm, k) * (A3(h1, i + 1, k, j) + A3(h1, i, k, j)) + A1(fnp, k) * (A3(h1, i + 1, k - 1, j) + A3(h1, i, k - 1, j)));
        h2a[k] = (real)0.5 * (A1(fnm, k) * (A3(h2, i, k, j + 1) + A3(h2, i, k, j)) + A1(fnp, k) * (A3(h2, i, k - 1, j + 1) + A3(h2, i, k - 1, j)));
      }
      h1a[1] = 0; h1a[ktf + 1] = 0; h2a[1] = 0; h2a[ktf + 1] = 0;
      for (int k = 1; k <= ktf; k++) {
        real zxm = (real)0.25 * (A3(zx, i, k, j) + A3(zx, i + 1, k, j) + A3(zx, i, k + 1, j) + A3(zx, i + 1, k + 1, j));
        real zym = (real)0.25 * (A3(zy, i, k, j) + A3(zy, i, k, j + 1) + A3(zy, i, k + 1, j) + A3(zy, i, k + 1, j + 1));
        real mrdx = A2(msftx, i, j) * rdx, mrdy = A2(msfty, i, j) * rdy;
        real inc = g / (A1(dnw, k) * A3(rdzw, i, k, j)) *
                   (mrdx * (A3(h1, i + 1, k, j) - A3(h1, i, k, j)) + mrdy * (A3(h2, i, k, j + 1) - A3(h2, i, k, j)) -
                    A2(msftx, i, j) * zxm * (h1a[k + 1] - h1a[k]) * A3(rdzw, i, k, j) - A2(msfty, i, j) * zym * (h2a[k + 1] - h2a[k]) * A3(rdzw, i, k, j));
        A3(tend, i, k, j) = A3(tend, i, k, j) + inc;
      }
    }
}

/* horizontal_diffusion_2 driver (module_diffusion_em.py:2988-3100):
 * u, v, w momentum, theta (t_2) and every moist / scalar species. */
void o_horizontal_diffusion_2(OCtx *x) {
  real *t1 = (real *)calloc(x->n3, sizeof(real)), *t2 = (real *)calloc(x->n3, sizeof(real));
  hdiff_u(x, t1, t2);
  hdiff_v(x, t1, t2);
  hdiff_w(x, t1, t2);
  o_hdiff_s(x, FLD(t_tendf), FLD(t_2), t1, t2);
  for (int s = 0; s < x->c.nmoist; s++) o_hdiff_s(x, SPC(moist_tend, s), SPC(moist, s), t1, t2);
  for (int s = 0; s < x->c.nscalar; s++) o_hdiff_s(x, SPC(scalar_tend, s), SPC(scalar, s), t1, t2);
  free(t1);
  free(t2);
}
