/*
 * Oracle: lateral / physical boundary routines (test infrastructure only).
 * Follows src/wrf_jax/boundary/ and src/wrf_jax/dyn_em/module_bc_em.py.
 */
#include "o_core.h"

/* Staggering-dependent edge limits used by every specified-zone routine:
 * spec_bdytend.py:24-116 / spec_bdyupdate.py:24-103 / zero_grad_bdy.py:44-60. */
typedef struct { int ibs, ibe, jbs, jbe, ktf; } ZDims;

static ZDims zdims(const OCtx *x, char var, int ktop) {
  ZDims d;
  d.ibs = 1; d.ibe = IDE - 1; d.jbs = 1; d.jbe = JDE - 1; d.ktf = KDE - 1;
  if (var == 'u') d.ibe = IDE;
  if (var == 'v') d.jbe = JDE;
  if (ktop) d.ktf = KDE;
  return d;
}

/* set_physical_bc3d / set_physical_bc2d for specified boundaries without
 * halo: physical_bc.py:79-129 (_set_physical_bc3d_no_halo_specified) with the
 * variable flags of physical_bc.py:171-212.  Copies plane ide-1 -> ide
 * (non x-staggered) then row jde-1 -> jde (non y-staggered). */
void o_bc3(OCtx *x, real *fl, char var) {
  int full_k = (var == 'e' || var == 'f' || var == 'w');
  int kend = full_k ? KDE : KDE - 1;
  int xst = (var == 'u' || var == 'x');
  int yst = (var == 'v' || var == 'y');
  if (!xst)
    for (int j = 1; j <= JDE; j++)
      for (int k = 1; k <= kend; k++) A3(fl, IDE, k, j) = A3(fl, IDE - 1, k, j);
  if (!yst)
    for (int k = 1; k <= kend; k++)
      for (int i = 1; i <= IDE; i++) A3(fl, i, k, JDE) = A3(fl, i, k, JDE - 1);
}

void o_bc2(OCtx *x, real *fl, char var) {
  int xst = (var == 'u' || var == 'x');
  int yst = (var == 'v' || var == 'y');
  if (!xst)
    for (int j = 1; j <= JDE; j++) A2(fl, IDE, j) = A2(fl, IDE - 1, j);
  if (!yst)
    for (int i = 1; i <= IDE; i++) A2(fl, i, JDE) = A2(fl, i, JDE - 1);
}

/* Value of a lateral-boundary array at zone position (edge e: 0=ys 1=ye
 * 2=xs 3=xe), along-edge index a, level k, width index b (1-based). */
static real bval(const OCtx *x, real *const b4[4], int e, int a, int k, int b, int is2d) {
  if (is2d) return (e < 2) ? BY2(b4[e], a, b) : BX2(b4[e], a, b);
  return (e < 2) ? BY(b4[e], a, k, b) : BX(b4[e], a, k, b);
}

/* spec_bdytend overlay: the specified-zone points of `tend` take the
 * boundary tendency (spec_bdytend.py:117-196, overlay runtime.py:6656). */
void o_spec_overlay(OCtx *x, real *tend, char var, int ktop, int is2d,
                    real *bt_xs, real *bt_xe, real *bt_ys, real *bt_ye) {
  real *const bt[4] = {bt_ys, bt_ye, bt_xs, bt_xe};
  ZDims d = zdims(x, var, ktop);
  int sz = x->c.spec_zone;
  int kt = is2d ? 1 : d.ktf;
#define TSET(i, k, j, val) do { if (is2d) A2(tend, i, j) = (val); else A3(tend, i, k, j) = (val); } while (0)
  for (int j = d.jbs; j <= imin(d.jbe, d.jbs + sz - 1); j++) {
    int bd = j - d.jbs;
    for (int k = 1; k <= kt; k++)
      for (int i = d.ibs + bd; i <= d.ibe - bd; i++) TSET(i, k, j, bval(x, bt, 0, i, k, bd + 1, is2d));
  }
  for (int j = imax(1, d.jbe - sz + 1); j <= d.jbe; j++) {
    int bd = d.jbe - j;
    for (int k = 1; k <= kt; k++)
      for (int i = d.ibs + bd; i <= d.ibe - bd; i++) TSET(i, k, j, bval(x, bt, 1, i, k, bd + 1, is2d));
  }
  for (int i = d.ibs; i <= imin(d.ibe, d.ibs + sz - 1); i++) {
    int bd = i - d.ibs;
    for (int k = 1; k <= kt; k++)
      for (int j = bd + d.jbs + 1; j <= d.jbe - bd - 1; j++) TSET(i, k, j, bval(x, bt, 2, j, k, bd + 1, is2d));
  }
  for (int i = imax(1, d.ibe - sz + 1); i <= d.ibe; i++) {
    int bd = d.ibe - i;
    for (int k = 1; k <= kt; k++)
      for (int j = bd + d.jbs + 1; j <= d.jbe - bd - 1; j++) TSET(i, k, j, bval(x, bt, 3, j, k, bd + 1, is2d));
  }
#undef TSET
}

/* Visit every specified-zone point (same zone as spec_bdyupdate.py:24-103). */
#define ZONE_LOOP(d, sz, kt, BODY)                                            \
  do {                                                                        \
    for (int j = (d).jbs; j <= imin((d).jbe, (d).jbs + (sz) - 1); j++) {      \
      int bd = j - (d).jbs;                                                   \
      for (int k = 1; k <= (kt); k++)                                         \
        for (int i = (d).ibs + bd; i <= (d).ibe - bd; i++) { BODY; }          \
    }                                                                         \
    for (int j = imax(1, (d).jbe - (sz) + 1); j <= (d).jbe; j++) {            \
      int bd = (d).jbe - j;                                                   \
      for (int k = 1; k <= (kt); k++)                                         \
        for (int i = (d).ibs + bd; i <= (d).ibe - bd; i++) { BODY; }          \
    }                                                                         \
    for (int i = (d).ibs; i <= imin((d).ibe, (d).ibs + (sz) - 1); i++) {      \
      int bd = i - (d).ibs;                                                   \
      for (int k = 1; k <= (kt); k++)                                         \
        for (int j = bd + (d).jbs + 1; j <= (d).jbe - bd - 1; j++) { BODY; }  \
    }                                                                         \
    for (int i = imax(1, (d).ibe - (sz) + 1); i <= (d).ibe; i++) {            \
      int bd = (d).ibe - i;                                                   \
      for (int k = 1; k <= (kt); k++)                                         \
        for (int j = bd + (d).jbs + 1; j <= (d).jbe - bd - 1; j++) { BODY; }  \
    }                                                                         \
  } while (0)

/* spec_bdyupdate: field += dt * tend on the specified zone
 * (spec_bdyupdate.py:104-180). */
void o_spec_bdyupdate(OCtx *x, real *fl, const real *tend, double dt, char var, int is2d) {
  ZDims d = zdims(x, var, var == 'm' || var == 'h');
  int kt = is2d ? 1 : d.ktf;
  real rdt = (real)dt;
  if (is2d) {
    ZONE_LOOP(d, x->c.spec_zone, kt, { (void)bd; A2(fl, i, j) = A2(fl, i, j) + rdt * A2(tend, i, j); });
  } else {
    ZONE_LOOP(d, x->c.spec_zone, kt, { (void)bd; A3(fl, i, k, j) = A3(fl, i, k, j) + rdt * A3(tend, i, k, j); });
  }
}

/* spec_bdyupdate_ph (module_bc_em.py:1429-1560): geopotential on the
 * specified zone, re-weighted by the column-mass change. */
void o_spec_bdyupdate_ph(OCtx *x, double dts) {
  real *ph = FLD(ph_2), *tend = FLD(ph_tend), *phs = FLD(ph_save);
  real *muts = FLD(muts), *mut_t = FLD(mu_tend), *c1 = FLD(c1f), *c2 = FLD(c2f);
  ZDims d = zdims(x, 'h', 1);
  real rdt = (real)dts;
  ZONE_LOOP(d, x->c.spec_zone, d.ktf, {
    (void)bd;
    real mts = A2(muts, i, j);
    real mu_old = mts - rdt * A2(mut_t, i, j);
    real den = A1(c1, k) * mts + A1(c2, k);
    real ratio = (A1(c1, k) * mu_old + A1(c2, k)) / den;
    A3(ph, i, k, j) = A3(ph, i, k, j) * ratio + rdt * A3(tend, i, k, j) / den + A3(phs, i, k, j) * (ratio - (real)1);
  });
}

/* zero_grad_bdy for w (zero_grad_bdy.py:21-128): copy the first interior
 * row/column (clipped along the edge) into the specified zone. */
void o_zero_grad_bdy_w(OCtx *x, real *w) {
  int sz = x->c.spec_zone;
  int ibs = 1, ibe = IDE - 1, jbs = 1, jbe = JDE - 1, ktf = KDE;
  for (int j = jbs; j <= imin(jbe, jbs + sz - 1); j++) {
    int bd = j - jbs;
    for (int i = ibs + bd; i <= ibe - bd; i++) {
      int ii = imin(imax(i, ibs + sz), ibe - sz);
      for (int k = 1; k <= ktf; k++) A3(w, i, k, j) = A3(w, ii, k, jbs + sz);
    }
  }
  for (int j = imax(1, jbe - sz + 1); j <= jbe; j++) {
    int bd = jbe - j;
    for (int i = ibs + bd; i <= ibe - bd; i++) {
      int ii = imin(imax(i, ibs + sz), ibe - sz);
      for (int k = 1; k <= ktf; k++) A3(w, i, k, j) = A3(w, ii, k, jbe - sz);
    }
  }
  for (int i = ibs; i <= imin(ibe, ibs + sz - 1); i++) {
    int bd = i - ibs;
    for (int j = bd + jbs + 1; j <= jbe - bd - 1; j++) {
      int jj = imin(imax(j, jbs + sz), jbe - sz);
      for (int k = 1; k <= ktf; k++) A3(w, i, k, j) = A3(w, ibs + sz, k, jj);
    }
  }
  for (int i = imax(1, ibe - sz + 1); i <= ibe; i++) {
    int bd = ibe - i;
    for (int j = bd + jbs + 1; j <= jbe - bd - 1; j++) {
      int jj = imin(imax(j, jbs + sz), jbe - sz);
      for (int k = 1; k <= ktf; k++) A3(w, i, k, j) = A3(w, ibe - sz, k, jj);
    }
  }
}

/* relax_bdytend_core (boundary/relax_bdytend_core.py:27-294): relaxation
 * tendency fcx*(bdy-fld) - gcx*Laplacian(bdy-fld) on the relaxation zone.
 * `fld` is the (mass-coupled) field, `tend` receives the increment. */
void o_relax_bdytend(OCtx *x, const real *fld, real *tend, char var, int ktop, int is2d,
                     real *b_xs, real *b_xe, real *b_ys, real *b_ye,
                     real *bt_xs, real *bt_xe, real *bt_ys, real *bt_ye) {
  real *const b[4] = {b_ys, b_ye, b_xs, b_xe};
  real *const bt[4] = {bt_ys, bt_ye, bt_xs, bt_xe};
  ZDims d = zdims(x, var, ktop);
  int sz = x->c.spec_zone, rz = x->c.relax_zone;
  int kt = is2d ? 1 : d.ktf;
  real dtbc = (real)x->c.dtbc;
#define FV(i, k, j) (is2d ? A2(fld, i, j) : A3(fld, i, k, j))
#define BD(e, a, k, bb) (bval(x, b, e, a, k, bb, is2d) + dtbc * bval(x, bt, e, a, k, bb, is2d))
#define TADD(i, k, j, v) do { if (is2d) A2(tend, i, j) += (v); else A3(tend, i, k, j) += (v); } while (0)
  /* south / north rows */
  for (int side = 0; side < 2; side++) {
    int j0 = side == 0 ? imax(1, d.jbs + sz) : imax(1, d.jbe - rz + 1);
    int j1 = side == 0 ? imin(d.jbe, d.jbs + rz - 1) : d.jbe - sz;
    for (int j = j0; j <= j1; j++) {
      int bd = side == 0 ? j - d.jbs : d.jbe - j;
      int jin = side == 0 ? j - 1 : j + 1;   /* neighbour toward the edge */
      int jout = side == 0 ? j + 1 : j - 1;
      for (int k = 1; k <= kt; k++)
        for (int i = d.ibs + bd; i <= d.ibe - bd; i++) {
          int im = imax(i - 1, d.ibs), ipp = imin(i + 1, d.ibe);
          real f0 = BD(side, i, k, bd + 1) - FV(i, k, j);
          real f1 = BD(side, im, k, bd + 1) - FV(im, k, j);
          real f2 = BD(side, ipp, k, bd + 1) - FV(ipp, k, j);
          real f3 = BD(side, i, k, bd) - FV(i, k, jin);
          real f4 = BD(side, i, k, bd + 2) - FV(i, k, jout);
          TADD(i, k, j, x->fcx[bd + 1] * f0 - x->gcx[bd + 1] * (f1 + f2 + f3 + f4 - (real)4 * f0));
        }
    }
  }
  /* west / east columns */
  for (int side = 2; side < 4; side++) {
    int i0 = side == 2 ? d.ibs + sz : imax(1, d.ibe - rz + 1);
    int i1 = side == 2 ? imin(d.ibe, d.ibs + rz - 1) : d.ibe - sz;
    for (int i = i0; i <= i1; i++) {
      int bd = side == 2 ? i - d.ibs : d.ibe - i;
      int iin = side == 2 ? i - 1 : i + 1;
      int iout = side == 2 ? i + 1 : i - 1;
      for (int k = 1; k <= kt; k++)
        for (int j = bd + d.jbs + 1; j <= d.jbe - bd - 1; j++) {
          real f0 = BD(side, j, k, bd + 1) - FV(i, k, j);
          real f1 = BD(side, j - 1, k, bd + 1) - FV(i, k, j - 1);
          real f2 = BD(side, j + 1, k, bd + 1) - FV(i, k, j + 1);
          real f3 = BD(side, j, k, bd) - FV(iin, k, j);
          real f4 = BD(side, j, k, bd + 2) - FV(iout, k, j);
          TADD(i, k, j, x->fcx[bd + 1] * f0 - x->gcx[bd + 1] * (f1 + f2 + f3 + f4 - (real)4 * f0));
        }
    }
  }
#undef FV
#undef BD
#undef TADD
}

/* flow_dep_bdy (boundary/flow_dep_bdy.py:27-176): inflow zero, outflow copy
 * of the first interior value, for unstaggered scalars. */
void o_flow_dep_bdy(OCtx *x, real *q, const real *u, const real *v) {
  int sz = x->c.spec_zone;
  int ibs = 1, ibe = IDE - 1, jbs = 1, jbe = JDE - 1, ktf = KDE - 1;
  for (int j = jbs; j <= imin(jbe, jbs + sz - 1); j++) {
    int bd = j - jbs;
    for (int i = ibs + bd; i <= ibe - bd; i++) {
      int ii = imin(imax(i, ibs + sz), ibe - sz);
      for (int k = 1; k <= ktf; k++)
        A3(q, i, k, j) = A3(v, i, k, j) < (real)0 ? A3(q, ii, k, jbs + sz) : (real)0;
    }
  }
  for (int j = imax(1, jbe - sz + 1); j <= jbe; j++) {
    int bd = jbe - j;
    for (int i = ibs + bd; i <= ibe - bd; i++) {
      int ii = imin(imax(i, ibs + sz), ibe - sz);
      for (int k = 1; k <= ktf; k++)
        A3(q, i, k, j) = A3(v, i, k, j + 1) > (real)0 ? A3(q, ii, k, jbe - sz) : (real)0;
    }
  }
  for (int i = ibs; i <= imin(ibe, ibs + sz - 1); i++) {
    int bd = i - ibs;
    for (int j = bd + jbs + 1; j <= jbe - bd - 1; j++) {
      int jj = imin(imax(j, jbs + sz), jbe - sz);
      for (int k = 1; k <= ktf; k++)
        A3(q, i, k, j) = A3(u, i, k, j) < (real)0 ? A3(q, ibs + sz, k, jj) : (real)0;
    }
  }
  for (int i = imax(1, ibe - sz + 1); i <= ibe; i++) {
    int bd = ibe - i;
    for (int j = bd + jbs + 1; j <= jbe - bd - 1; j++) {
      int jj = imin(imax(j, jbs + sz), jbe - sz);
      for (int k = 1; k <= ktf; k++)
        A3(q, i, k, j) = A3(u, i + 1, k, j) > (real)0 ? A3(q, ibe - sz, k, jj) : (real)0;
    }
  }
}

/* spec_bdy_final (boundary/spec_bdy_final.py:27-191): specified-zone values
 * from the boundary arrays, decoupled by map factor and column mass. */
void o_spec_bdy_final(OCtx *x, real *fl, char var, int is2d, const real *mu, const real *msf,
                      const real *c1, const real *c2,
                      real *b_xs, real *b_xe, real *b_ys, real *b_ye,
                      real *bt_xs, real *bt_xe, real *bt_ys, real *bt_ye) {
  real *const b[4] = {b_ys, b_ye, b_xs, b_xe};
  real *const bt[4] = {bt_ys, bt_ye, bt_xs, bt_xe};
  ZDims d = zdims(x, var, var == 'm' || var == 'h' || var == 'w');
  int msfc = (var == 'u' || var == 'v' || var == 'w');
  int muc = (var != 'm');
  int kt = is2d ? 1 : d.ktf;
  int sz = x->c.spec_zone;
  real dtbc = (real)x->c.dtbc;
  for (int e = 0; e < 4; e++) {
    if (e < 2) {
      int j0 = e == 0 ? d.jbs : imax(1, d.jbe - sz + 1);
      int j1 = e == 0 ? imin(d.jbe, d.jbs + sz - 1) : d.jbe;
      for (int j = j0; j <= j1; j++) {
        int bd = e == 0 ? j - d.jbs : d.jbe - j;
        for (int k = 1; k <= kt; k++)
          for (int i = d.ibs + bd; i <= d.ibe - bd; i++) {
            real v = bval(x, b, e, i, k, bd + 1, is2d) + dtbc * bval(x, bt, e, i, k, bd + 1, is2d);
            if (msfc) v = v * A2(msf, i, j);
            if (muc) v = v / (A1(c1, k) * A2(mu, i, j) + A1(c2, k));
            if (is2d) A2(fl, i, j) = v; else A3(fl, i, k, j) = v;
          }
      }
    } else {
      int i0 = e == 2 ? d.ibs : imax(1, d.ibe - sz + 1);
      int i1 = e == 2 ? imin(d.ibe, d.ibs + sz - 1) : d.ibe;
      for (int i = i0; i <= i1; i++) {
        int bd = e == 2 ? i - d.ibs : d.ibe - i;
        for (int k = 1; k <= kt; k++)
          for (int j = bd + d.jbs + 1; j <= d.jbe - bd - 1; j++) {
            real v = bval(x, b, e, j, k, bd + 1, is2d) + dtbc * bval(x, bt, e, j, k, bd + 1, is2d);
            if (msfc) v = v * A2(msf, i, j);
            if (muc) v = v / (A1(c1, k) * A2(mu, i, j) + A1(c2, k));
            if (is2d) A2(fl, i, j) = v; else A3(fl, i, k, j) = v;
          }
      }
    }
  }
}

/* set_w_surface (module_bc_em.py:448-560) without the fill option:
 * terrain-following lower-boundary w from the extrapolated near-surface
 * winds (neighbour indices clipped to the domain). */
void o_set_w_surface(OCtx *x, real *w, const real *u, const real *v) {
  real *ht = FLD(ht), *msftx = FLD(msftx), *msfty = FLD(msfty);
  real cf1 = (real)x->c.cf1, cf2 = (real)x->c.cf2, cf3 = (real)x->c.cf3;
  real rdx = (real)(1.0 / x->c.dx), rdy = (real)(1.0 / x->c.dy);
  for (int j = 1; j <= JDE - 1; j++)
    for (int i = 1; i <= IDE - 1; i++) {
      int im = imax(i - 1, 1), ipp = imin(i + 1, IDE - 1);
      int jm = imax(j - 1, 1), jp = imin(j + 1, JDE - 1);
      real vjp = cf1 * A3(v, i, 1, jp) + cf2 * A3(v, i, 2, jp) + cf3 * A3(v, i, 3, jp);
      real vj = cf1 * A3(v, i, 1, j) + cf2 * A3(v, i, 2, j) + cf3 * A3(v, i, 3, j);
      real uip = cf1 * A3(u, ipp, 1, j) + cf2 * A3(u, ipp, 2, j) + cf3 * A3(u, ipp, 3, j);
      real ui = cf1 * A3(u, i, 1, j) + cf2 * A3(u, i, 2, j) + cf3 * A3(u, i, 3, j);
      real yp = A2(msfty, i, j) * (real)0.5 * rdy * ((A2(ht, i, jp) - A2(ht, i, j)) * vjp + (A2(ht, i, j) - A2(ht, i, jm)) * vj);
      real xp = A2(msftx, i, j) * (real)0.5 * rdx * ((A2(ht, ipp, j) - A2(ht, i, j)) * uip + (A2(ht, i, j) - A2(ht, im, j)) * ui);
      A3(w, i, 1, j) = yp + xp;
    }
}
