// Boundary kernels: physical-boundary copies, specified zone, relaxation
// zone, flow-dependent scalar boundary, final specified values, surface w.
// Semantics follow src/wrf_jax/boundary/ and src/wrf_jax/dyn_em/module_bc_em.py
// for specified lateral boundaries with spec_zone = 1.
#pragma once
#include "wrfdc_common.h"

// ---- physical BC copy (set_physical_bc3d, no-halo specified) -------------
// plane i = ide-1 -> ide for every j (skipped for x-staggered variables)
__global__ void k_bc_x(const DevCtx xv, float *fl, int kend, int is2d) {
  const DevCtx *__restrict__ x = &xv;
  const int j = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  if (j > x->jde || k > kend) return;
  if (is2d) A2(fl, x->ide, j) = A2(fl, x->ide - 1, j);
  else A3(fl, x->ide, k, j) = A3(fl, x->ide - 1, k, j);
}
// row j = jde-1 -> jde for every i (skipped for y-staggered variables)
__global__ void k_bc_y(const DevCtx xv, float *fl, int kend, int is2d) {
  const DevCtx *__restrict__ x = &xv;
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  if (i > x->ide || k > kend) return;
  if (is2d) A2(fl, i, x->jde) = A2(fl, i, x->jde - 1);
  else A3(fl, i, k, x->jde) = A3(fl, i, k, x->jde - 1);
}

// Batched physical BC copies: up to kBcMax fields per launch, blockIdx.z =
// field.  Column ide and row jde are written in one pass from their interior
// neighbours with the sequential result of "plane, then row": the corner
// (ide, jde) takes (ide-1, jde-1) when both copies apply.
constexpr int kBcMax = 24;
struct BcList {
  float *f[kBcMax];
  int kend[kBcMax];
  int flags[kBcMax];  // bit 0: copy column ide, bit 1: copy row jde, bit 2: 2-D
};
__global__ void k_bc_batch(const DevCtx xv, BcList L) {
  const DevCtx *__restrict__ x = &xv;
  const int a = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  const int q = (int)blockIdx.z;
  const int fl = L.flags[q];
  if (k > L.kend[q]) return;
  float *f = L.f[q];
  const int ide = x->ide, jde = x->jde;
  const bool cx = fl & 1, cy = fl & 2, is2d = fl & 4;
#define AT(ii, jj) (is2d ? A2(f, ii, jj) : A3(f, ii, k, jj))
  if (cx && a <= jde) {
    if (a == jde && cy) AT(ide, jde) = AT(ide - 1, jde - 1);
    else AT(ide, a) = AT(ide - 1, a);
  }
  if (cy && a <= ide && !(a == ide && cx)) AT(a, jde) = AT(a, jde - 1);
#undef AT
}

// Batched fills: blockIdx.y = field, grid-stride over its elements.
constexpr int kFillMax = 16;
struct FillList {
  float *p[kFillMax];
  size_t n[kFillMax];
};
__global__ void k_fill_batch(FillList L) {
  float *p = L.p[blockIdx.y];
  const size_t n = L.n[blockIdx.y];
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = 0.f;
}

// ---- specified zone (spec_zone = 1): ring of the variable's grid ----------
// zone point t -> (i, j, edge) with edge 0=south 1=north 2=west 3=east.
// south/north rows span i in [1, ibe]; west/east columns j in [2, jbe-1].
__device__ __forceinline__ bool zone_point(int t, int ibe, int jbe, int &i, int &j, int &edge) {
  const int ni = ibe, nj = jbe - 2;
  if (t < ni) { i = 1 + t; j = 1; edge = 0; return true; }
  t -= ni;
  if (t < ni) { i = 1 + t; j = jbe; edge = 1; return true; }
  t -= ni;
  if (t < nj) { i = 1; j = 2 + t; edge = 2; return true; }
  t -= nj;
  if (t < nj) { i = ibe; j = 2 + t; edge = 3; return true; }
  return false;
}
__device__ __forceinline__ float bval(const DevCtx *__restrict__ x, const float *const b4[4], int e, int a, int k, int b, int is2d) {
  if (is2d) return (e < 2) ? BY2(b4[e], a, b) : BX2(b4[e], a, b);
  return (e < 2) ? BY(b4[e], a, k, b) : BX(b4[e], a, k, b);
}
struct BPtrs {
  const float *e[4];  // ys, ye, xs, xe
};

// spec_bdytend overlay: tend on the zone = boundary tendency (b = 1)
__global__ void k_spec_overlay(const DevCtx xv, float *tend, int ibe, int jbe, int kt, int is2d, BPtrs bt) {
  const DevCtx *__restrict__ x = &xv;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  int i, j, e;
  if (k > kt || !zone_point(t, ibe, jbe, i, j, e)) return;
  const int a = (e < 2) ? i : j;
  float v = bval(x, bt.e, e, a, k, 1, is2d);
  if (is2d) A2(tend, i, j) = v;
  else A3(tend, i, k, j) = v;
}

// several overlays in one launch (blockIdx.z selects the field): the five
// dry tendencies of one stage
constexpr int kOverlayMax = 5;
struct OverlayList {
  float *tend[kOverlayMax];
  BPtrs bt[kOverlayMax];
  int ibe[kOverlayMax], jbe[kOverlayMax], kt[kOverlayMax], is2d[kOverlayMax];
};
__global__ void k_spec_overlay_batch(const DevCtx xv, OverlayList L) {
  const DevCtx *__restrict__ x = &xv;
  const int f = (int)blockIdx.z;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  int i, j, e;
  if (k > L.kt[f] || !zone_point(t, L.ibe[f], L.jbe[f], i, j, e)) return;
  const int a = (e < 2) ? i : j;
  const int is2d = L.is2d[f];
  const float v = bval(x, L.bt[f].e, e, a, k, 1, is2d);
  if (is2d) A2(L.tend[f], i, j) = v;
  else A3(L.tend[f], i, k, j) = v;
}

// spec_bdyupdate: field += dt * tend on the zone
__global__ void k_spec_update(const DevCtx xv, float *fl, const float *tend, float dt, int ibe, int jbe, int kt, int is2d) {
  const DevCtx *__restrict__ x = &xv;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  int i, j, e;
  if (k > kt || !zone_point(t, ibe, jbe, i, j, e)) return;
  if (is2d) A2(fl, i, j) = A2(fl, i, j) + dt * A2(tend, i, j);
  else A3(fl, i, k, j) = A3(fl, i, k, j) + dt * A3(tend, i, k, j);
}

// spec_bdyupdate_ph (module_bc_em.py:1429-1560)
__global__ void k_spec_update_ph(const DevCtx xv, float dts) {
  const DevCtx *__restrict__ x = &xv;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  int i, j, e;
  if (k > x->kde || !zone_point(t, x->ide - 1, x->jde - 1, i, j, e)) return;
  const float *muts = FL(muts), *mt = FL(mu_tend), *c1 = FL(c1f), *c2 = FL(c2f);
  float *ph = FL(ph_2);
  float mts = A2(muts, i, j);
  float mu_old = mts - dts * A2(mt, i, j);
  float den = A1(c1, k) * mts + A1(c2, k);
  float ratio = (A1(c1, k) * mu_old + A1(c2, k)) / den;
  A3(ph, i, k, j) = A3(ph, i, k, j) * ratio + dts * A3(FL(ph_tend), i, k, j) / den + A3(FL(ph_save), i, k, j) * (ratio - 1.f);
}

// zero_grad_bdy for w (zero_grad_bdy.py:21-128): rows first, then columns
__global__ void k_zero_grad_rows(const DevCtx xv, float *w) {
  const DevCtx *__restrict__ x = &xv;
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  const int side = (int)blockIdx.z;
  const int ibe = x->ide - 1, jbe = x->jde - 1;
  if (i > ibe || k > x->kde) return;
  const int ii = imin_(imax_(i, 2), ibe - 1);
  if (side == 0) A3(w, i, k, 1) = A3(w, ii, k, 2);
  else A3(w, i, k, jbe) = A3(w, ii, k, jbe - 1);
}
__global__ void k_zero_grad_cols(const DevCtx xv, float *w) {
  const DevCtx *__restrict__ x = &xv;
  const int j = 2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  const int side = (int)blockIdx.z;
  const int ibe = x->ide - 1, jbe = x->jde - 1;
  if (j > jbe - 1 || k > x->kde) return;
  const int jj = imin_(imax_(j, 2), jbe - 1);
  if (side == 0) A3(w, 1, k, j) = A3(w, 2, k, jj);
  else A3(w, ibe, k, j) = A3(w, ibe - 1, k, jj);
}

// ---- relaxation zone (relax_bdytend_core.py:27-294) -----------------------
// side 0..3 = south, north, west, east; bd = 1 .. rz-1 rows into the zone.
// fld is the mass-coupled field, tend receives the increment.
// the relaxed field at one point: the field itself (mw 0), or mass_weight's
// coupled value (module_bc_em.py:561-624) formed on the spot, the field times
// c1*mut + c2 (mw 1: half levels, c1h/c2h, k <= kde-1; mw 2: w levels,
// c1f/c2f), zero outside the coupled range
__device__ __forceinline__ float relax_fv(const DevCtx *__restrict__ x, const float *fld, int ii, int k, int jj, int is2d, int mw) {
  if (is2d) return A2(fld, ii, jj);
  const float f = A3(fld, ii, k, jj);
  if (mw == 0) return f;
  const bool in = (ii <= x->ide - 1 && jj <= x->jde - 1 && (mw == 2 || k <= x->kde - 1));
  const float *c1 = mw == 2 ? FL(c1f) : FL(c1h), *c2 = mw == 2 ? FL(c2f) : FL(c2h);
  return in ? f * (A1(c1, k) * A2(FL(mut), ii, jj) + A1(c2, k)) : 0.f;
}
// the fields of one launch: blockIdx.z = field * 4*(rz-1) + side*(rz-1) + bd-1
constexpr int kRelaxMax = 5;
struct RelaxList {
  const float *fld[kRelaxMax];
  float *tend[kRelaxMax];
  BPtrs b[kRelaxMax], bt[kRelaxMax];
  int ibe[kRelaxMax], jbe[kRelaxMax], kt[kRelaxMax], is2d[kRelaxMax], mw[kRelaxMax];
};
// dt_bc_p: the time since the lateral-boundary update, read from device memory
// (set before each step, so that a captured step reads the current value)
__global__ void k_relax(const DevCtx xv, RelaxList R, const float *__restrict__ dt_bc_p) {
  const DevCtx *__restrict__ x = &xv;
  const float dt_bc = *dt_bc_p;
  const int a = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  const int rz = x->relax_zone;
  const int zs = 4 * (rz - 1);
  const int fi = (int)blockIdx.z / zs, zz = (int)blockIdx.z % zs;
  const float *fld = R.fld[fi];
  float *tend = R.tend[fi];
  const int ibe = R.ibe[fi], jbe = R.jbe[fi], kt = R.kt[fi], is2d = R.is2d[fi], mw = R.mw[fi];
  const BPtrs &b = R.b[fi], &bt = R.bt[fi];
  const int side = zz / (rz - 1);
  const int bd = 1 + zz % (rz - 1);
  if (k > kt) return;
  int i, j, iin, iout, jin, jout, along;
  if (side < 2) {
    j = side == 0 ? 1 + bd : jbe - bd;
    i = a;
    if (i < 1 + bd || i > ibe - bd) return;
    jin = side == 0 ? j - 1 : j + 1;
    jout = side == 0 ? j + 1 : j - 1;
    iin = iout = i;
    along = i;
  } else {
    i = side == 2 ? 1 + bd : ibe - bd;
    j = a;
    if (j < bd + 2 || j > jbe - bd - 1) return;
    iin = side == 2 ? i - 1 : i + 1;
    iout = side == 2 ? i + 1 : i - 1;
    jin = jout = j;
    along = j;
  }
  (void)along;
#define FV(ii_, jj_) relax_fv(x, fld, ii_, k, jj_, is2d, mw)
#define BD(e_, a_, bb_) (bval(x, b.e, e_, a_, k, bb_, is2d) + dt_bc * bval(x, bt.e, e_, a_, k, bb_, is2d))
  float f0, f1, f2, f3, f4;
  if (side < 2) {
    const int im = imax_(i - 1, 1), ipp = imin_(i + 1, ibe);
    f0 = BD(side, i, bd + 1) - FV(i, j);
    f1 = BD(side, im, bd + 1) - FV(im, j);
    f2 = BD(side, ipp, bd + 1) - FV(ipp, j);
    f3 = BD(side, i, bd) - FV(i, jin);
    f4 = BD(side, i, bd + 2) - FV(i, jout);
  } else {
    f0 = BD(side, j, bd + 1) - FV(i, j);
    f1 = BD(side, j - 1, bd + 1) - FV(i, j - 1);
    f2 = BD(side, j + 1, bd + 1) - FV(i, j + 1);
    f3 = BD(side, j, bd) - FV(iin, j);
    f4 = BD(side, j, bd + 2) - FV(iout, j);
  }
#undef FV
#undef BD
  const float inc = x->fcx[bd + 1] * f0 - x->gcx[bd + 1] * (f1 + f2 + f3 + f4 - 4.f * f0);
  if (is2d) A2(tend, i, j) += inc;
  else A3(tend, i, k, j) += inc;
}

// flow_dep_bdy (flow_dep_bdy.py:27-176): rows then columns
__global__ void k_flow_dep_rows(const DevCtx xv, float *q, const float *v) {
  const DevCtx *__restrict__ x = &xv;
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  const int side = (int)blockIdx.z;
  const int ibe = x->ide - 1, jbe = x->jde - 1;
  if (i > ibe || k > x->kde - 1) return;
  const int ii = imin_(imax_(i, 2), ibe - 1);
  if (side == 0) A3(q, i, k, 1) = A3(v, i, k, 1) < 0.f ? A3(q, ii, k, 2) : 0.f;
  else A3(q, i, k, jbe) = A3(v, i, k, jbe + 1) > 0.f ? A3(q, ii, k, jbe - 1) : 0.f;
}
__global__ void k_flow_dep_cols(const DevCtx xv, float *q, const float *u) {
  const DevCtx *__restrict__ x = &xv;
  const int j = 2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  const int side = (int)blockIdx.z;
  const int ibe = x->ide - 1, jbe = x->jde - 1;
  if (j > jbe - 1 || k > x->kde - 1) return;
  const int jj = imin_(imax_(j, 2), jbe - 1);
  if (side == 0) A3(q, 1, k, j) = A3(u, 1, k, j) < 0.f ? A3(q, 2, k, jj) : 0.f;
  else A3(q, ibe, k, j) = A3(u, ibe + 1, k, j) > 0.f ? A3(q, ibe - 1, k, jj) : 0.f;
}

// spec_bdy_final (spec_bdy_final.py:27-191)
// the fields of one launch (blockIdx.z selects the field)
constexpr int kFinalMax = 6;
struct FinalList {
  float *fl[kFinalMax];
  const float *mu[kFinalMax], *msf[kFinalMax], *c1[kFinalMax], *c2[kFinalMax];
  BPtrs b[kFinalMax], bt[kFinalMax];
  int ibe[kFinalMax], jbe[kFinalMax], kt[kFinalMax], is2d[kFinalMax], msfc[kFinalMax], muc[kFinalMax];
};
__global__ void k_spec_final(const DevCtx xv, FinalList L, const float *__restrict__ dt_bc_p) {
  const DevCtx *__restrict__ x = &xv;
  const float dt_bc = *dt_bc_p;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  const int f = (int)blockIdx.z;
  float *fl = L.fl[f];
  const float *mu = L.mu[f], *msf = L.msf[f], *c1 = L.c1[f], *c2 = L.c2[f];
  const BPtrs &b = L.b[f], &bt = L.bt[f];
  const int ibe = L.ibe[f], jbe = L.jbe[f], kt = L.kt[f], is2d = L.is2d[f], msfc = L.msfc[f], muc = L.muc[f];
  int i, j, e;
  if (k > kt || !zone_point(t, ibe, jbe, i, j, e)) return;
  const int a = (e < 2) ? i : j;
  float v = bval(x, b.e, e, a, k, 1, is2d) + dt_bc * bval(x, bt.e, e, a, k, 1, is2d);
  if (msfc) v = v * A2(msf, i, j);
  if (muc) v = v / (A1(c1, k) * A2(mu, i, j) + A1(c2, k));
  if (is2d) A2(fl, i, j) = v;
  else A3(fl, i, k, j) = v;
}

// set_w_surface (module_bc_em.py:448-560) without the fill option
__global__ void k_set_w_surface(const DevCtx xv) {
  const DevCtx *__restrict__ x = &xv;
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = 1 + (int)blockIdx.y;
  if (i > x->ide - 1 || j > x->jde - 1) return;
  const float *u = FL(u_2), *v = FL(v_2), *ht = FL(ht);
  const float cf1 = x->cf1, cf2 = x->cf2, cf3 = x->cf3;
  const int im = imax_(i - 1, 1), ipp = imin_(i + 1, x->ide - 1);
  const int jm = imax_(j - 1, 1), jp = imin_(j + 1, x->jde - 1);
  float vjp = cf1 * A3(v, i, 1, jp) + cf2 * A3(v, i, 2, jp) + cf3 * A3(v, i, 3, jp);
  float vj = cf1 * A3(v, i, 1, j) + cf2 * A3(v, i, 2, j) + cf3 * A3(v, i, 3, j);
  float uip = cf1 * A3(u, ipp, 1, j) + cf2 * A3(u, ipp, 2, j) + cf3 * A3(u, ipp, 3, j);
  float ui = cf1 * A3(u, i, 1, j) + cf2 * A3(u, i, 2, j) + cf3 * A3(u, i, 3, j);
  float yp = A2(FL(msfty), i, j) * 0.5f * x->rdy * ((A2(ht, i, jp) - A2(ht, i, j)) * vjp + (A2(ht, i, j) - A2(ht, i, jm)) * vj);
  float xp = A2(FL(msftx), i, j) * 0.5f * x->rdx * ((A2(ht, ipp, j) - A2(ht, i, j)) * uip + (A2(ht, i, j) - A2(ht, im, j)) * ui);
  A3(FL(w_2), i, 1, j) = yp + xp;
}
