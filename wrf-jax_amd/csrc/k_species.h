// Fused species stage: every moist and scalar species of one RK stage in one
// pass (moist_scalar_advance / scalar_advance, src/wrf_jax/runtime.py:
// 2715-3117, 13900-14830, with rk_update_scalar, module_em.py:1437-1598).
//
// Per mass point the mass fluxes ru_m / rv_m / ww_m of the six faces are
// loaded once; each species then evaluates its 5th/3rd-order face fluxes,
// the advective tendency and the mass-weighted RK update.  Inputs are read
// from one buffer and results written to another (the caller alternates the
// species buffers between stages), so no thread reads a value another
// thread of the same launch writes.
//
// Positive-definite advection on the last stage runs as two passes over all
// PD species: the renormalisation factor of every limited cell, then the
// tendency from the limited faces (each face scaled by the cell its
// antidiffusive flux leaves) followed by the update.  Face fluxes are
// recomputed in each pass instead of being stored.
#pragma once
#include "wrfdc_common.h"
#include "k_advect.h"

constexpr int kMaxSpec = 16;
struct SpecList {
  const float *in[kMaxSpec];  // stencil source (time level being advected)
  float *out[kMaxSpec];       // updated species
  float *old[kMaxSpec];       // stage-start species (q_old)
  float *tend[kMaxSpec];      // accumulated physics / boundary tendency
  float *scale[kMaxSpec];     // PD renormalisation factors (scratch)
  int n;
};

// scalar face fluxes with the face velocity given (same arithmetic as
// fy_s / fx_s / fz_s of k_advect.h)
__device__ __forceinline__ float sfy(const DevCtx *__restrict__ x, const float *q, float vel, int i, int k, int j) {
  const int jde = x->jde;
  if (j >= 4 && j <= jde - 3)
    return vel * flux5(A3(q, i, k, j - 3), A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), A3(q, i, k, j + 2), vel);
  if (j == 2 || j == jde - 1) return 0.5f * vel * (A3(q, i, k, j) + A3(q, i, k, j - 1));
  return vel * flux3(A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), vel);
}
__device__ __forceinline__ float sfx(const DevCtx *__restrict__ x, const float *q, float vel, int i, int k, int j) {
  const int ide = x->ide;
  if (i >= 4 && i <= ide - 3)
    return vel * flux5(A3(q, i - 3, k, j), A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), A3(q, i + 2, k, j), vel);
  if (i == 2 || i == ide - 1) return 0.5f * vel * (A3(q, i, k, j) + A3(q, i - 1, k, j));
  return vel * flux3(A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), vel);
}
__device__ __forceinline__ float sfz(const DevCtx *__restrict__ x, const float *q, float vel, int i, int k, int j) {
  const int ktf = x->kde - 1;
  if (k <= 1 || k >= x->kde) return 0.f;
  if (k == 2 || k == ktf) return vel * (A1(FL(fnm), k) * A3(q, i, k, j) + A1(FL(fnp), k) * A3(q, i, k - 1, j));
  return vel * flux3(A3(q, i, k - 2, j), A3(q, i, k - 1, j), A3(q, i, k, j), A3(q, i, k + 1, j), -vel);
}

// mass-weighted RK update of one point (rk_update_scalar)
struct UpdCoef {
  float oldm, newm, msfty;
  bool inner;
};
__device__ __forceinline__ UpdCoef upd_coef(const DevCtx *__restrict__ x, int i, int k, int j) {
  UpdCoef c;
  const int sz = x->spec_zone;
  c.inner = (i >= 1 + sz && i <= x->ide - sz - 1 && j >= 1 + sz && j <= x->jde - sz - 1);
  const float mub = A2(FL(mub), i, j);
  const float muold = A2(FL(mu_1), i, j) + mub, munew = A2(FL(mu_2), i, j) + mub;
  const float c1 = A1(FL(c1h), k), c2 = A1(FL(c2h), k);
  c.oldm = c1 * muold + c2;
  c.newm = c1 * munew + c2;
  c.msfty = A2(FL(msfty), i, j);
  return c;
}

// plain (non-PD) stages: advection with the time-averaged mass fluxes and
// the update, all species of the list.  At rk = 1 `in` is the stage-start
// copy q_old (equal to q there), so old = in.
__global__ void __launch_bounds__(256) k_species_update(const DevCtx xv, SpecList L, int rk, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *ru = FL(ru_m), *rv = FL(rv_m), *rom = FL(ww_m);
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const bool doy = (j >= 2 && j <= jde - 2), dox = (i >= 2 && i <= ide - 2);
  const float vn = doy ? A3(rv, i, k, j + 1) : 0.f, vs = doy ? A3(rv, i, k, j) : 0.f;
  const float ue = dox ? A3(ru, i + 1, k, j) : 0.f, uw = dox ? A3(ru, i, k, j) : 0.f;
  const float wt = (k + 1 <= kde - 1) ? A3(rom, i, k + 1, j) : 0.f, wb = (k >= 2) ? A3(rom, i, k, j) : 0.f;
  const float msftx = A2(FL(msftx), i, j), rdzw = A1(FL(rdnw), k);
  const UpdCoef c = upd_coef(x, i, k, j);
  const size_t o = XI3(i, k, j);
  for (int s = 0; s < L.n; s++) {
    const float *q = L.in[s];
    float adv = 0.f;
    if (doy) adv -= msftx * x->rdy * (sfy(x, q, vn, i, k, j + 1) - sfy(x, q, vs, i, k, j));
    if (dox) adv -= msftx * x->rdx * (sfx(x, q, ue, i + 1, k, j) - sfx(x, q, uw, i, k, j));
    adv -= rdzw * (sfz(x, q, wt, i, k + 1, j) - sfz(x, q, wb, i, k, j));
    float tend = c.inner ? adv * c.msfty : 0.f;
    tend = tend + L.tend[s][o];
    const float old = (rk == 1) ? q[o] : L.old[s][o];
    L.out[s][o] = (c.oldm * old + dt * tend) / c.newm;
  }
}

// rk_update_scalar_pd (module_em.py:1601-1669) for every species of the list:
// last-stage physics update of q_old with mu_old = mu_new = mu_1, tendency
// cleared.
__global__ void k_species_pd_prep(const DevCtx xv, SpecList L, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const size_t o = XI3(i, k, j);
  const float m = A1(FL(c1h), k) * (A2(FL(mu_1), i, j) + A2(FL(mub), i, j)) + A1(FL(c2h), k);
  for (int s = 0; s < L.n; s++) {
    L.old[s][o] = (m * L.old[s][o] + dt * L.tend[s][o]) / m;
    L.tend[s][o] = 0.f;
  }
}

// ---- positive-definite faces (same arithmetic as k_pd_fluxes) ------------
// antidiffusive (high - low) and upwind (low) flux of the y face j
__device__ __forceinline__ void pd_face_y(const DevCtx *__restrict__ x, const float *q, const float *qo, int i, int k, int j, float dt,
                                          float &anti, float &low) {
  anti = 0.f;
  low = 0.f;
  const int jde = x->jde;
  if (!(j >= 2 && j <= jde - 1 && i <= x->ide - 1 && k <= x->kde - 1)) return;
  const float *c1 = FL(c1h), *c2 = FL(c2h), *msftx = FL(msftx), *mut = FL(muts);
  const float dy = 2.f / (A2(msftx, i, j) + A2(msftx, i, j - 1)) / x->rdy;
  const float mu = 0.5f * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i, j - 1) + A1(c2, k)));
  const float vel = A3(FL(rv_m), i, k, j);
  const float cr = vel * dt / dy / mu;
  low = mu * (dy / dt) * flux_up(A3(qo, i, k, j - 1), A3(qo, i, k, j), cr);
  float high;
  if (j >= 4 && j <= jde - 3)
    high = vel * flux5pd(A3(q, i, k, j - 3), A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), A3(q, i, k, j + 2), vel);
  else if (j == 2 || j == jde - 1)
    high = 0.5f * vel * (A3(q, i, k, j) + A3(q, i, k, j - 1));
  else
    high = vel * flux3pd(A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), vel);
  anti = high - low;
}
__device__ __forceinline__ void pd_face_x(const DevCtx *__restrict__ x, const float *q, const float *qo, int i, int k, int j, float dt,
                                          float &anti, float &low) {
  anti = 0.f;
  low = 0.f;
  const int ide = x->ide;
  if (!(i >= 2 && i <= ide - 1 && j <= x->jde - 1 && k <= x->kde - 1)) return;
  const float *c1 = FL(c1h), *c2 = FL(c2h), *msfty = FL(msfty), *mut = FL(muts);
  const float dx = 2.f / (A2(msfty, i, j) + A2(msfty, i - 1, j)) / x->rdx;
  const float mu = 0.5f * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i - 1, j) + A1(c2, k)));
  const float vel = A3(FL(ru_m), i, k, j);
  const float cr = vel * dt / dx / mu;
  low = mu * (dx / dt) * flux_up(A3(qo, i - 1, k, j), A3(qo, i, k, j), cr);
  float high;
  if (i >= 4 && i <= ide - 3)
    high = vel * flux5pd(A3(q, i - 3, k, j), A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), A3(q, i + 2, k, j), vel);
  else if (i == 2 || i == ide - 1)
    high = 0.5f * vel * (A3(q, i, k, j) + A3(q, i - 1, k, j));
  else
    high = vel * flux3pd(A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), vel);
  anti = high - low;
}
__device__ __forceinline__ void pd_face_z(const DevCtx *__restrict__ x, const float *q, const float *qo, int i, int k, int j, float dt,
                                          float &anti, float &low) {
  anti = 0.f;
  low = 0.f;
  const int ktf = x->kde - 1;
  if (!(k >= 2 && k <= ktf && i <= x->ide - 1 && j <= x->jde - 1)) return;
  const float *c1 = FL(c1h), *c2 = FL(c2h), *rdzw = FL(rdnw), *mut = FL(muts);
  const float dz = 2.f / (A1(rdzw, k) + A1(rdzw, k - 1));
  const float mu = A1(c1, k) * A2(mut, i, j) + A1(c2, k);
  const float vel = A3(FL(ww_m), i, k, j);
  const float cr = vel * dt / dz / mu;
  low = mu * (dz / dt) * flux_up(A3(qo, i, k - 1, j), A3(qo, i, k, j), cr);
  float high;
  if (k == 2 || k == ktf) high = vel * (A1(FL(fnm), k) * A3(q, i, k, j) + A1(FL(fnp), k) * A3(q, i, k - 1, j));
  else high = vel * flux3pd(A3(q, i, k - 2, j), A3(q, i, k - 1, j), A3(q, i, k, j), A3(q, i, k + 1, j), -vel);
  anti = high - low;
}

// pass 1: renormalisation factor of every cell (1 where not limited);
// full extent so the factor array is defined wherever pass 2 reads it.
__global__ void __launch_bounds__(256) k_species_pd_scale(const DevCtx xv, SpecList L, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const size_t o = XI3(i, k, j);
  const bool lim = (i >= 2 && i <= x->ide - 2 && j >= 2 && j <= x->jde - 2 && k <= x->kde - 1);
  if (!lim) {
    for (int s = 0; s < L.n; s++) L.scale[s][o] = 1.f;
    return;
  }
  const float *c1 = FL(c1h), *c2 = FL(c2h), *rdzw = FL(rdnw);
  const float rdx = x->rdx, rdy = x->rdy;
  const float mm = A2(FL(msftx), i, j) * A2(FL(msfty), i, j), msfty = A2(FL(msfty), i, j);
  const float mold = (A1(c1, k) * A2(FL(mub), i, j) + A1(c2, k)) + (A1(c1, k) * A2(FL(mu_1), i, j));
  for (int s = 0; s < L.n; s++) {
    const float *q = L.in[s], *qo = L.old[s];
    float ax0, lx0, ax1, lx1, ay0, ly0, ay1, ly1, az0, lz0, az1, lz1;
    pd_face_x(x, q, qo, i, k, j, dt, ax0, lx0);
    pd_face_x(x, q, qo, i + 1, k, j, dt, ax1, lx1);
    pd_face_y(x, q, qo, i, k, j, dt, ay0, ly0);
    pd_face_y(x, q, qo, i, k, j + 1, dt, ay1, ly1);
    pd_face_z(x, q, qo, i, k, j, dt, az0, lz0);
    pd_face_z(x, q, qo, i, k + 1, j, dt, az1, lz1);
    const float ph_low = mold * A3(qo, i, k, j) - dt * (mm * (rdx * (lx1 - lx0) + rdy * (ly1 - ly0)) + msfty * A1(rdzw, k) * (lz1 - lz0));
    const float flux_out = dt * (mm * (rdx * (fmaxf(0.f, ax1) - fminf(0.f, ax0)) + rdy * (fmaxf(0.f, ay1) - fminf(0.f, ay0))) +
                                 msfty * A1(rdzw, k) * (fminf(0.f, az1) - fmaxf(0.f, az0)));
    float sc = 1.f;
    if (flux_out > ph_low) sc = fmaxf(0.f, ph_low / (flux_out + 1.0e-20f));
    L.scale[s][o] = sc;
  }
}

// pass 2: tendency from the renormalised faces, then the mass-weighted update
// (old = q_old after k_species_pd_prep; the cleared physics tendency adds 0).
__global__ void __launch_bounds__(256) k_species_pd_update(const DevCtx xv, SpecList L, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const size_t o = XI3(i, k, j);
  const float rdx = x->rdx, rdy = x->rdy, rdzw = A1(FL(rdnw), k), msftx = A2(FL(msftx), i, j);
  const bool dox = (i >= 2 && i <= x->ide - 2), doy = (j >= 2 && j <= x->jde - 2);
  const UpdCoef c = upd_coef(x, i, k, j);
  for (int s = 0; s < L.n; s++) {
    const float *q = L.in[s], *qo = L.old[s], *sc = L.scale[s];
    float az0, lz0, az1, lz1;
    pd_face_z(x, q, qo, i, k, j, dt, az0, lz0);
    pd_face_z(x, q, qo, i, k + 1, j, dt, az1, lz1);
    float zt = az1, zb = az0;
    if (zt < 0.f) zt *= A3(sc, i, k, j);
    else if (zt > 0.f && k + 1 <= x->kde - 1) zt *= A3(sc, i, k + 1, j);
    if (zb > 0.f) zb *= A3(sc, i, k, j);
    else if (zb < 0.f && k - 1 >= 1) zb *= A3(sc, i, k - 1, j);
    float t = 0.f;
    t += -rdzw * (zt - zb + lz1 - lz0);
    if (dox) {
      float ax0, lx0, ax1, lx1;
      pd_face_x(x, q, qo, i, k, j, dt, ax0, lx0);
      pd_face_x(x, q, qo, i + 1, k, j, dt, ax1, lx1);
      float xe = ax1, xw = ax0;
      if (xe > 0.f) xe *= A3(sc, i, k, j);
      else if (xe < 0.f) xe *= A3(sc, i + 1, k, j);
      if (xw < 0.f) xw *= A3(sc, i, k, j);
      else if (xw > 0.f) xw *= A3(sc, i - 1, k, j);
      t += -msftx * (rdx * (xe - xw + lx1 - lx0));
    }
    if (doy) {
      float ay0, ly0, ay1, ly1;
      pd_face_y(x, q, qo, i, k, j, dt, ay0, ly0);
      pd_face_y(x, q, qo, i, k, j + 1, dt, ay1, ly1);
      float yn = ay1, ys = ay0;
      if (yn > 0.f) yn *= A3(sc, i, k, j);
      else if (yn < 0.f) yn *= A3(sc, i, k, j + 1);
      if (ys < 0.f) ys *= A3(sc, i, k, j);
      else if (ys > 0.f) ys *= A3(sc, i, k, j - 1);
      t += -msftx * (rdy * (yn - ys + ly1 - ly0));
    }
    float tend = c.inner ? t * c.msfty : 0.f;
    tend = tend + L.tend[s][o];
    L.out[s][o] = (c.oldm * qo[o] + dt * tend) / c.newm;
  }
}

// flow_dep_bdy (flow_dep_bdy.py:27-176) of several species in one launch:
// inflow 0, outflow copies the first interior value; every destination (the
// zone ring) reads an interior point, so rows and columns go together.
struct FlowList {
  float *q[kMaxSpec];
  int n;
};
__global__ void k_flow_dep_batch(const DevCtx xv, FlowList L) {
  const DevCtx *__restrict__ x = &xv;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  const int s = (int)blockIdx.z;
  const int ibe = x->ide - 1, jbe = x->jde - 1;
  int i, j, e;
  if (k > x->kde - 1 || !zone_point(t, ibe, jbe, i, j, e)) return;
  float *q = L.q[s];
  const float *u = FL(ru_m), *v = FL(rv_m);
  if (e == 0) {
    const int ii = imin_(imax_(i, 2), ibe - 1);
    A3(q, i, k, 1) = A3(v, i, k, 1) < 0.f ? A3(q, ii, k, 2) : 0.f;
  } else if (e == 1) {
    const int ii = imin_(imax_(i, 2), ibe - 1);
    A3(q, i, k, jbe) = A3(v, i, k, jbe + 1) > 0.f ? A3(q, ii, k, jbe - 1) : 0.f;
  } else if (e == 2) {
    const int jj = imin_(imax_(j, 2), jbe - 1);
    A3(q, 1, k, j) = A3(u, 1, k, j) < 0.f ? A3(q, 2, k, jj) : 0.f;
  } else {
    const int jj = imin_(imax_(j, 2), jbe - 1);
    A3(q, ibe, k, j) = A3(u, ibe + 1, k, j) > 0.f ? A3(q, ibe - 1, k, jj) : 0.f;
  }
}
