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
// antidiffusive flux leaves) followed by the update.  Both passes form the
// face fluxes from the species themselves with the same arithmetic; only the
// factor (one value per cell) goes through memory between them.
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
  // species whose last-stage physics update of q_old (rk_update_scalar_pd)
  // the second stage's update also does, with the last stage's dt: bit s
  // set for species s (k_species_update at rk = 2 only)
  unsigned prep_mask;
  float prep_dt;
};

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
// copy q_old (equal to q there), so old = in.  Face-sharing blocks (block
// (64, kFaceRows), 1-D grid of face_tiles(r).count() blocks): the south faces of all
// species go to LDS first, one barrier, then each species' update.
//
// At rk = 2 the species of L.prep_mask also get the last stage's
// rk_update_scalar_pd here (k_species_pd_prep's arithmetic, over the same
// points): q_old and the tendency it reads are not written again before the
// last stage, and c.oldm is that kernel's column mass (c1h (mu_1 + mub) +
// c2h), so the last stage finds the same q_old and a cleared tendency.  Each
// thread reads q_old and the tendency at its own point only, so updating them
// in place is race-free.
__global__ void __launch_bounds__(64 * kFaceRows) k_species_update(const DevCtx xv, SpecList L, int rk, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  extern __shared__ float ys_[];  // L.n x kFaceRows x 64: south face flux of every species
  GET_FACE_BLOCK(r);
#define YS(s, w) ys_[((s) * kFaceRows + (w)) * 64 + lane]
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const float *ru = FL(ru_m), *rv = FL(rv_m), *rom = FL(ww_m);
  const unsigned nb = (unsigned)x->n3, sj = (unsigned)(kde * x->ip), sk = (unsigned)x->ip;
  {
    // rows past the range still form their south face (the row below reads
    // it); its window may fall outside the field, the face is then 0
    const bool yf = (j >= 2 && j <= jde - 1);
    const float vs = yf ? A3(rv, ic, k, j) : 0.f;
    const unsigned os = XI3(ic, k, j);
    for (int s = 0; s < L.n; s++) {
      float w[6];
      win6b(bufv(L.in[s], nb), os, sj, w);
      pin6(w);
      YS(s, wv) = yf ? hflux_br(w, vs, j, jde) : 0.f;
    }
  }
  __syncthreads();
  if (!jrow) return;
  const bool xface = (i >= 2 && i <= ide - 1);
  const bool yface_n = (wv == kFaceRows - 1 && j + 1 <= jde - 1);
  const bool doy = (j >= 2 && j <= jde - 2), dox = (ic >= 2 && ic <= ide - 2);
  const float uw = xface ? A3(ru, i, k, j) : 0.f;
  const float vn = yface_n ? A3(rv, ic, k, j + 1) : 0.f;
  const float wt = (k + 1 <= kde - 1) ? A3(rom, ic, k + 1, j) : 0.f, wb = (k >= 2) ? A3(rom, ic, k, j) : 0.f;
  const float msftx = A2(FL(msftx), ic, j), rdzw = A1(FL(rdnw), k);
  const UpdCoef c = upd_coef(x, ic, k, j);
  const unsigned o = XI3(ic, k, j), ox = XI3(i, k, j);
  const int k1 = imin_(k + 1, kde);
  const float fm0 = A1(FL(fnm), k), fp0 = A1(FL(fnp), k), fm1 = A1(FL(fnm), k1), fp1 = A1(FL(fnp), k1);
  for (int s = 0; s < L.n; s++) {
    const BufF bq = bufv(L.in[s], nb);
    // every load of this species first: x stencil of the west face, the
    // column window k-2 .. k+2 of both vertical faces, tendency and old
    // value; the top wave also the stencil of its north face
    float xw[6], yw[6], zw[5];
    win6b(bq, ox, 1u, xw);
    win5z(bq, o, sk, zw);
    const float tnd = elem(L.tend[s], o);
    const float old = elem(rk == 1 ? L.in[s] : L.old[s], o);
    if (wv == kFaceRows - 1) {
      win6b(bq, o + sj, sj, yw);
      pin6(yw);
    }
    pin6(xw);
    pin5(zw);
    // fluxes
    const float fw = xface ? hflux_br(xw, uw, i, ide) : 0.f;
    const float fe = lane_next(fw, lane);
    const float fs = YS(s, wv);
    float fn;
    if (wv < kFaceRows - 1) fn = YS(s, wv + 1);
    else fn = yface_n ? hflux_br(yw, vn, j + 1, jde) : 0.f;
    float adv = 0.f;
    if (doy) adv -= msftx * x->rdy * (fn - fs);
    if (dox) adv -= msftx * x->rdx * (fe - fw);
    adv -= rdzw * (zflux_br(zw[1], zw[2], zw[3], zw[4], wt, k + 1, kde, fm1, fp1) - zflux_br(zw[0], zw[1], zw[2], zw[3], wb, k, kde, fm0, fp0));
    float tend = c.inner ? adv * c.msfty : 0.f;
    tend = tend + tnd;
    const float out = (c.oldm * old + dt * tend) / c.newm;
    if (own) {
      elem(L.out[s], o) = out;
      if (rk == 2 && ((L.prep_mask >> s) & 1u)) {
        elem(L.old[s], o) = (c.oldm * old + L.prep_dt * tnd) / c.oldm;
        elem(L.tend[s], o) = 0.f;
      }
    }
  }
#undef YS
}

// rk_update_scalar_pd (module_em.py:1601-1669) for every species of the list:
// last-stage physics update of q_old with mu_old = mu_new = mu_1, tendency
// cleared.
__global__ void k_species_pd_prep(const DevCtx xv, SpecList L, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const size_t o = XI3(i, k, j);
  const float m = A1(FL(c1h), k) * (A2(FL(mu_1), i, j) + A2(FL(mub), i, j)) + A1(FL(c2h), k);
  for (int s = 0; s < L.n; s++) {
    L.old[s][o] = (m * L.old[s][o] + dt * L.tend[s][o]) / m;
    L.tend[s][o] = 0.f;
  }
}

// ---- positive-definite faces (same arithmetic as k_pd_fluxes) ------------
// Species-independent part of one face: velocity, Courant number, the
// upwind-flux weight mu*(d/dt) and the order of the high-order flux
// (5: 5th order, 3: 3rd order, 2: centred, 0: no face).
struct PdFace {
  float vel, cr, muf;
  int mode;
};
__device__ __forceinline__ PdFace pd_face_geom(const DevCtx *__restrict__ x, int axis, int i, int k, int j, float dt) {
  PdFace f{0.f, 0.f, 0.f, 0};
  const float *c1 = FL(c1h), *c2 = FL(c2h), *mut = FL(muts);
  const int ide = x->ide, jde = x->jde, ktf = x->kde - 1;
  if (axis == 0) {
    if (!(i >= 2 && i <= ide - 1 && j <= jde - 1 && k <= ktf)) return f;
    const float dx = 2.f / (A2(FL(msfty), i, j) + A2(FL(msfty), i - 1, j)) / x->rdx;
    const float mu = 0.5f * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i - 1, j) + A1(c2, k)));
    f.vel = A3(FL(ru_m), i, k, j);
    f.cr = f.vel * dt / dx / mu;
    f.muf = mu * (dx / dt);
    f.mode = (i >= 4 && i <= ide - 3) ? 5 : ((i == 2 || i == ide - 1) ? 2 : 3);
  } else if (axis == 1) {
    if (!(j >= 2 && j <= jde - 1 && i <= ide - 1 && k <= ktf)) return f;
    const float dy = 2.f / (A2(FL(msftx), i, j) + A2(FL(msftx), i, j - 1)) / x->rdy;
    const float mu = 0.5f * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i, j - 1) + A1(c2, k)));
    f.vel = A3(FL(rv_m), i, k, j);
    f.cr = f.vel * dt / dy / mu;
    f.muf = mu * (dy / dt);
    f.mode = (j >= 4 && j <= jde - 3) ? 5 : ((j == 2 || j == jde - 1) ? 2 : 3);
  } else {
    if (!(k >= 2 && k <= ktf && i <= ide - 1 && j <= jde - 1)) return f;
    const float *rdzw = FL(rdnw);
    const float dz = 2.f / (A1(rdzw, k) + A1(rdzw, k - 1));
    const float mu = A1(c1, k) * A2(mut, i, j) + A1(c2, k);
    f.vel = A3(FL(ww_m), i, k, j);
    f.cr = f.vel * dt / dz / mu;
    f.muf = mu * (dz / dt);
    f.mode = (k == 2 || k == ktf) ? 2 : 3;
  }
  return f;
}
// Antidiffusive and upwind flux of one species through a face from values
// already in registers (every load of a species is issued and pinned first,
// then the order branches).  Horizontal: w = q at slots -3 .. +2 of the
// face, om1 / o0 = q_old on either side.
__device__ __forceinline__ void pd_hflux_br(const PdFace &f, const float w[6], float om1, float o0, float &anti, float &low) {
  if (f.mode == 0) {
    anti = 0.f;
    low = 0.f;
    return;
  }
  low = f.muf * flux_up(om1, o0, f.cr);
  float high;
  if (f.mode == 5) high = f.vel * flux5pd(w[0], w[1], w[2], w[3], w[4], w[5], f.vel);
  else if (f.mode == 2) high = 0.5f * f.vel * (w[3] + w[2]);
  else high = f.vel * flux3pd(w[1], w[2], w[3], w[4], f.vel);
  anti = high - low;
}
// vertical face at level kk: q(kk-2 .. kk+1), q_old(kk-1), q_old(kk) and the
// centred weights fnm(kk), fnp(kk)
__device__ __forceinline__ void pd_vflux_sel(const PdFace &f, float qm2, float qm1, float q0, float qp1, float om1, float o0, float fm, float fp,
                                             float &anti, float &low) {
  const float lo = f.muf * flux_up(om1, o0, f.cr);
  const float h2 = f.vel * (fm * q0 + fp * qm1);
  const float h3 = f.vel * flux3pd(qm2, qm1, q0, qp1, -f.vel);
  anti = f.mode == 0 ? 0.f : (f.mode == 2 ? h2 : h3) - lo;
  low = f.mode == 0 ? 0.f : lo;
}

// pass 1: renormalisation factor of every cell (1 where not limited), over
// the full extent so the factor array is defined wherever pass 2 reads it.
// Face-sharing blocks as k_species_update: the west face comes from this
// lane, the east face from the next lane, the south face from this wave and
// the north face from the wave above (LDS).
__global__ void __launch_bounds__(64 * kFaceRows) k_species_pd_scale(const DevCtx xv, SpecList L, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  extern __shared__ float yal_[];  // 2 x L.n x kFaceRows x 64: south antidiffusive / upwind flux of every species
  GET_FACE_BLOCK_LG(r, (unsigned)x->face_lg[2]);
#define YA(s, w) yal_[((s) * kFaceRows + (w)) * 64 + lane]
#define YL(s, w) yal_[((L.n + (s)) * kFaceRows + (w)) * 64 + lane]
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const unsigned nb = (unsigned)x->n3, sj = (unsigned)(kde * x->ip), sk = (unsigned)x->ip;
  {
    // rows past the range still form their south face (the row below reads
    // it); its window may fall outside the field, the face is then 0
    const PdFace fs = pd_face_geom(x, 1, ic, k, j, dt);
    const unsigned os = XI3(ic, k, j);
    for (int s = 0; s < L.n; s++) {
      const BufF bo = bufv(L.old[s], nb);
      float w[6], a, l;
      win6b(bufv(L.in[s], nb), os, sj, w);
      const float om1 = bld(bo, os - sj), o0 = bld(bo, os);
      pin6(w);
      pd_hflux_br(fs, w, om1, o0, a, l);
      YA(s, wv) = a;
      YL(s, wv) = l;
    }
  }
  __syncthreads();
  if (!jrow) return;
  const bool lim = own && (ic >= 2 && ic <= ide - 2 && j >= 2 && j <= jde - 2 && k <= x->kde - 1);
  const unsigned o = XI3(ic, k, j);
  const float *c1 = FL(c1h), *c2 = FL(c2h);
  const float rdx = x->rdx, rdy = x->rdy, rz = A1(FL(rdnw), imin_(k, x->kde - 1));
  const float mm = A2(FL(msftx), ic, j) * A2(FL(msfty), ic, j), msfty = A2(FL(msfty), ic, j);
  const float mold = (A1(c1, k) * A2(FL(mub), ic, j) + A1(c2, k)) + (A1(c1, k) * A2(FL(mu_1), ic, j));
  const PdFace fx = pd_face_geom(x, 0, i, k, j, dt);
  const PdFace fyn = (wv == kFaceRows - 1) ? pd_face_geom(x, 1, ic, k, j + 1, dt) : PdFace{0.f, 0.f, 0.f, 0};
  const PdFace fz0 = pd_face_geom(x, 2, ic, k, j, dt), fz1 = pd_face_geom(x, 2, ic, k + 1, j, dt);
  const int k1 = imin_(k + 1, kde);
  const float fm0 = A1(FL(fnm), k), fp0 = A1(FL(fnp), k), fm1 = A1(FL(fnm), k1), fp1 = A1(FL(fnp), k1);
  const unsigned ox = XI3(i, k, j);
  for (int s = 0; s < L.n; s++) {
    const BufF bq = bufv(L.in[s], nb), bo = bufv(L.old[s], nb);
    // every load of this species first: x stencil of the west face, the
    // column windows of both vertical faces and the north face stencil (top
    // wave), with the old values of each
    float xw[6], yw[6], zw[5], zo[3];
    win6b(bq, ox, 1u, xw);
    const float xom = bld(bo, ox - 1u), xo0 = bld(bo, ox);
    win5z(bq, o, sk, zw);
#pragma unroll
    for (int n = 0; n < 3; n++) zo[n] = bld(bo, o + (unsigned)(n - 1) * sk);
    float yon = 0.f;
    if (wv == kFaceRows - 1) {
      win6b(bq, o + sj, sj, yw);
      yon = bld(bo, o + sj);
      pin6(yw);
    }
    pin6(xw);
    pin5(zw);
    float ax0, lx0, ay1, ly1, az0, lz0, az1, lz1;
    pd_hflux_br(fx, xw, xom, xo0, ax0, lx0);
    const float ax1 = lane_next(ax0, lane), lx1 = lane_next(lx0, lane);
    if (!own) continue;
    if (!lim) {
      L.scale[s][o] = 1.f;
      continue;
    }
    const float ay0 = YA(s, wv), ly0 = YL(s, wv);
    if (wv < kFaceRows - 1) {
      ay1 = YA(s, wv + 1);
      ly1 = YL(s, wv + 1);
    } else {
      pd_hflux_br(fyn, yw, zo[1], yon, ay1, ly1);
    }
    pd_vflux_sel(fz0, zw[0], zw[1], zw[2], zw[3], zo[0], zo[1], fm0, fp0, az0, lz0);
    pd_vflux_sel(fz1, zw[1], zw[2], zw[3], zw[4], zo[1], zo[2], fm1, fp1, az1, lz1);
    const float ph_low = mold * zo[1] - dt * (mm * (rdx * (lx1 - lx0) + rdy * (ly1 - ly0)) + msfty * rz * (lz1 - lz0));
    const float flux_out = dt * (mm * (rdx * (fmaxf(0.f, ax1) - fminf(0.f, ax0)) + rdy * (fmaxf(0.f, ay1) - fminf(0.f, ay0))) +
                                 msfty * rz * (fminf(0.f, az1) - fmaxf(0.f, az0)));
    float sc = 1.f;
    if (flux_out > ph_low) sc = fmaxf(0.f, ph_low / (flux_out + 1.0e-20f));
    L.scale[s][o] = sc;
  }
#undef YA
#undef YL
}

// pass 2: tendency from the renormalised faces, then the mass-weighted update
// (old = q_old after k_species_pd_prep, which also cleared the physics
// tendency over the same range: it is added as the +0 it holds).  Only the
// interior cells i [2, ide-2], j [2, jde-2] (the cells pass 1 limits: the
// specified zone is one cell wide) take an advective tendency.  The face
// fluxes are formed again as in pass 1 (face-sharing blocks over the mass
// points, the same loads and arithmetic, so the same values); each
// antidiffusive flux is scaled by the factor of the cell it leaves.
__global__ void __launch_bounds__(64 * kFaceRows) k_species_pd_update(const DevCtx xv, SpecList L, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  extern __shared__ float yal_[];  // 2 x L.n x kFaceRows x 64: south antidiffusive / upwind flux of every species
  GET_FACE_BLOCK_LG(r, (unsigned)x->face_lg[2]);
#define YA(s, w) yal_[((s) * kFaceRows + (w)) * 64 + lane]
#define YL(s, w) yal_[((L.n + (s)) * kFaceRows + (w)) * 64 + lane]
  const int kde = x->kde;
  const unsigned nb = (unsigned)x->n3, sj = (unsigned)(kde * x->ip), sk = (unsigned)x->ip;
  {
    const PdFace fs = pd_face_geom(x, 1, ic, k, j, dt);
    const unsigned os = XI3(ic, k, j);
    for (int s = 0; s < L.n; s++) {
      const BufF bo = bufv(L.old[s], nb);
      float w[6], a, l;
      win6b(bufv(L.in[s], nb), os, sj, w);
      const float om1 = bld(bo, os - sj), o0 = bld(bo, os);
      pin6(w);
      pd_hflux_br(fs, w, om1, o0, a, l);
      YA(s, wv) = a;
      YL(s, wv) = l;
    }
  }
  __syncthreads();
  if (!jrow) return;
  const unsigned o = XI3(ic, k, j), ox = XI3(i, k, j);
  const UpdCoef c = upd_coef(x, ic, k, j);
  const float rdx = x->rdx, rdy = x->rdy, rdzw = A1(FL(rdnw), k), msftx = A2(FL(msftx), ic, j);
  const bool up_ok = (k + 1 <= kde - 1), dn_ok = (k - 1 >= 1);
  const PdFace fx = pd_face_geom(x, 0, i, k, j, dt);
  const PdFace fyn = (wv == kFaceRows - 1) ? pd_face_geom(x, 1, ic, k, j + 1, dt) : PdFace{0.f, 0.f, 0.f, 0};
  const PdFace fz0 = pd_face_geom(x, 2, ic, k, j, dt), fz1 = pd_face_geom(x, 2, ic, k + 1, j, dt);
  const int k1 = imin_(k + 1, kde);
  const float fm0 = A1(FL(fnm), k), fp0 = A1(FL(fnp), k), fm1 = A1(FL(fnm), k1), fp1 = A1(FL(fnp), k1);
  for (int s = 0; s < L.n; s++) {
    const BufF bq = bufv(L.in[s], nb), bo = bufv(L.old[s], nb), bs = bufv(L.scale[s], nb);
    float xw[6], yw[6], zw[5], zo[3];
    win6b(bq, ox, 1u, xw);
    const float xom = bld(bo, ox - 1u), xo0 = bld(bo, ox);
    win5z(bq, o, sk, zw);
#pragma unroll
    for (int n = 0; n < 3; n++) zo[n] = bld(bo, o + (unsigned)(n - 1) * sk);
    float yon = 0.f;
    if (wv == kFaceRows - 1) {
      win6b(bq, o + sj, sj, yw);
      yon = bld(bo, o + sj);
      pin6(yw);
    }
    const float s0 = bld(bs, o), sw = bld(bs, o - 1u), se = bld(bs, o + 1u), ss = bld(bs, o - sj), sn = bld(bs, o + sj);
    const float sb = bld(bs, o - sk), st = bld(bs, o + sk);
    pin6(xw);
    pin5(zw);
    float ax0, lx0, ay1, ly1, az0, lz0, az1, lz1;
    pd_hflux_br(fx, xw, xom, xo0, ax0, lx0);
    const float ax1 = lane_next(ax0, lane), lx1 = lane_next(lx0, lane);
    if (!own) continue;
    if (!c.inner) {
      const float tend = 0.f + 0.f;
      elem(L.out[s], o) = (c.oldm * zo[1] + dt * tend) / c.newm;
      continue;
    }
    const float ay0 = YA(s, wv), ly0 = YL(s, wv);
    if (wv < kFaceRows - 1) {
      ay1 = YA(s, wv + 1);
      ly1 = YL(s, wv + 1);
    } else {
      pd_hflux_br(fyn, yw, zo[1], yon, ay1, ly1);
    }
    pd_vflux_sel(fz0, zw[0], zw[1], zw[2], zw[3], zo[0], zo[1], fm0, fp0, az0, lz0);
    pd_vflux_sel(fz1, zw[1], zw[2], zw[3], zw[4], zo[1], zo[2], fm1, fp1, az1, lz1);
    const float zt = az1 < 0.f ? az1 * s0 : ((az1 > 0.f && up_ok) ? az1 * st : az1);
    const float zb = az0 > 0.f ? az0 * s0 : ((az0 < 0.f && dn_ok) ? az0 * sb : az0);
    float t = 0.f;
    t += -rdzw * (zt - zb + lz1 - lz0);
    {
      const float xe = ax1 > 0.f ? ax1 * s0 : (ax1 < 0.f ? ax1 * se : ax1);
      const float xwf = ax0 < 0.f ? ax0 * s0 : (ax0 > 0.f ? ax0 * sw : ax0);
      t += -msftx * (rdx * (xe - xwf + lx1 - lx0));
    }
    {
      const float yn = ay1 > 0.f ? ay1 * s0 : (ay1 < 0.f ? ay1 * sn : ay1);
      const float ys = ay0 < 0.f ? ay0 * s0 : (ay0 > 0.f ? ay0 * ss : ay0);
      t += -msftx * (rdy * (yn - ys + ly1 - ly0));
    }
    float tend = t * c.msfty;
    tend = tend + 0.f;
    elem(L.out[s], o) = (c.oldm * zo[1] + dt * tend) / c.newm;
  }
#undef YA
#undef YL
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
