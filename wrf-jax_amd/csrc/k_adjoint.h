// Tangent-linear and adjoint of the scalar transport of one RK stage:
// advect_scalar (flux form, 5th-order horizontal / 3rd-order vertical,
// specified boundaries; module_advect_em.py:90-369, _fast.py:1437-1769)
// followed by rk_update_scalar (module_em.py:1437-1598) with the stage's
// recorded mass fluxes (ru_m, rv_m, ww_m) and column masses (mu_1, mu_2):
//   tend  = A q_in                   (A: the advect_scalar operator)
//   q_out = (m_old q_old + dt msfty [inner] tend) / m_new
// with m = c1h (mu + mub) + c2h.  For fixed fluxes the stage is linear in
// (q_old, q_in), so the tangent-linear stage is the stage itself (the
// dycore's own advect_scalar kernel and k_tl_update), and the adjoint is its
// transpose: k_ad_update then k_ad_advect (q_in_bar += A^T tend_bar).
#pragma once

#include "wrfdc_common.h"

struct StageMass {
  const float *mu_1, *mu_2;  // recorded column masses of the stage (perturbation; + mub)
};

// q_out over the mass points [1, ide-1] x [1, kde-1] x [1, jde-1]
__global__ void k_tl_update(const DevCtx xv, R3 r, StageMass sm, const float *__restrict__ q_old, const float *__restrict__ tend,
                            float *__restrict__ q_out, float dt) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const int sz = x->spec_zone;
  const bool inner = i >= 1 + sz && i <= x->ide - sz - 1 && j >= 1 + sz && j <= x->jde - sz - 1;
  const float mub = A2(FL(mub), i, j), c1 = A1(FL(c1h), k), c2 = A1(FL(c2h), k);
  const float oldm = c1 * (A2(sm.mu_1, i, j) + mub) + c2, newm = c1 * (A2(sm.mu_2, i, j) + mub) + c2;
  const float t = inner ? A3(tend, i, k, j) * A2(FL(msfty), i, j) : 0.f;
  A3(q_out, i, k, j) = (oldm * A3(q_old, i, k, j) + dt * t) / newm;
}

// transpose of k_tl_update: q_old_bar += (m_old / m_new) q_out_bar,
// tend_bar = dt msfty [inner] / m_new q_out_bar
__global__ void k_ad_update(const DevCtx xv, R3 r, StageMass sm, const float *__restrict__ qbar_out, float *__restrict__ qbar_old,
                            float *__restrict__ tbar, float dt) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const int sz = x->spec_zone;
  const bool inner = i >= 1 + sz && i <= x->ide - sz - 1 && j >= 1 + sz && j <= x->jde - sz - 1;
  const float mub = A2(FL(mub), i, j), c1 = A1(FL(c1h), k), c2 = A1(FL(c2h), k);
  const float oldm = c1 * (A2(sm.mu_1, i, j) + mub) + c2, newm = c1 * (A2(sm.mu_2, i, j) + mub) + c2;
  const float qb = A3(qbar_out, i, k, j);
  A3(qbar_old, i, k, j) += oldm / newm * qb;
  A3(tbar, i, k, j) = inner ? dt * A2(FL(msfty), i, j) / newm * qb : 0.f;
}

// coefficient of q(f + m) in the flux through face f (flux5 / flux3 /
// second order of module_advect_em.py:29-57), for offset m in [-3, 2] and the
// sign s of the face velocity
__device__ __forceinline__ float face_coef(int order, float s, int m) {
  if (order == 5) {
    switch (m) {
      case -3: return (1.f + s) / 60.f;
      case -2: return (-8.f - 5.f * s) / 60.f;
      case -1: return (37.f + 10.f * s) / 60.f;
      case 0: return (37.f - 10.f * s) / 60.f;
      case 1: return (-8.f + 5.f * s) / 60.f;
      case 2: return (1.f - s) / 60.f;
      default: return 0.f;
    }
  }
  if (order == 3) {
    switch (m) {
      case -2: return (-1.f - s) / 12.f;
      case -1: return (7.f + 3.f * s) / 12.f;
      case 0: return (7.f - 3.f * s) / 12.f;
      case 1: return (-1.f + s) / 12.f;
      default: return 0.f;
    }
  }
  return (m == -1 || m == 0) ? 0.5f : 0.f;
}

// q_in_bar += A^T tend_bar at the mass points [1, ide-1] x [1, kde-1] x
// [1, jde-1] (gather: each point collects the faces whose fluxes read it).
// Horizontal faces f in [2, e-1] (e = ide or jde) carry order 2 at f = 2 and
// e-1, 5 in [4, e-3], else 3; tendencies exist at [2, e-2].  Vertical faces
// f in [3, ktf-1] are 3rd order with the velocity's sign reversed (eta falls
// upward), f = 2 and ktf second order through fzm / fzp, 1 and kde zero.
__global__ void k_ad_advect(const DevCtx xv, R3 r, const float *__restrict__ tbar, const float *__restrict__ ru,
                            const float *__restrict__ rv, const float *__restrict__ rom, float *__restrict__ qbar) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const int ide = x->ide, jde = x->jde, ktf = x->kde - 1;
  const float rdx = x->rdx, rdy = x->rdy;
  const float *msftx = FL(msftx), *rdzw = FL(rdnw), *fzm = FL(fnm), *fzp = FL(fnp);
  float acc = 0.f;
  // x faces f = i - m, m in [-3, 2]
  for (int m = -3; m <= 2; m++) {
    const int f = i - m;
    if (f < 2 || f > ide - 1) continue;
    const int order = (f >= 4 && f <= ide - 3) ? 5 : ((f == 2 || f == ide - 1) ? 2 : 3);
    if (face_coef(order, 1.f, m) == 0.f && face_coef(order, -1.f, m) == 0.f) continue;
    // F_bar(f) = msftx(f) rdx tbar(f) [f tendency] - msftx(f-1) rdx tbar(f-1) [f-1 tendency]
    float fb = 0.f;
    if (f <= ide - 2) fb += A2(msftx, f, j) * rdx * A3(tbar, f, k, j);
    if (f - 1 >= 2) fb -= A2(msftx, f - 1, j) * rdx * A3(tbar, f - 1, k, j);
    const float vel = A3(ru, f, k, j);
    acc += vel * face_coef(order, sgn1(vel), m) * fb;
  }
  // y faces
  for (int m = -3; m <= 2; m++) {
    const int f = j - m;
    if (f < 2 || f > jde - 1) continue;
    const int order = (f >= 4 && f <= jde - 3) ? 5 : ((f == 2 || f == jde - 1) ? 2 : 3);
    if (face_coef(order, 1.f, m) == 0.f && face_coef(order, -1.f, m) == 0.f) continue;
    float fb = 0.f;
    if (f <= jde - 2) fb += A2(msftx, i, f) * rdy * A3(tbar, i, k, f);
    if (f - 1 >= 2) fb -= A2(msftx, i, f - 1) * rdy * A3(tbar, i, k, f - 1);
    const float vel = A3(rv, i, k, f);
    acc += vel * face_coef(order, sgn1(vel), m) * fb;
  }
  // vertical faces f = k - m, m in [-2, 1]
  for (int m = -2; m <= 1; m++) {
    const int f = k - m;
    if (f < 2 || f > ktf) continue;
    float coef;
    const float vel = A3(rom, i, f, j);
    if (f == 2 || f == ktf) {
      coef = m == 0 ? A1(fzm, f) : (m == -1 ? A1(fzp, f) : 0.f);
    } else {
      coef = face_coef(3, sgn1(-vel), m);
    }
    if (coef == 0.f) continue;
    // V_bar(f) = rdzw(f) tbar(f) - rdzw(f-1) tbar(f-1), tendencies at [1, ktf]
    float vb = 0.f;
    if (f <= ktf) vb += A1(rdzw, f) * A3(tbar, i, f, j);
    if (f - 1 >= 1) vb -= A1(rdzw, f - 1) * A3(tbar, i, f - 1, j);
    acc += vel * coef * vb;
  }
  A3(qbar, i, k, j) += acc;
}
