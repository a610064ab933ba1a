// diff_opt = 2 / km_opt = 4 kernels of the first RK stage: metric terms,
// deformation / divergence, 2-D Smagorinsky diffusivity and the horizontal
// diffusion of momentum, theta and the transported species.
// Semantics: src/wrf_jax/dyn_em/module_diffusion_em.py (specified
// boundaries, no SFS/NBA terms, mix_full_fields = false).
#pragma once
#include "wrfdc_common.h"

// rho = (1 + qv) / alt on mass points (phy_prep,
// module_big_step_utilities_em.py:1823-2081, rho at :1965)
__global__ void k_rho(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  A3(FL(rho), i, k, j) = (1.f / A3(FL(alt), i, k, j)) * (1.f + A3(SP(moist, 0), i, k, j));
}

// compute_diff_metrics (module_diffusion_em.py:1825-2043), full extent.
__global__ void k_diff_metrics(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const int ide = x->ide, jde = x->jde, kde = x->kde, ktf = kde - 1;
  const float *ph = FL(ph_2), *phb = FL(phb);
  const float g = x->g;
#define ZW(ii, kk, jj) ((A3(ph, ii, kk, jj) + A3(phb, ii, kk, jj)) / g)
  if (i <= ide - 1) {
    if (k <= ktf) A3(FL(rdzw), i, k, j) = 1.f / (ZW(i, k + 1, j) - ZW(i, k, j));
    if (k >= 2 && k <= ktf) A3(FL(rdz), i, k, j) = 2.f / (ZW(i, k + 1, j) - ZW(i, k - 1, j));
    if (k == 1) A3(FL(rdz), i, 1, j) = 2.f / (ZW(i, 2, j) - ZW(i, 1, j));
  }
#undef ZW
  if (j <= jde - 1) {
    if (i >= 2 && i <= ide - 1) {
      const float zb = (x->rdx * (A3(phb, i, k, j) - A3(phb, i - 1, k, j))) / g;
      const float zp = (x->rdx * (A3(ph, i, k, j) - A3(ph, i - 1, k, j))) / g;
      A3(FL(zx), i, k, j) = zb + zp;
    } else if (k <= ktf) {
      A3(FL(zx), i, k, j) = 0.f;
    }
  }
  if (i <= ide - 1) {
    if (j >= 2 && j <= jde - 1) {
      const float zb = (x->rdy * (A3(phb, i, k, j) - A3(phb, i, k, j - 1))) / g;
      const float zp = (x->rdy * (A3(ph, i, k, j) - A3(ph, i, k, j - 1))) / g;
      A3(FL(zy), i, k, j) = zb + zp;
    } else if (k <= ktf) {
      A3(FL(zy), i, k, j) = 0.f;
    }
  }
  if (i <= ide - 1 && j <= jde - 1 && k <= ktf) {
    float s = A3(ph, i, k, j) + A3(phb, i, k, j);
    s = s + A3(ph, i, k + 1, j);
    s = s + A3(phb, i, k + 1, j);
    A3(FL(z), i, k, j) = (0.5f * s) / g;
  }
}

// value of a staggered field scaled by a map factor
__device__ __forceinline__ float hatv(const DevCtx *__restrict__ x, const float *f, const float *m, int i, int k, int j) {
  return A3(f, i, k, j) / A2(m, i, j);
}
// average of the two columns A and B of hat = f / m onto w level k
// (the hatavg construction shared by the deformation components)
__device__ __forceinline__ float hat_avg(const DevCtx *__restrict__ x, const float *f, const float *m, int ia, int ja, int ib, int jb, int k) {
  const int kte = x->kde, ktes1 = kte - 1, ktes2 = kte - 2;
  if (k == 1)
    return 0.5f * (x->cf1 * hatv(x, f, m, ia, 1, ja) + x->cf2 * hatv(x, f, m, ia, 2, ja) + x->cf3 * hatv(x, f, m, ia, 3, ja) +
                   x->cf1 * hatv(x, f, m, ib, 1, jb) + x->cf2 * hatv(x, f, m, ib, 2, jb) + x->cf3 * hatv(x, f, m, ib, 3, jb));
  if (k == kte) {
    const float cft2 = -0.5f * A1(FL(dnw), ktes1) / A1(FL(dn), ktes1), cft1 = 1.f - cft2;
    return 0.5f * (cft1 * (hatv(x, f, m, ia, ktes1, ja) + hatv(x, f, m, ib, ktes1, jb)) +
                   cft2 * (hatv(x, f, m, ia, ktes2, ja) + hatv(x, f, m, ib, ktes2, jb)));
  }
  return 0.5f * (A1(FL(fnm), k) * (hatv(x, f, m, ia, k, ja) + hatv(x, f, m, ib, k, jb)) +
                 A1(FL(fnp), k) * (hatv(x, f, m, ia, k - 1, ja) + hatv(x, f, m, ib, k - 1, jb)));
}

// cal_deform_and_div (module_diffusion_em.py:366-891): D11, D22, D33 and the
// divergence on mass points.
__global__ void k_deform_diag(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *u = FL(u_2), *v = FL(v_2), *w = FL(w_2), *zx = FL(zx), *zy = FL(zy), *rdzw = FL(rdzw);
  const float *msfuy = FL(msfuy), *msfvx = FL(msfvx);
  const float mm = A2(FL(msftx), i, j) * A2(FL(msfty), i, j);
  float t1;
  {
    const float tmpzx = 0.25f * (A3(zx, i, k, j) + A3(zx, i + 1, k, j) + A3(zx, i, k + 1, j) + A3(zx, i + 1, k + 1, j));
    t1 = (hat_avg(x, u, msfuy, i, j, i + 1, j, k + 1) - hat_avg(x, u, msfuy, i, j, i + 1, j, k)) * tmpzx * A3(rdzw, i, k, j);
    t1 = mm * (x->rdx * (hatv(x, u, msfuy, i + 1, k, j) - hatv(x, u, msfuy, i, k, j)) - t1);
    A3(FL(defor11), i, k, j) = 2.f * t1;
  }
  float dv = t1;
  {
    const float tmpzy = 0.25f * (A3(zy, i, k, j) + A3(zy, i, k, j + 1) + A3(zy, i, k + 1, j) + A3(zy, i, k + 1, j + 1));
    t1 = (hat_avg(x, v, msfvx, i, j, i, j + 1, k + 1) - hat_avg(x, v, msfvx, i, j, i, j + 1, k)) * tmpzy * A3(rdzw, i, k, j);
    t1 = mm * (x->rdy * (hatv(x, v, msfvx, i, k, j + 1) - hatv(x, v, msfvx, i, k, j)) - t1);
    A3(FL(defor22), i, k, j) = 2.f * t1;
  }
  dv = dv + t1;
  t1 = (A3(w, i, k + 1, j) - A3(w, i, k, j)) * A3(rdzw, i, k, j);
  A3(FL(defor33), i, k, j) = 2.f * t1;
  dv = dv + t1;
  A3(FL(div), i, k, j) = dv;
}

// D12 at (u, v) corner points i [2, ide-1], j [2, jde-1], k [1, kde-1]
__global__ void k_deform12(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *u = FL(u_2), *v = FL(v_2), *zx = FL(zx), *zy = FL(zy), *rdzw = FL(rdzw), *msfux = FL(msfux), *msfvy = FL(msfvy);
  const float mm = 0.25f * (A2(msfux, i, j - 1) + A2(msfux, i, j)) * (A2(msfvy, i - 1, j) + A2(msfvy, i, j));
  float d;
  {
    const float tmpzy = 0.25f * (A3(zy, i - 1, k, j) + A3(zy, i, k, j) + A3(zy, i - 1, k + 1, j) + A3(zy, i, k + 1, j));
    const float t1 = (hat_avg(x, u, msfux, i, j - 1, i, j, k + 1) - hat_avg(x, u, msfux, i, j - 1, i, j, k)) * 0.25f * tmpzy *
                     (A3(rdzw, i, k, j) + A3(rdzw, i - 1, k, j) + A3(rdzw, i - 1, k, j - 1) + A3(rdzw, i, k, j - 1));
    d = mm * (x->rdy * (hatv(x, u, msfux, i, k, j) - hatv(x, u, msfux, i, k, j - 1)) - t1);
  }
  {
    const float tmpzx = 0.25f * (A3(zx, i, k, j - 1) + A3(zx, i, k, j) + A3(zx, i, k + 1, j - 1) + A3(zx, i, k + 1, j));
    const float t1 = (hat_avg(x, v, msfvy, i - 1, j, i, j, k + 1) - hat_avg(x, v, msfvy, i - 1, j, i, j, k)) * 0.25f * tmpzx *
                     (A3(rdzw, i, k, j) + A3(rdzw, i, k, j - 1) + A3(rdzw, i - 1, k, j - 1) + A3(rdzw, i - 1, k, j));
    const float second = mm * (x->rdx * (hatv(x, v, msfvy, i, k, j) - hatv(x, v, msfvy, i - 1, k, j)) - t1);
    d = d + second;
  }
  A3(FL(defor12), i, k, j) = d;
}

// D13 / D23 at (u, w) and (v, w) points i [2, ide-1], j [2, jde-1], k [1, kde]
__global__ void k_deform13_23(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const int ktf = x->kde - 1;
  float *d13 = FL(defor13), *d23 = FL(defor23);
  if (k == 1 || k == ktf + 1) {
    A3(d13, i, k, j) = 0.f;
    A3(d23, i, k, j) = 0.f;
    return;
  }
  const float *w = FL(w_2), *u = FL(u_2), *v = FL(v_2), *zx = FL(zx), *zy = FL(zy), *rdz = FL(rdz);
  const float *msfty = FL(msfty), *msftx = FL(msftx), *ub = FL(u_base), *vb = FL(v_base);
  {
#define H13(ii, kk) hatv(x, w, msfty, ii, kk, j)
#define AV13(kk) (0.25f * (H13(i, kk) + H13(i, (kk) + 1) + H13(i - 1, kk) + H13(i - 1, (kk) + 1)))
    const float mm = A2(FL(msfux), i, j) * A2(FL(msfuy), i, j);
    const float t1 = (AV13(k) - AV13(k - 1)) * A3(zx, i, k, j) * 0.5f * (A3(rdz, i, k, j) + A3(rdz, i - 1, k, j));
    float d = mm * (x->rdx * (H13(i, k) - H13(i - 1, k)) - t1);
    const float val = (A3(u, i, k, j) - A1(ub, k) - A3(u, i, k - 1, j) + A1(ub, k - 1)) * 0.5f * (A3(rdz, i, k, j) + A3(rdz, i - 1, k, j));
    A3(d13, i, k, j) = d + val;
#undef AV13
#undef H13
  }
  {
#define H23(jj, kk) hatv(x, w, msftx, i, kk, jj)
#define AV23(kk) (0.25f * (H23(j, kk) + H23(j, (kk) + 1) + H23(j - 1, kk) + H23(j - 1, (kk) + 1)))
    const float mm = A2(FL(msfvx), i, j) * A2(FL(msfvy), i, j);
    const float t1 = (AV23(k) - AV23(k - 1)) * A3(zy, i, k, j) * 0.5f * (A3(rdz, i, k, j) + A3(rdz, i, k, j - 1));
    float d = mm * (x->rdy * (H23(j, k) - H23(j - 1, k)) - t1);
    const float val = (A3(v, i, k, j) - A1(vb, k) - A3(v, i, k - 1, j) + A1(vb, k - 1)) * 0.5f * (A3(rdz, i, k, j) + A3(rdz, i, k, j - 1));
    A3(d23, i, k, j) = d + val;
#undef AV23
#undef H23
  }
}

// edge copy of whole planes: i = idst <- isrc for every (k, j) (axis 0) or
// j = jdst <- jsrc for every (i, k) (axis 1); up to two fields per launch.
__global__ void k_edge_copy(const DevCtx xv, float *f0, float *f1, int axis, int dst, int src, int kend) {
  const DevCtx *__restrict__ x = &xv;
  const int a = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  if (k > kend) return;
  if (axis == 0) {
    if (a > x->jde) return;
    A3(f0, dst, k, a) = A3(f0, src, k, a);
    if (f1) A3(f1, dst, k, a) = A3(f1, src, k, a);
  } else {
    if (a > x->ide) return;
    A3(f0, a, k, dst) = A3(f0, a, k, src);
    if (f1) A3(f1, a, k, dst) = A3(f1, a, k, src);
  }
}

// smag2d_km (module_diffusion_em.py:1502-1646) with the slope correction;
// i [2, ide-2], j [2, jde-2], k [1, kde-1].  With the learned closure on,
// |D| = sqrt(def2) and the slope divisor are stored too (defmag, divisor).
__global__ void k_smag(const DevCtx xv, R3 r, float *defmag = nullptr, float *divout = nullptr) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *d11 = FL(defor11), *d22 = FL(defor22), *d12 = FL(defor12), *zx = FL(zx), *zy = FL(zy), *rdzw = FL(rdzw);
  const float dx = x->dx, dy = x->dy, c_s = x->c_s;
  const float mx = A2(FL(msftx), i, j), my = A2(FL(msfty), i, j);
  const float dd = A3(d11, i, k, j) - A3(d22, i, k, j);
  float def2 = 0.25f * dd * dd;
  const float tmp = 0.25f * (A3(d12, i, k, j) + A3(d12, i, k, j + 1) + A3(d12, i + 1, k, j) + A3(d12, i + 1, k, j + 1));
  def2 = def2 + tmp * tmp;
  const float mlen = sqrtf(dx / mx * dy / my);
  const float sd = sqrtf(def2);
  float km = c_s * c_s * mlen * mlen * sd;
  km = fminf(km, 10.f * mlen);
  const float dxm = dx / mx, dym = dy / my;
  const float tzx = 0.25f * (fabsf(A3(zx, i, k, j)) + fabsf(A3(zx, i + 1, k, j)) + fabsf(A3(zx, i, k + 1, j)) + fabsf(A3(zx, i + 1, k + 1, j))) *
                    A3(rdzw, i, k, j) * dxm;
  const float tzy = 0.25f * (fabsf(A3(zy, i, k, j)) + fabsf(A3(zy, i, k, j + 1)) + fabsf(A3(zy, i, k + 1, j)) + fabsf(A3(zy, i, k + 1, j + 1))) *
                    A3(rdzw, i, k, j) * dym;
  const float alpha = fmaxf(sqrtf(tzx * tzx + tzy * tzy), 1.f);
  const float lim = fmaxf(10.f / mlen, 1.0e-3f);
  const float divisor = sd > lim ? alpha * alpha : alpha;
  km = km / divisor;
  if (defmag) {
    A3(defmag, i, k, j) = sd;
    A3(divout, i, k, j) = divisor;
  }
  A3(FL(xkmh), i, k, j) = km;
  A3(FL(xkmv), i, k, j) = km;
  A3(FL(xkhh), i, k, j) = km / x->prandtl;
  A3(FL(xkhv), i, k, j) = 0.f;
}

// phy_prep's full theta, pressure and temperature at one mass point
// (module_big_step_utilities_em.py:1823-2081, use_theta_m 0)
struct Thermo {
  float th, p, t;
};
__device__ __forceinline__ Thermo thermo_at(const DevCtx *__restrict__ x, int i, int k, int j) {
  Thermo r;
  r.th = A3(FL(t_2), i, k, j) + x->t0;
  r.p = A3(FL(p), i, k, j) + A3(FL(pb), i, k, j);
  const float rcp = (float)(287.0 / (7.0 * 287.0 / 2.0));
  r.t = r.th * powf(r.p / 100000.0f, rcp);
  return r;
}
// saturation mixing ratio of calculate_km_kh
__device__ __forceinline__ float qvs_of(float t, float p) {
  const float es = 1000.0f * 0.6112f * expf(17.67f * (t - 273.15f) / (t - 29.65f));
  return (float)(287.0 / 461.6) * es / (p - es);
}
__device__ __forceinline__ float qtot_at(const DevCtx *__restrict__ x, int i, int k, int j) {
  float q = A3(SP(moist, 0), i, k, j);
  if (x->nmoist > 1) q = q + A3(SP(moist, 1), i, k, j);
  if (x->nmoist > 3) q = q + A3(SP(moist, 3), i, k, j);
  return q;
}
__device__ __forceinline__ float qc_at(const DevCtx *__restrict__ x, int i, int k, int j) {
  return x->nmoist > 1 ? A3(SP(moist, 1), i, k, j) : 0.f;
}

// N^2 of calculate_km_kh (module_diffusion_em_fast.py:933-1005): moist where
// saturated (qv >= qvs or qc >= 1e-5), dry otherwise; level 1 from the
// surface values (pressure and temperature on w level 1 through mass levels
// 1 and 2, linear in height), level ktf repeats ktf-1.  i [2, ide-2],
// k [1, kde-1], j [2, jde-2]; one thread per point.
__global__ void k_n2(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const float G = 9.81f, R_D = 287.0f, CP = (float)(7.0 * 287.0 / 2.0), XLV = 2.5e6f, RV = 461.6f, C161 = 1.61f;
  const float *qv = SP(moist, 0), *rdz = FL(rdz), *rdzw = FL(rdzw);
  const int ktf = x->kde - 1;
  float val;
  if (k >= 2) {
    const int kk = k < ktf ? k : ktf - 1;  // level ktf repeats ktf-1
    const Thermo a = thermo_at(x, i, kk - 1, j), b = thermo_at(x, i, kk, j), c = thermo_at(x, i, kk + 1, j);
    const float tmpdz = 1.0f / A3(rdz, i, kk, j) + 1.0f / A3(rdz, i, kk + 1, j);
    const float qvk = A3(qv, i, kk, j);
    const bool sat = (qvk >= qvs_of(b.t, b.p)) || (qc_at(x, i, kk, j) >= 0.00001f);
    const float dqt = qtot_at(x, i, kk + 1, j) - qtot_at(x, i, kk - 1, j);
    if (sat) {
      const float xlvqv = XLV * qvk;
      const float coefa = (1.0f + xlvqv / R_D / b.t) / (1.0f + XLV * xlvqv / CP / RV / b.t / b.t) / b.th;
      const float thetaep1 = c.th * (1.0f + XLV * qvs_of(c.t, c.p) / CP / c.t);
      const float thetaem1 = a.th * (1.0f + XLV * qvs_of(a.t, a.p) / CP / a.t);
      val = G * (coefa * (thetaep1 - thetaem1) / tmpdz - dqt / tmpdz);
    } else {
      val = G * ((c.th - a.th) / b.th / tmpdz + C161 * (A3(qv, i, kk + 1, j) - A3(qv, i, kk - 1, j)) / tmpdz - dqt / tmpdz);
    }
  } else {
    const float g = x->g, cf1 = x->cf1, cf2 = x->cf2, cf3 = x->cf3;
    const Thermo a = thermo_at(x, i, 1, j), b = thermo_at(x, i, 2, j), c = thermo_at(x, i, 3, j);
    const float zw1 = (A3(FL(phb), i, 1, j) + A3(FL(ph_2), i, 1, j)) / g, zw2 = (A3(FL(phb), i, 2, j) + A3(FL(ph_2), i, 2, j)) / g;
    const float zw3 = (A3(FL(phb), i, 3, j) + A3(FL(ph_2), i, 3, j)) / g;
    const float z1 = 0.5f * (zw1 + zw2), z2 = 0.5f * (zw2 + zw3);
    const float w1 = (zw1 - z2) / (z1 - z2), w2 = 1.0f - w1;
    const float p8w = w1 * a.p + w2 * b.p, t8w = w1 * a.t + w2 * b.t;
    const float qvs1 = qvs_of(a.t, a.p), qvs2 = qvs_of(b.t, b.p), qvs3 = qvs_of(c.t, c.p);
    const float qv1 = A3(qv, i, 1, j), qv2 = A3(qv, i, 2, j), qv3 = A3(qv, i, 3, j);
    float qtsfc = cf1 * qv1 + cf2 * qv2 + cf3 * qv3;
    if (x->nmoist > 1) {
      const float *qc = SP(moist, 1);
      qtsfc = qtsfc + cf1 * A3(qc, i, 1, j) + cf2 * A3(qc, i, 2, j) + cf3 * A3(qc, i, 3, j);
    }
    if (x->nmoist > 3) {
      const float *qi = SP(moist, 3);
      qtsfc = qtsfc + cf1 * A3(qi, i, 1, j) + cf2 * A3(qi, i, 2, j) + cf3 * A3(qi, i, 3, j);
    }
    const float dqt = qtot_at(x, i, 2, j) - qtsfc;
    const bool sat = (qv1 >= qvs1) || (qc_at(x, i, 1, j) >= 0.00001f);
    if (sat) {
      const float tmpdz = 1.0f / A3(rdz, i, 2, j) + 0.5f / A3(rdzw, i, 1, j);
      const float thsfc = t8w / powf(p8w / 100000.0f, R_D / CP);
      const float qvsfc_sat = cf1 * qvs1 + cf2 * qvs2 + cf3 * qvs3;
      const float xlvqv = XLV * qv1;
      const float coefa = (1.0f + xlvqv / R_D / a.t) / (1.0f + XLV * xlvqv / CP / RV / a.t / a.t) / a.th;
      const float thetaep1 = b.th * (1.0f + XLV * qvs2 / CP / b.t);
      const float thetaesfc = thsfc * (1.0f + XLV * qvsfc_sat / CP / t8w);
      val = G * (coefa * (thetaep1 - thetaesfc) / tmpdz - dqt / tmpdz);
    } else {
      const float tmpdz = 1.0f / A3(rdzw, i, 1, j);
      const float qvsfc = cf1 * qv1 + cf2 * qv2 + cf3 * qv3;
      val = G * ((b.th - a.th) / a.th / tmpdz + C161 * (qv2 - qvsfc) / tmpdz - dqt / tmpdz);
    }
  }
  A3(FL(bn2), i, k, j) = val;
}

// titau (module_diffusion_em.py:894-1316): stress component with rho and the
// diffusivity averaged to the component's point; zero outside [i0,i1]x[j0,j1]
// (and at k = 1, kde for the 13 / 23 components).  kind 11, 12, 13, 23.
// one stress component at (i, k, j): zero outside [i0, i1] x [j0, j1] x [1, kde-1]
__device__ __forceinline__ float titau_val(const DevCtx *__restrict__ x, const float *xk, const float *defor, int kind, int i0, int i1,
                                           int j0, int j1, int i, int k, int j) {
  const float *rho = FL(rho);
  const int ktf = x->kde - 1;
  float val = 0.f;
  if (i >= i0 && i <= i1 && j >= j0 && j <= j1 && k <= ktf) {
    if (kind == 11) {
      val = -A3(rho, i, k, j) * A3(xk, i, k, j) * A3(defor, i, k, j);
    } else if (kind == 12) {
      const float ra = 0.25f * (A3(rho, i - 1, k, j) + A3(rho, i, k, j) + A3(rho, i - 1, k, j - 1) + A3(rho, i, k, j - 1));
      const float xa = ra * 0.25f * (A3(xk, i - 1, k, j) + A3(xk, i, k, j) + A3(xk, i - 1, k, j - 1) + A3(xk, i, k, j - 1));
      val = -xa * A3(defor, i, k, j);
    } else if (k >= 2) {
      const float *fnm = FL(fnm), *fnp = FL(fnp);
      const int di = kind == 13 ? 1 : 0, dj = kind == 13 ? 0 : 1;
      const float ra = 0.5f * (A1(fnm, k) * (A3(rho, i - di, k, j - dj) + A3(rho, i, k, j)) + A1(fnp, k) * (A3(rho, i - di, k - 1, j - dj) + A3(rho, i, k - 1, j)));
      const float xa = ra * 0.5f * (A1(fnm, k) * (A3(xk, i, k, j) + A3(xk, i - di, k, j - dj)) + A1(fnp, k) * (A3(xk, i, k - 1, j) + A3(xk, i - di, k - 1, j - dj)));
      val = -xa * A3(defor, i, k, j);
    }
  }
  return val;
}
// the two stress components one diffusion kernel reads, in one pass
struct TitauPair {
  float *tt[2];
  const float *xk[2], *defor[2];
  int kind[2], i0[2], i1[2], j0[2], j1[2];
};
__global__ void k_titau(const DevCtx xv, TitauPair P, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float v0 = titau_val(x, P.xk[0], P.defor[0], P.kind[0], P.i0[0], P.i1[0], P.j0[0], P.j1[0], i, k, j);
  const float v1 = titau_val(x, P.xk[1], P.defor[1], P.kind[1], P.i0[1], P.i1[1], P.j0[1], P.j1[1], i, k, j);
  A3(P.tt[0], i, k, j) = v0;
  A3(P.tt[1], i, k, j) = v1;
}

// horizontal_diffusion_u_2 (module_diffusion_em.py:2302-2532):
// u points i [2, ide-1], j [2, jde-2], k [1, kde-1].
__device__ __forceinline__ float hd_a1u(const DevCtx *__restrict__ x, const float *t1, int i, int k, int j) {
  if (k == 1 || k == x->kde) return 0.f;
  return 0.5f * (A1(FL(fnm), k) * (A3(t1, i - 1, k, j) + A3(t1, i, k, j)) + A1(FL(fnp), k) * (A3(t1, i - 1, k - 1, j) + A3(t1, i, k - 1, j)));
}
__device__ __forceinline__ float hd_a2u(const DevCtx *__restrict__ x, const float *t2, int i, int k, int j) {
  if (k == 1 || k == x->kde) return 0.f;
  return 0.5f * (A1(FL(fnm), k) * (A3(t2, i, k, j + 1) + A3(t2, i, k, j)) + A1(FL(fnp), k) * (A3(t2, i, k - 1, j + 1) + A3(t2, i, k - 1, j)));
}
__global__ void k_hdiff_u(const DevCtx xv, const float *t1, const float *t2, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *zx = FL(zx), *zy = FL(zy), *rdzw = FL(rdzw);
  const float zxu = 0.5f * (A3(zx, i, k, j) + A3(zx, i, k + 1, j));
  const float zyu = 0.125f * (A3(zy, i - 1, k, j) + A3(zy, i, k, j) + A3(zy, i - 1, k, j + 1) + A3(zy, i, k, j + 1) + A3(zy, i - 1, k + 1, j) +
                              A3(zy, i, k + 1, j) + A3(zy, i - 1, k + 1, j + 1) + A3(zy, i, k + 1, j + 1));
  const float msfux = A2(FL(msfux), i, j), msfuy = A2(FL(msfuy), i, j);
  const float mrdx = msfux * x->rdx, mrdy = msfuy * x->rdy;
  const float tmpdz = (1.f / A3(rdzw, i, k, j) + 1.f / A3(rdzw, i - 1, k, j)) / 2.f;
  const float inc = x->g * tmpdz / A1(FL(dnw), k) *
                    (mrdx * (A3(t1, i, k, j) - A3(t1, i - 1, k, j)) + mrdy * (A3(t2, i, k, j + 1) - A3(t2, i, k, j)) -
                     msfux * zxu * (hd_a1u(x, t1, i, k + 1, j) - hd_a1u(x, t1, i, k, j)) / tmpdz -
                     msfuy * zyu * (hd_a2u(x, t2, i, k + 1, j) - hd_a2u(x, t2, i, k, j)) / tmpdz);
  A3(FL(ru_tendf), i, k, j) = A3(FL(ru_tendf), i, k, j) + inc;
}

// horizontal_diffusion_v_2 (module_diffusion_em.py:2535-2765):
// v points i [2, ide-2], j [2, jde-1], k [1, kde-1].
__device__ __forceinline__ float hd_a1v(const DevCtx *__restrict__ x, const float *t1, int i, int k, int j) {
  if (k == 1 || k == x->kde) return 0.f;
  return 0.5f * (A1(FL(fnm), k) * (A3(t1, i + 1, k, j) + A3(t1, i, k, j)) + A1(FL(fnp), k) * (A3(t1, i + 1, k - 1, j) + A3(t1, i, k - 1, j)));
}
__device__ __forceinline__ float hd_a2v(const DevCtx *__restrict__ x, const float *t2, int i, int k, int j) {
  if (k == 1 || k == x->kde) return 0.f;
  return 0.5f * (A1(FL(fnm), k) * (A3(t2, i, k, j - 1) + A3(t2, i, k, j)) + A1(FL(fnp), k) * (A3(t2, i, k - 1, j - 1) + A3(t2, i, k - 1, j)));
}
__global__ void k_hdiff_v(const DevCtx xv, const float *t1, const float *t2, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *zx = FL(zx), *zy = FL(zy), *rdzw = FL(rdzw);
  const float zxv = 0.125f * (A3(zx, i, k, j) + A3(zx, i + 1, k, j) + A3(zx, i, k, j - 1) + A3(zx, i + 1, k, j - 1) + A3(zx, i, k + 1, j) +
                              A3(zx, i + 1, k + 1, j) + A3(zx, i, k + 1, j - 1) + A3(zx, i + 1, k + 1, j - 1));
  const float zyv = 0.5f * (A3(zy, i, k, j) + A3(zy, i, k + 1, j));
  const float msfvx = A2(FL(msfvx), i, j), msfvy = A2(FL(msfvy), i, j);
  const float mrdx = msfvx * x->rdx, mrdy = msfvy * x->rdy;
  const float tmpdz = (1.f / A3(rdzw, i, k, j) + 1.f / A3(rdzw, i, k, j - 1)) / 2.f;
  const float inc = x->g * tmpdz / A1(FL(dnw), k) *
                    (mrdy * (A3(t2, i, k, j) - A3(t2, i, k, j - 1)) + mrdx * (A3(t1, i + 1, k, j) - A3(t1, i, k, j)) -
                     msfvx * zxv * (hd_a1v(x, t1, i, k + 1, j) - hd_a1v(x, t1, i, k, j)) / tmpdz -
                     msfvy * zyv * (hd_a2v(x, t2, i, k + 1, j) - hd_a2v(x, t2, i, k, j)) / tmpdz);
  A3(FL(rv_tendf), i, k, j) = A3(FL(rv_tendf), i, k, j) + inc;
}

// horizontal_diffusion_w_2 (module_diffusion_em.py:2768-2985):
// w points i [2, ide-2], j [2, jde-2], k [2, kde-1].
__global__ void k_hdiff_w(const DevCtx xv, const float *t1, const float *t2, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *zx = FL(zx), *zy = FL(zy), *rdz = FL(rdz);
#define HA1(kk) (0.25f * (A3(t1, i + 1, (kk) + 1, j) + A3(t1, i, (kk) + 1, j) + A3(t1, i + 1, kk, j) + A3(t1, i, kk, j)))
#define HA2(kk) (0.25f * (A3(t2, i, (kk) + 1, j + 1) + A3(t2, i, (kk) + 1, j) + A3(t2, i, kk, j + 1) + A3(t2, i, kk, j)))
  const float zxw = 0.5f * (A3(zx, i, k, j) + A3(zx, i + 1, k, j));
  const float zyw = 0.5f * (A3(zy, i, k, j) + A3(zy, i, k, j + 1));
  const float mrdx = A2(FL(msftx), i, j) * x->rdx, mrdy = A2(FL(msfty), i, j) * x->rdy;
  const float rz = A3(rdz, i, k, j);
  const float inc = x->g / (A1(FL(dn), k) * rz) *
                    (mrdx * (A3(t1, i + 1, k, j) - A3(t1, i, k, j)) + mrdy * (A3(t2, i, k, j + 1) - A3(t2, i, k, j)) -
                     A2(FL(msfty), i, j) * rz * (zxw * (HA1(k) - HA1(k - 1)) + zyw * (HA2(k) - HA2(k - 1))));
#undef HA1
#undef HA2
  A3(FL(rw_tendf), i, k, j) = A3(FL(rw_tendf), i, k, j) + inc;
}

// horizontal_diffusion_s (module_diffusion_em.py:2046-2291) of every
// diffused scalar (theta, moist, scalars), in two passes over all of them:
// the face fluxes H1 (u faces) and H2 (v faces), then the tendency of the
// mass points i [2, ide-2], j [2, jde-2], k [1, kde-1] from H1/H2 at levels
// k-1, k, k+1.  The face geometry (diffusivity x density, slope, inverse
// layer depth) and the tendency geometry do not depend on the scalar and are
// formed once per point for all of them.  Every value is the reference's
// expression in the reference's operation order.

// The level averages of a scalar on the two columns A and B of a face, from
// the values at levels k-1, k, k+1, k+2 of each column (qa[0..3], qb[0..3]):
// at level k (surface extrapolation at k = 1) and at level k+1 (the top
// extrapolation at kde); hs_hav's three forms with the same arithmetic.
__device__ __forceinline__ float hs_hav_dn(const DevCtx *__restrict__ x, const float qa[4], const float qb[4], int k, float fm, float fp) {
  if (k == 1) return 0.5f * (x->cf1 * qa[1] + x->cf2 * qa[2] + x->cf3 * qa[3] + x->cf1 * qb[1] + x->cf2 * qb[2] + x->cf3 * qb[3]);
  return 0.5f * (fm * (qb[1] + qa[1]) + fp * (qb[0] + qa[0]));
}
__device__ __forceinline__ float hs_hav_up(const DevCtx *__restrict__ x, const float qa[4], const float qb[4], int k, float fm, float fp,
                                           float dnw1, float dn1) {
  if (k + 1 == x->kde)
    return 0.5f * (qa[1] + (qa[1] - qa[0]) * 0.5f * dnw1 / dn1 + qb[1] + (qb[1] - qb[0]) * 0.5f * dnw1 / dn1);
  return 0.5f * (fm * (qb[2] + qa[2]) + fp * (qb[1] + qa[1]));
}

// scalar-independent part of one face at level k: -msf * xk * rho (c),
// the face slope averaged to the level (tz) and the inverse layer depth at
// the face (rdz), for the face between columns A (a) and B (b)
struct HsFace {
  float c, tz, rdz;
};
__device__ __forceinline__ HsFace hs_face(const DevCtx *__restrict__ x, const float *zs, float msf, int ia, int ja, int ib, int jb, int k) {
  const float *xkhh = FL(xkhh), *rho = FL(rho), *rdzw = FL(rdzw);
  HsFace f;
  const float xk = 0.5f * (A3(xkhh, ib, k, jb) + A3(xkhh, ia, k, ja)) * 0.5f * (A3(rho, ib, k, jb) + A3(rho, ia, k, ja));
  f.c = -msf * xk;
  f.tz = 0.5f * (A3(zs, ia, k, ja) + A3(zs, ia, k + 1, ja));
  f.rdz = 2.f / (1.f / A3(rdzw, ia, k, ja) + 1.f / A3(rdzw, ib, k, jb));
  return f;
}
// H1 / H2 of one face: c * (rd * (var_A - var_B) - tz * (hav(k+1) - hav(k)) * rdz)
__device__ __forceinline__ float hs_flux(const HsFace &f, float rd, float va, float vb, float hav_up, float hav_dn) {
  return f.c * (rd * (va - vb) - f.tz * (hav_up - hav_dn) * f.rdz);
}

constexpr int kMaxDiff = 17;  // theta and up to 16 advected species
struct HdsList {
  const float *var[kMaxDiff];
  float *tend[kMaxDiff];
  float *h1[kMaxDiff], *h2[kMaxDiff];  // face fluxes (scratch)
  int n;
};

// pass 1: H1 at u faces i [2, ide-1], j [2, jde-2] and H2 at v faces
// i [2, ide-2], j [2, jde-1] of every scalar; r = i [2, ide-1], k [1, kde-1],
// j [2, jde-1]
__global__ void __launch_bounds__(256) k_hdiff_s_faces(const DevCtx xv, HdsList L, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const int ide = x->ide, jde = x->jde;
  const bool f1 = (i <= ide - 1 && j <= jde - 2), f2 = (i <= ide - 2 && j <= jde - 1);
  const unsigned o = XI3(i, k, j);
  HsFace a{0.f, 0.f, 0.f}, b{0.f, 0.f, 0.f};
  if (f1) a = hs_face(x, FL(zx), A2(FL(msfux), i, j), i, j, i - 1, j, k);
  if (f2) b = hs_face(x, FL(zy), A2(FL(msfvy), i, j), i, j, i, j - 1, k);
  const int kde = x->kde, k1 = imin_(k + 1, kde);
  const float fm0 = A1(FL(fnm), k), fp0 = A1(FL(fnp), k), fm1 = A1(FL(fnm), k1), fp1 = A1(FL(fnp), k1);
  const float dnw1 = A1(FL(dnw), kde - 1), dn1 = A1(FL(dn), kde - 1);
  const unsigned nb = (unsigned)x->n3, sk = (unsigned)x->ip, sj = (unsigned)(kde * x->ip);
  for (int s = 0; s < L.n; s++) {
    // levels k-1 .. k+2 of the point's column and of its west and south
    // neighbours, all loaded first (buffer loads: the levels outside the
    // column read 0 and are not used)
    const BufF bq = bufv(L.var[s], nb);
    float qc[4], qw[4], qs[4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const unsigned on = o + (unsigned)(n - 1) * sk;
      qc[n] = bld(bq, on);
      qw[n] = bld(bq, on - 1u);
      qs[n] = bld(bq, on - sj);
    }
#pragma unroll
    for (int n = 0; n < 4; n++) {
      pin1(qc[n]);
      pin1(qw[n]);
      pin1(qs[n]);
    }
    if (f1) elem(L.h1[s], o) = hs_flux(a, x->rdx, qc[1], qw[1], hs_hav_up(x, qc, qw, k, fm1, fp1, dnw1, dn1), hs_hav_dn(x, qc, qw, k, fm0, fp0));
    if (f2) elem(L.h2[s], o) = hs_flux(b, x->rdy, qc[1], qs[1], hs_hav_up(x, qc, qs, k, fm1, fp1, dnw1, dn1), hs_hav_dn(x, qc, qs, k, fm0, fp0));
  }
}

// pass 2: tendency of every scalar on mass points i [2, ide-2], j [2, jde-2],
// k [1, kde-1] (r)
__global__ void __launch_bounds__(256) k_hdiff_s_tend(const DevCtx xv, HdsList L, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *zx = FL(zx), *zy = FL(zy), *rdzw = FL(rdzw), *fnm = FL(fnm), *fnp = FL(fnp);
  const int ktf = x->kde - 1;
  const float zxm = 0.25f * (A3(zx, i, k, j) + A3(zx, i + 1, k, j) + A3(zx, i, k + 1, j) + A3(zx, i + 1, k + 1, j));
  const float zym = 0.25f * (A3(zy, i, k, j) + A3(zy, i, k, j + 1) + A3(zy, i, k + 1, j) + A3(zy, i, k + 1, j + 1));
  const float msftx = A2(FL(msftx), i, j), msfty = A2(FL(msfty), i, j);
  const float mrdx = msftx * x->rdx, mrdy = msfty * x->rdy;
  const float rz = A3(rdzw, i, k, j);
  const float fac = x->g / (A1(FL(dnw), k) * rz);
  const float cx = msftx * zxm, cy = msfty * zym;
  const bool lo = (k != 1), hi = (k + 1 != ktf + 1);  // H*A(k), H*A(k+1) nonzero
  const float fm0 = A1(fnm, k), fp0 = A1(fnp, k), fm1 = A1(fnm, k + 1), fp1 = A1(fnp, k + 1);
  const unsigned o = XI3(i, k, j), sk = (unsigned)x->ip, sj = (unsigned)(x->kde * x->ip);
  const unsigned nb = (unsigned)x->n3;
  for (int s = 0; s < L.n; s++) {
    // every load of this scalar first (buffer loads: the level below 1 and
    // above kde-1 read 0 and are not used)
    const BufF b1 = bufv(L.h1[s], nb), b2 = bufv(L.h2[s], nb);
    float *t = L.tend[s];
    const float a0 = bld(b1, o), a1 = bld(b1, o + 1u), b0 = bld(b2, o), b1v = bld(b2, o + sj);
    const float a0m = bld(b1, o - sk), a1m = bld(b1, o + 1u - sk), a0p = bld(b1, o + sk), a1p = bld(b1, o + 1u + sk);
    const float b0m = bld(b2, o - sk), b1m = bld(b2, o + sj - sk), b0p = bld(b2, o + sk), b1p = bld(b2, o + sj + sk);
    const float t0 = elem(t, o);
    pin1(a0m);
    pin1(a1m);
    pin1(a0p);
    pin1(a1p);
    pin1(b0m);
    pin1(b1m);
    pin1(b0p);
    pin1(b1p);
    const float h1a0 = lo ? 0.5f * (fm0 * (a1 + a0) + fp0 * (a1m + a0m)) : 0.f;
    const float h1a1 = hi ? 0.5f * (fm1 * (a1p + a0p) + fp1 * (a1 + a0)) : 0.f;
    const float h2a0 = lo ? 0.5f * (fm0 * (b1v + b0) + fp0 * (b1m + b0m)) : 0.f;
    const float h2a1 = hi ? 0.5f * (fm1 * (b1p + b0p) + fp1 * (b1v + b0)) : 0.f;
    const float inc = fac * (mrdx * (a1 - a0) + mrdy * (b1v - b0) - cx * (h1a1 - h1a0) * rz - cy * (h2a1 - h2a0) * rz);
    elem(t, o) = t0 + inc;
  }
}
