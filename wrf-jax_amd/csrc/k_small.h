// Acoustic (small) step kernels: small_step_prep, calc_p_rho, calc_coef_w,
// advance_uv, advance_mu_t, advance_w, sumflux, small_step_finish, and the
// species update of the RK stage (rk_update_scalar, rk_update_scalar_pd).
// Semantics: src/wrf_jax/dyn_em/module_small_step_em.py and module_em.py,
// argument mapping of the stage adapters in src/wrf_jax/runtime.py.
#pragma once
#include "wrfdc_common.h"

// small_step_prep, 2-D part (module_small_step_em.py:272-420): stage-start
// column masses.  i [1, ide], j [1, jde].
__global__ void k_ssp_mass(const DevCtx xv, int rk) {
  const DevCtx *__restrict__ x = &xv;
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = 1 + (int)blockIdx.y;
  const int ide = x->ide, jde = x->jde;
  if (i > ide || j > jde) return;
  const float *mub = FL(mub), *mu_1 = FL(mu_1);
  float *mu_2 = FL(mu_2);
  const bool mass = (i <= ide - 1 && j <= jde - 1);
  if (rk == 1) {
    if (j <= jde - 1) A2(FL(muus), i, j) = A2(FL(muu), i, j);
    if (i <= ide - 1) A2(FL(muvs), i, j) = A2(FL(muv), i, j);
    if (mass) {
      const float m2 = A2(mu_2, i, j);
      A2(FL(mu_1), i, j) = m2;
      A2(FL(mudf), i, j) = 0.f;
      A2(FL(muts), i, j) = A2(mub, i, j) + m2;
      A2(FL(mu_save), i, j) = m2;
      A2(mu_2, i, j) = 0.f;
    }
  } else {
    if (j <= jde - 1) {
      const int iw = (i == 1) ? 1 : i - 1;
      float s = A2(mub, i, j) + A2(mu_1, i, j);
      s = s + A2(mub, iw, j);
      s = s + A2(mu_1, iw, j);
      A2(FL(muus), i, j) = 0.5f * s;
    }
    if (i <= ide - 1) {
      const int js = (j == 1) ? 1 : j - 1;
      float s = A2(mub, i, j) + A2(mu_1, i, j);
      s = s + A2(mub, i, js);
      s = s + A2(mu_1, i, js);
      A2(FL(muvs), i, j) = 0.5f * s;
    }
    if (mass) {
      A2(FL(muts), i, j) = A2(mub, i, j) + A2(mu_1, i, j);
      const float m2 = A2(mu_2, i, j);
      A2(FL(mu_save), i, j) = m2;
      A2(mu_2, i, j) = A2(mu_1, i, j) - m2;
    }
  }
}

// small_step_prep, 3-D part (module_small_step_em.py:420-717): save the
// stage-start state, form the acoustic perturbation variables and c2a.
// Full extent i [1, ide], k [1, kde], j [1, jde].
// Every field of the point is loaded first (all addresses are points of the
// full-extent fields), then the stores of the staggerings the point belongs
// to; RK1: the first RK stage (stage-start copies, perturbations zero).
template <bool RK1>
__global__ void k_ssp_3d(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const unsigned o = XI3(i, k, j), o2 = XI2(i, j);
  const bool up = (j <= jde - 1 && k <= kde - 1), vp = (i <= ide - 1 && k <= kde - 1), mp = (i <= ide - 1 && j <= jde - 1);
  const bool tp = mp && k <= kde - 1;
  const int kh = imin_(k, kde - 1);  // in-range index of the half-level coefficients
  const float c1h = A1(FL(c1h), kh), c2h = A1(FL(c2h), kh), c1f = A1(FL(c1f), k), c2f = A1(FL(c2f), k);
  const float pb = elem(FL(pb), o), pp = elem(FL(p), o), alt = elem(FL(alt), o);
  const float u2 = elem(FL(u_2), o), v2 = elem(FL(v_2), o), t2 = elem(FL(t_2), o), w2 = elem(FL(w_2), o), ph2 = elem(FL(ph_2), o);
  const float wwv = elem(FL(ww), o);
  float u1 = u2, v1 = v2, t1 = t2, w1 = w2, ph1 = ph2;
  float muus = 0.f, muu = 0.f, msfuy = 1.f, muvs = 0.f, muv = 0.f, msfvxi = 1.f, mts = 0.f, mt = 0.f, msfty = 1.f;
  if (!RK1) {
    u1 = elem(FL(u_1), o);
    v1 = elem(FL(v_1), o);
    t1 = elem(FL(t_1), o);
    w1 = elem(FL(w_1), o);
    ph1 = elem(FL(ph_1), o);
    muus = elem(FL(muus), o2);
    muu = elem(FL(muu), o2);
    msfuy = elem(FL(msfuy), o2);
    muvs = elem(FL(muvs), o2);
    muv = elem(FL(muv), o2);
    msfvxi = elem(FL(msfvx_inv), o2);
    mts = elem(FL(muts), o2);
    mt = elem(FL(mut), o2);
    msfty = elem(FL(msfty), o2);
  }
  pin1(pb);
  pin1(pp);
  pin1(alt);
  if (tp) elem(FL(c2a), o) = (x->cpovcv * (pb + pp)) / alt;
  if (up) {  // u points
    if (RK1) elem(FL(u_1), o) = u2;
    elem(FL(u_save), o) = u2;
    elem(FL(u_2), o) = RK1 ? 0.f : ((c1h * muus + c2h) * u1 - (c1h * muu + c2h) * u2) / msfuy;
  }
  if (vp) {  // v points
    if (RK1) elem(FL(v_1), o) = v2;
    elem(FL(v_save), o) = v2;
    elem(FL(v_2), o) = RK1 ? 0.f : ((c1h * muvs + c2h) * v1 - (c1h * muv + c2h) * v2) * msfvxi;
  }
  if (mp) {
    if (k <= kde - 1) {  // theta
      if (RK1) elem(FL(t_1), o) = t2;
      elem(FL(t_save), o) = t2;
      elem(FL(t_2), o) = RK1 ? 0.f : (c1h * mts + c2h) * t1 - (c1h * mt + c2h) * t2;
    }
    // w / ph / ww levels
    if (RK1) elem(FL(w_1), o) = w2;
    elem(FL(w_save), o) = w2;
    elem(FL(w_2), o) = RK1 ? 0.f : ((c1f * mts + c2f) * w1 - (c1f * mt + c2f) * w2) / msfty;
    if (RK1) elem(FL(ph_1), o) = ph2;
    elem(FL(ph_save), o) = ph2;
    elem(FL(ph_2), o) = ph1 - ph2;
    elem(FL(ww_1), o) = wwv;
  }
}

// calc_p_rho, non-hydrostatic (module_small_step_em.py:720-1034) with
// ph = ph_2, t_1 = t_save, mu = mu_2, mut = muts; mass points.
__global__ void k_calc_p_rho(const DevCtx xv, int step, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const unsigned o = XI3(i, k, j);
  // every load first (the stores below could alias them as far as the
  // compiler knows)
  const float muts = A2(FL(muts), i, j), mu2 = A2(FL(mu_2), i, j);
  const float alt = elem(FL(alt), o), t1 = elem(FL(t_save), o), t2 = elem(FL(t_2), o), c2a = elem(FL(c2a), o);
  const float ph1 = elem(FL(ph_2), o + (unsigned)x->ip), ph0 = elem(FL(ph_2), o);
  const float pm1 = step == 0 ? 0.f : elem(FL(pm1), o);
  const float mass = A1(FL(c1h), k) * muts + A1(FL(c2h), k);
  const float dry = A1(FL(c1h), k) * mu2;
  const float bracket = alt * dry + A1(FL(rdnw), k) * (ph1 - ph0);
  const float alv = (-1.f / mass) * bracket;
  const float tdel = t2 - dry * t1;
  const float frac = (alt * tdel) / (mass * (x->t0 + t1));
  const float pr = c2a * (frac - alv);
  elem(FL(al), o) = alv;
  if (step == 0) elem(FL(p), o) = pr;
  else elem(FL(p), o) = pr + x->smdiv * (pr - pm1);
  elem(FL(pm1), o) = pr;
}

// calc_coef_w (module_small_step_em.py:1359-1586): LU factors of the
// implicit w-phi system, one thread per mass column.
__global__ void k_calc_coef_w(const DevCtx xv, float dts, R2 r) {
  const DevCtx *__restrict__ x = &xv;
  // the column coefficients in LDS: read per level inside the recurrence, and
  // from global memory every such read would wait behind the previous level's
  // stores (the compiler cannot prove they do not alias)
  constexpr int kMaxLev = 152;  // kde <= 151 (wrfdc_create checks nz <= 150)
  __shared__ float c1h[kMaxLev], c2h[kMaxLev], c1f[kMaxLev], c2f[kMaxLev], rdn[kMaxLev], rdnw[kMaxLev];
  const int kde = x->kde;
  for (int q = (int)threadIdx.x; q < kde; q += (int)blockDim.x) {
    c1h[q] = FL(c1h)[q];
    c2h[q] = FL(c2h)[q];
    c1f[q] = FL(c1f)[q];
    c2f[q] = FL(c2f)[q];
    rdn[q] = FL(rdn)[q];
    rdnw[q] = FL(rdnw)[q];
  }
  __syncthreads();
  GET_IJ(r);
  float *a = FL(a), *alpha = FL(alpha), *gamma = FL(gamma);
  const float *cqw = FL(cqw), *c2a = FL(c2a);
  const float lid = x->top_lid ? 0.f : 1.f;
  float cof = 0.5f * dts * x->g * (1.f + x->epssm);
  cof = cof * cof;
  const float m = A2(FL(mut), i, j);
  const float c2a_top = A3(c2a, i, kde - 1, j);
  const float atop = (-2.f * cof * (A1(rdnw, kde - 1) * A1(rdnw, kde - 1)) * c2a_top * lid) /
                     ((A1(c1h, kde - 1) * m + A1(c2h, kde - 1)) * (A1(c1f, kde - 1) * m + A1(c2f, kde - 1)));
  A3(a, i, 2, j) = 0.f;
  A3(a, i, kde, j) = atop;
  A3(gamma, i, 1, j) = 0.f;
  float gprev = 0.f;
  // levels in chunks of 8: the chunk's cqw and c2a are read before its
  // stores (which the compiler cannot move later loads past), so a column
  // waits for memory once per chunk instead of once per level
  for (int k0 = 2; k0 <= kde - 1; k0 += 8) {
    float r_cqw[8], r_c2a[8], r_c2am[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = imin_(k0 + q, kde - 1);
      r_cqw[q] = A3(cqw, i, k, j);
      r_c2a[q] = A3(c2a, i, k, j);
      r_c2am[q] = A3(c2a, i, k - 1, j);
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = k0 + q;
      if (k > kde - 1) break;
      const float mh = A1(c1h, k) * m + A1(c2h, k), mf = A1(c1f, k) * m + A1(c2f, k), mhb = A1(c1h, k - 1) * m + A1(c2h, k - 1);
      const float mfa = A1(c1f, k + 1) * m + A1(c2f, k + 1);
      float ak = 0.f;
      if (k >= 3) {
        ak = (-r_cqw[q] * cof * A1(rdn, k) * A1(rdnw, k - 1) * r_c2am[q]) / ((A1(c1h, k - 1) * m + A1(c2h, k - 1)) * (A1(c1f, k - 1) * m + A1(c2f, k - 1)));
        A3(a, i, k, j) = ak;
      }
      const float bu = (A1(rdnw, k) * r_c2a[q]) / (mh * mf);
      const float bl = (A1(rdnw, k - 1) * r_c2am[q]) / (mhb * mf);
      const float b = 1.f + r_cqw[q] * cof * A1(rdn, k) * (bu + bl);
      const float c = (-r_cqw[q] * cof * A1(rdn, k) * A1(rdnw, k) * r_c2a[q]) / (mh * mfa);
      const float al_ = 1.f / (b - ak * gprev);
      const float ga = c * al_;
      A3(alpha, i, k, j) = al_;
      A3(gamma, i, k, j) = ga;
      gprev = ga;
    }
  }
  const float btop = 1.f + (2.f * cof * (A1(rdnw, kde - 1) * A1(rdnw, kde - 1)) * c2a_top) /
                               ((A1(c1h, kde - 1) * m + A1(c2h, kde - 1)) * (A1(c1f, kde) * m + A1(c2f, kde)));
  A3(alpha, i, kde, j) = 1.f / (btop - atop * gprev);
  A3(gamma, i, kde, j) = 0.f;
}

// interface pressure of a u/v face (module_small_step_em.py:1627-1645)
__device__ __forceinline__ float uv_iface(const DevCtx *__restrict__ x, const float *p, int i, int j, int di, int dj, int level) {
#define PP(kk) (A3(p, i, kk, j) + A3(p, i - di, kk, j - dj))
  if (level == 1) return 0.5f * (x->cf1 * PP(1) + x->cf2 * PP(2) + x->cf3 * PP(3));
  if (level >= 2 && level <= x->kde - 1) return 0.5f * (A1(FL(fnm), level) * PP(level) + A1(FL(fnp), level) * PP(level - 1));
#undef PP
  return 0.f;
}

// advance_uv (module_small_step_em.py:1589-1987), specified boundaries.
// axis 0: u on i [2, ide-1], j [2, jde-2]; axis 1: v on i [2, ide-2], j [2, jde-1].
__global__ void k_advance_uv(const DevCtx xv, int axis, float dts, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
  float *vel = axis == 0 ? FL(u_2) : FL(v_2);
  const float *tend = axis == 0 ? FL(ru_tend) : FL(rv_tend);
  const float *mass = axis == 0 ? FL(muu) : FL(muv);
  const float *cq = axis == 0 ? FL(cqu) : FL(cqv);
  const float *p = FL(p), *pb = FL(pb), *ph = FL(ph_2), *alt = FL(alt), *al = FL(al), *php = FL(php), *mu = FL(mu_2), *mudf = FL(mudf);
  const float *c1h = FL(c1h), *c2h = FL(c2h);
  const float spacing = axis == 0 ? x->rdx : x->rdy;
  const float grid_length = 1.f / spacing;
  const size_t o = XI3(i, k, j);
  const float adv = vel[o] + dts * tend[o];
  float damping = -1.f * x->emdiv * grid_length * (A2(mudf, i, j) - A2(mudf, i - di, j - dj));
  damping = axis == 0 ? damping / A2(FL(msfuy), i, j) : damping * A2(FL(msfvx_inv), i, j);
  const float phi = (A3(ph, i, k + 1, j) - A3(ph, i - di, k + 1, j - dj)) + (A3(ph, i, k, j) - A3(ph, i - di, k, j - dj));
  const float pres = (A3(alt, i, k, j) + A3(alt, i - di, k, j - dj)) * (A3(p, i, k, j) - A3(p, i - di, k, j - dj));
  const float base = (A3(al, i, k, j) + A3(al, i - di, k, j - dj)) * (A3(pb, i, k, j) - A3(pb, i - di, k, j - dj));
  const float term = (phi + pres) + base;
  const float ratio = axis == 0 ? A2(FL(msfux), i, j) / A2(FL(msfuy), i, j) : A2(FL(msfvy), i, j) / A2(FL(msfvx), i, j);
  const float m = A1(c1h, k) * A2(mass, i, j) + A1(c2h, k);
  float grad = (ratio * 0.5f * spacing * m) * term;
  const float vert = A1(FL(rdnw), k) * (uv_iface(x, p, i, j, di, dj, k + 1) - uv_iface(x, p, i, j, di, dj, k));
  const float muv_ = 0.5f * (A1(c1h, k) * A2(mu, i - di, j - dj) + A1(c1h, k) * A2(mu, i, j));
  grad = grad + (ratio * spacing * (A3(php, i, k, j) - A3(php, i - di, k, j - dj))) * (vert - muv_);
  vel[o] = (adv - dts * cq[o] * grad) + A1(c1h, k) * damping;
}

// horizontal mass divergence of one level for advance_mu_t
__device__ __forceinline__ float mut_dvdxi(const DevCtx *__restrict__ x, int i, int k, int j) {
  const float *u = FL(u_2), *u1 = FL(u_save), *v = FL(v_2), *v1 = FL(v_save), *muu = FL(muu), *muv = FL(muv);
  const float *msfuy = FL(msfuy), *msfvx_inv = FL(msfvx_inv);
  const float c1 = A1(FL(c1h), k), c2 = A1(FL(c2h), k);
  const float vn = A3(v, i, k, j + 1) + ((c1 * A2(muv, i, j + 1) + c2) * A3(v1, i, k, j + 1)) * A2(msfvx_inv, i, j + 1);
  const float vs = A3(v, i, k, j) + ((c1 * A2(muv, i, j) + c2) * A3(v1, i, k, j)) * A2(msfvx_inv, i, j);
  const float ue = A3(u, i + 1, k, j) + ((c1 * A2(muu, i + 1, j) + c2) * A3(u1, i + 1, k, j)) / A2(msfuy, i + 1, j);
  const float uw = A3(u, i, k, j) + ((c1 * A2(muu, i, j) + c2) * A3(u1, i, k, j)) / A2(msfuy, i, j);
  const float hf = x->rdy * (vn - vs) + x->rdx * (ue - uw);
  return (A2(FL(msftx), i, j) * A2(FL(msfty), i, j)) * hf;
}

// advance_mu_t (module_small_step_em.py:2024-2403) with u_1 = u_save,
// v_1 = v_save, t_1 = t_save, t_ave = t_2ave, ft = t_tend; one thread per
// column i [2, ide-2], j [2, jde-2].
__global__ void k_advance_mu_t(const DevCtx xv, float dts, R2 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IJ(r);
  float *ww = FL(ww), *t = FL(t_2), *t_ave = FL(t_2ave);
  const float *ww_1 = FL(ww_1), *t_1 = FL(t_save), *ft = FL(t_tend), *u = FL(u_2), *v = FL(v_2);
  const float *c1h = FL(c1h), *dnw = FL(dnw), *fnm = FL(fnm), *fnp = FL(fnp), *rdnw = FL(rdnw);
  const int k_end = x->kde - 1;
  // column mass tendency: FP32 level terms summed in FP64, rounded once
  double dmdt_acc = 0.0;
  for (int k = 1; k <= k_end; k++) dmdt_acc += (double)(A1(dnw, k) * mut_dvdxi(x, i, k, j));
  const float dmdt = (float)dmdt_acc;
  float *mu = FL(mu_2);
  const float old = A2(mu, i, j);
  const float mtend = A2(FL(mu_tend), i, j);
  const float tend = dmdt + mtend;
  const float nmu = old + dts * tend;
  A2(mu, i, j) = nmu;
  A2(FL(mudf), i, j) = tend;
  A2(FL(muts), i, j) = A2(FL(mut), i, j) + nmu;
  A2(FL(muave), i, j) = 0.5f * ((1.f + x->epssm) * nmu + (1.f - x->epssm) * old);
  // omega recurrence on the stage-start omega, then remove it
  const float msfty = A2(FL(msfty), i, j);
  float prev = A3(ww, i, 1, j);
  A3(ww, i, 1, j) = prev - A3(ww_1, i, 1, j);
  for (int k = 2; k <= k_end; k++) {
    float integ = A1(c1h, k - 1) * dmdt + mut_dvdxi(x, i, k - 1, j);
    integ = integ + A1(c1h, k - 1) * mtend;
    const float cur = prev - (A1(dnw, k - 1) * integ) / msfty;
    A3(ww, i, k, j) = cur - A3(ww_1, i, k, j);
    prev = cur;
  }
  // theta
  const float msftx = A2(FL(msftx), i, j);
  const float rdx = x->rdx, rdy = x->rdy;
  float wlo = 0.f;
  for (int k = 1; k <= k_end; k++) {
    const float whi = (k + 1 <= k_end) ? A3(ww, i, k + 1, j) * (A1(fnm, k + 1) * A3(t_1, i, k + 1, j) + A1(fnp, k + 1) * A3(t_1, i, k, j)) : 0.f;
    const float tk = A3(t, i, k, j);
    A3(t_ave, i, k, j) = tk;
    const float ta = tk + (msfty * dts) * A3(ft, i, k, j);
    const float vnorth = A3(v, i, k, j + 1) * (A3(t_1, i, k, j + 1) + A3(t_1, i, k, j));
    const float vsouth = A3(v, i, k, j) * (A3(t_1, i, k, j) + A3(t_1, i, k, j - 1));
    const float vft = (0.5f * rdy) * (vnorth - vsouth);
    const float ueast = A3(u, i + 1, k, j) * (A3(t_1, i + 1, k, j) + A3(t_1, i, k, j));
    const float uwest = A3(u, i, k, j) * (A3(t_1, i, k, j) + A3(t_1, i - 1, k, j));
    const float uft = (0.5f * rdx) * (ueast - uwest);
    const float horiz = msftx * (vft + uft);
    const float vert = A1(rdnw, k) * (whi - wlo);
    A3(t, i, k, j) = ta - (dts * msfty) * (horiz + vert);
    wlo = whi;
  }
}

// advance_w (module_small_step_em.py:2406-3131) with t_1 = t_save,
// ph_1 = ph_save; one thread per column i [2, ide-2], j [2, jde-2].  The
// right-hand side of the phi equation is staged in the scratch array `rhs`
// (w levels); the forward sweep of the tridiagonal solve runs in w itself.
__global__ void k_advance_w(const DevCtx xv, float dts, float *__restrict__ rhs, R2 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IJ(r);
  float *w = FL(w_2), *ph = FL(ph_2), *t_2ave = FL(t_2ave);
  const float *rw_tend = FL(rw_tend), *ww = FL(ww), *u = FL(u_2), *v = FL(v_2), *t_2 = FL(t_2), *t_1 = FL(t_save);
  const float *ph_1 = FL(ph_save), *phb = FL(phb), *ph_tend = FL(ph_tend), *ht = FL(ht), *c2a = FL(c2a), *cqw = FL(cqw), *alt = FL(alt);
  const float *a = FL(a), *alpha = FL(alpha), *gamma = FL(gamma);
  const float *c1h = FL(c1h), *c2h = FL(c2h), *c1f = FL(c1f), *c2f = FL(c2f), *fnm = FL(fnm), *fnp = FL(fnp), *rdnw = FL(rdnw), *rdn = FL(rdn);
  const float g = x->g, epssm = x->epssm, t0 = x->t0;
  const float ope = 1.f + epssm, ome = 1.f - epssm;
  const int k_end = x->kde - 1, kde = x->kde;
  const float m = A2(FL(mut), i, j), mts = A2(FL(muts), i, j), mav = A2(FL(muave), i, j);
  const float msfty = A2(FL(msfty), i, j), msft_inv = 1.f / msfty;
  for (int k = 1; k <= k_end; k++) {
    float ta = 0.5f * (ope * A3(t_2, i, k, j) + ome * A3(t_2ave, i, k, j));
    ta = (ta + (A1(c1h, k) * mav) * t0) / ((A1(c1h, k) * mts + A1(c2h, k)) * (t0 + A3(t_1, i, k, j)));
    A3(t_2ave, i, k, j) = ta;
  }
  // rhs of the phi equation, levels 2..kde
#define WDWN(kk)                                                                                         \
  (((0.5f * (A3(ww, i, kk, j) + A3(ww, i, (kk) - 1, j))) * A1(rdnw, (kk) - 1)) *                         \
   (A3(ph_1, i, kk, j) - A3(ph_1, i, (kk) - 1, j) + A3(phb, i, kk, j) - A3(phb, i, (kk) - 1, j)))
  A3(rhs, i, 1, j) = 0.f;
  for (int k = 2; k <= k_end + 1; k++) {
    float rk = dts * (A3(ph_tend, i, k, j) + ((0.5f * g) * ome) * A3(w, i, k, j));
    if (k <= k_end) rk = rk - dts * (A1(fnm, k) * WDWN(k + 1) + A1(fnp, k) * WDWN(k));
    rk = A3(ph, i, k, j) + (msfty * rk) / (A1(c1f, k) * m + A1(c2f, k));
    if (k == k_end + 1 && x->top_lid) rk = 0.f;
    A3(rhs, i, k, j) = rk;
  }
#undef WDWN
  // surface w from the terrain slope
  const float cf1 = x->cf1, cf2 = x->cf2, cf3 = x->cf3;
  const float lv = (A2(ht, i, j + 1) - A2(ht, i, j)) * (cf1 * A3(v, i, 1, j + 1) + cf2 * A3(v, i, 2, j + 1) + cf3 * A3(v, i, 3, j + 1)) +
                   (A2(ht, i, j) - A2(ht, i, j - 1)) * (cf1 * A3(v, i, 1, j) + cf2 * A3(v, i, 2, j) + cf3 * A3(v, i, 3, j));
  const float lu = (A2(ht, i + 1, j) - A2(ht, i, j)) * (cf1 * A3(u, i + 1, 1, j) + cf2 * A3(u, i + 1, 2, j) + cf3 * A3(u, i + 1, 3, j)) +
                   (A2(ht, i, j) - A2(ht, i - 1, j)) * (cf1 * A3(u, i, 1, j) + cf2 * A3(u, i, 2, j) + cf3 * A3(u, i, 3, j));
  float wprev = ((msfty * 0.5f) * x->rdy) * lv + ((A2(FL(msftx), i, j) * 0.5f) * x->rdx) * lu;
  A3(w, i, 1, j) = wprev;
  // explicit w and forward elimination
  for (int k = 2; k <= k_end; k++) {
    const float rkp = A3(rhs, i, k + 1, j), rk0 = A3(rhs, i, k, j), rkm = A3(rhs, i, k - 1, j);
    const float pu = ((A3(c2a, i, k, j) * A1(rdnw, k)) / (A1(c1h, k) * m + A1(c2h, k))) *
                     (ope * (rkp - rk0) + ome * (A3(ph, i, k + 1, j) - A3(ph, i, k, j)));
    const float pl = ((A3(c2a, i, k - 1, j) * A1(rdnw, k - 1)) / (A1(c1h, k - 1) * m + A1(c2h, k - 1))) *
                     (ope * (rk0 - rkm) + ome * (A3(ph, i, k, j) - A3(ph, i, k - 1, j)));
    const float ppart = (((0.5f * dts) * g) * A1(rdn, k)) * (pu - pl);
    const float tu = (A3(c2a, i, k, j) * A3(alt, i, k, j)) * A3(t_2ave, i, k, j);
    const float tl = (A3(c2a, i, k - 1, j) * A3(alt, i, k - 1, j)) * A3(t_2ave, i, k - 1, j);
    const float bpart = A1(rdn, k) * (tu - tl) - A1(c1f, k) * mav;
    float wl = ((A3(w, i, k, j) + dts * A3(rw_tend, i, k, j)) + (msft_inv * A3(cqw, i, k, j)) * ppart) + ((dts * g) * msft_inv) * bpart;
    wl = (wl - A3(a, i, k, j) * wprev) * A3(alpha, i, k, j);
    A3(w, i, k, j) = wl;
    wprev = wl;
  }
  {
    const int k = kde;
    float fac = ((-0.5f * dts) * g) / (A1(c1h, k - 1) * m + A1(c2h, k - 1));
    fac = fac * (A1(rdnw, k - 1) * A1(rdnw, k - 1));
    fac = fac * 2.f;
    fac = fac * A3(c2a, i, k - 1, j);
    const float ptop = fac * (ope * (A3(rhs, i, k, j) - A3(rhs, i, k - 1, j)) + ome * (A3(ph, i, k, j) - A3(ph, i, k - 1, j)));
    const float btop = (-dts * g) * (((2.f * A1(rdnw, k - 1)) * A3(c2a, i, k - 1, j)) * A3(alt, i, k - 1, j) * A3(t_2ave, i, k - 1, j) + A1(c1f, k) * mav);
    float wl = A3(w, i, k, j) + dts * A3(rw_tend, i, k, j) + (1.f / msfty) * (ptop + btop);
    if (x->top_lid) wl = 0.f;
    wl = (wl - A3(a, i, k, j) * wprev) * A3(alpha, i, k, j);
    A3(w, i, k, j) = wl;
    wprev = wl;
  }
  // back substitution and the phi update
  const float pcoef = (((msfty * 0.5f) * dts) * g) * ope;
  A3(ph, i, kde, j) = A3(rhs, i, kde, j) + pcoef * wprev / (A1(c1f, kde) * mts + A1(c2f, kde));
  for (int k = k_end; k >= 2; k--) {
    const float wl = A3(w, i, k, j) - A3(gamma, i, k, j) * wprev;
    A3(w, i, k, j) = wl;
    A3(ph, i, k, j) = A3(rhs, i, k, j) + pcoef * wl / (A1(c1f, k) * mts + A1(c2f, k));
    wprev = wl;
  }
}

// sumflux (module_small_step_em.py:3134-3310) with ru = u_2, rv = v_2,
// u_lin = u_save, v_lin = v_save, ww_lin = ww_1; full extent so the first
// iteration also clears points outside the flux ranges.
__global__ void k_sumflux(const DevCtx xv, int first, int fin, float ns, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const size_t o = XI3(i, k, j);
  const float *c1h = FL(c1h), *c2h = FL(c2h);
  float *ru_m = FL(ru_m), *rv_m = FL(rv_m), *ww_m = FL(ww_m);
  if (j <= jde - 1 && k <= kde - 1) {
    float acc = (first ? 0.f : ru_m[o]) + FL(u_2)[o];
    if (fin) acc = acc / ns + ((A1(c1h, k) * A2(FL(muu), i, j) + A1(c2h, k)) * FL(u_save)[o]) / A2(FL(msfuy), i, j);
    ru_m[o] = acc;
  } else if (first) {
    ru_m[o] = 0.f;
  }
  if (i <= ide - 1 && k <= kde - 1) {
    float acc = (first ? 0.f : rv_m[o]) + FL(v_2)[o];
    if (fin) acc = acc / ns + ((A1(c1h, k) * A2(FL(muv), i, j) + A1(c2h, k)) * FL(v_save)[o]) * A2(FL(msfvx_inv), i, j);
    rv_m[o] = acc;
  } else if (first) {
    rv_m[o] = 0.f;
  }
  if (i <= ide - 1 && j <= jde - 1) {
    float acc = (first ? 0.f : ww_m[o]) + FL(ww)[o];
    if (fin) acc = acc / ns + FL(ww_1)[o];
    ww_m[o] = acc;
  } else if (first) {
    ww_m[o] = 0.f;
  }
}

// small_step_finish (module_small_step_em.py:1119-1332): recouple the
// acoustic perturbations onto the stage-start state.  Full extent.
// (loads first, as k_ssp_3d; FINAL: the last RK stage, which also removes
// the diabatic heating)
template <bool FINAL>
__global__ void k_ssf(const DevCtx xv, float dts, float nsm, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const unsigned o = XI3(i, k, j), o2 = XI2(i, j);
  const int kh = imin_(k, kde - 1);
  const float c1h = A1(FL(c1h), kh), c2h = A1(FL(c2h), kh), c1f = A1(FL(c1f), k), c2f = A1(FL(c2f), k);
  const float v2 = elem(FL(v_2), o), vs = elem(FL(v_save), o), u2 = elem(FL(u_2), o), us = elem(FL(u_save), o);
  const float w2 = elem(FL(w_2), o), ws = elem(FL(w_save), o), ph2 = elem(FL(ph_2), o), phs = elem(FL(ph_save), o);
  const float wwv = elem(FL(ww), o), ww1 = elem(FL(ww_1), o), t2 = elem(FL(t_2), o), ts = elem(FL(t_save), o);
  const float hd = FINAL ? elem(FL(h_diabatic), o) : 0.f;
  const float muv = elem(FL(muv), o2), muvs = elem(FL(muvs), o2), msfvx = elem(FL(msfvx), o2);
  const float muu = elem(FL(muu), o2), muus = elem(FL(muus), o2), msfuy = elem(FL(msfuy), o2);
  const float mt = elem(FL(mut), o2), mts = elem(FL(muts), o2), msfty = elem(FL(msfty), o2);
  if (i <= ide - 1 && k <= kde - 1) {
    const float old = c1h * muv + c2h, nw = c1h * muvs + c2h;
    elem(FL(v_2), o) = (msfvx * v2 + vs * old) / nw;
  }
  if (j <= jde - 1 && k <= kde - 1) {
    const float old = c1h * muu + c2h, nw = c1h * muus + c2h;
    elem(FL(u_2), o) = (msfuy * u2 + us * old) / nw;
  }
  if (i <= ide - 1 && j <= jde - 1) {
    {
      const float old = c1f * mt + c2f, nw = c1f * mts + c2f;
      elem(FL(w_2), o) = (msfty * w2 + ws * old) / nw;
      elem(FL(ph_2), o) = ph2 + phs;
      elem(FL(ww), o) = wwv + ww1;
    }
    if (k <= kde - 1) {
      const float old = c1h * mt + c2h, nw = c1h * mts + c2h;
      float tt = t2;
      if (FINAL) tt = tt - ((dts * nsm) * old) * hd;
      elem(FL(t_2), o) = (tt + ts * old) / nw;
    }
    if (k == 1) A2(FL(mu_2), i, j) = A2(FL(mu_2), i, j) + A2(FL(mu_save), i, j);
  }
}

// rk_update_scalar_pd (module_em.py:1601-1669): last-stage physics update of
// the old scalar with mu_old = mu_new = mu_1, then the tendency is cleared.
__global__ void k_update_scalar_pd(const DevCtx xv, float *q_old, float *q_tend, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const size_t o = XI3(i, k, j);
  const float m = A1(FL(c1h), k) * (A2(FL(mu_1), i, j) + A2(FL(mub), i, j)) + A1(FL(c2h), k);
  q_old[o] = (m * q_old[o] + dt * q_tend[o]) / m;
  q_tend[o] = 0.f;
}

// rk_update_scalar (module_em.py:1437-1598): advective tendency (interior
// only, times msfty) plus the physics tendency, mass-weighted update.
__global__ void k_update_scalar(const DevCtx xv, float *q, float *q_old, const float *q_tend, int rk, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const size_t o = XI3(i, k, j);
  const int sz = x->spec_zone;
  const bool inner = (i >= 1 + sz && i <= x->ide - sz - 1 && j >= 1 + sz && j <= x->jde - sz - 1);
  float tend = inner ? FL(advect_tend)[o] * A2(FL(msfty), i, j) : 0.f;
  tend = tend + q_tend[o];
  const float mub = A2(FL(mub), i, j);
  const float muold = A2(FL(mu_1), i, j) + mub, munew = A2(FL(mu_2), i, j) + mub;
  const float c1 = A1(FL(c1h), k), c2 = A1(FL(c2h), k);
  const float oldm = c1 * muold + c2, newm = c1 * munew + c2;
  float old;
  if (rk == 1) {
    old = q[o];
    q_old[o] = old;
  } else {
    old = q_old[o];
  }
  q[o] = (oldm * old + dt * tend) / newm;
}
