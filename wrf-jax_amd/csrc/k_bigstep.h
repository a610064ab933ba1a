// RK-stage (big step) kernels: rk_step_prep, rhs_ph, horizontal pressure
// gradient, pg_buoy_w, Coriolis, curvature, rk_addtend_dry, calc_p_rho_phi.
// Semantics: src/wrf_jax/dyn_em/module_big_step_utilities_em.py,
// module_big_step_utilities_em_fast.py and module_em.py.
#pragma once
#include "wrfdc_common.h"
#include "k_f64math.h"

__global__ void k_fill(float *p, float v, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) p[t] = v;
}
__global__ void k_copy(float *dst, const float *src, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) dst[t] = src[t];
}

// mut, muu, muv (calc_mu_uv, module_big_step_utilities_em.py:174-266) and
// muus/muvs from muts (calc_mu_uv_1).  2-D over i [1, ide], j [1, jde].
__global__ void k_mu_uv(const DevCtx xv, const float *mu, const float *mub, float *muu, float *muv, float *mut) {
  const DevCtx *__restrict__ x = &xv;
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = 1 + (int)blockIdx.y;
  const int ide = x->ide, jde = x->jde;
  if (i > ide || j > jde) return;
  // mub == nullptr: calc_mu_uv_1 on the full column mass (adding the zero
  // base state of the reference form leaves every sum unchanged)
  if (mut) A2(mut, i, j) = A2(mub, i, j) + A2(mu, i, j);
  if (j <= jde - 1) {
    int ia = (i == ide) ? ide - 1 : i;
    int ib = (i == 1) ? 1 : ((i == ide) ? ide - 1 : i - 1);
    float s = A2(mu, ia, j) + A2(mu, ib, j);
    if (mub) s = s + A2(mub, ia, j) + A2(mub, ib, j);
    A2(muu, i, j) = 0.5f * s;
  }
  if (i <= ide - 1) {
    int ja = (j == jde) ? jde - 1 : j;
    int jb = (j == 1) ? 1 : ((j == jde) ? jde - 1 : j - 1);
    float s = A2(mu, i, ja) + A2(mu, i, jb);
    if (mub) s = s + A2(mub, i, ja) + A2(mub, i, jb);
    A2(muv, i, j) = 0.5f * s;
  }
}

// couple_momentum + calc_cq + calc_alt + calc_php over the full 3-D extent
// (zero outside each field's range, as the reference's fresh outputs).
__global__ void k_rk_prep3(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const unsigned o = XI3(i, k, j), o2 = XI2(i, j);
  const unsigned ok1 = k < kde ? o + (unsigned)x->ip : o;  // level k+1 (k at the top, not used there)
  const int kh = imin_(k, kde - 1);
  const float c1h = A1(FL(c1h), kh), c2h = A1(FL(c2h), kh), c1f = A1(FL(c1f), k), c2f = A1(FL(c2f), k);
  // every load of the point first (the moist species follow in groups)
  const float u = elem(FL(u_2), o), v = elem(FL(v_2), o), w = elem(FL(w_2), o);
  const float muu = elem(FL(muu), o2), muv = elem(FL(muv), o2), mut = elem(FL(mut), o2);
  const float msfuy = elem(FL(msfuy), o2), msfvxi = elem(FL(msfvx_inv), o2), msfty = elem(FL(msfty), o2);
  const float al = elem(FL(al), o), alb = elem(FL(alb), o);
  const float phb0 = elem(FL(phb), o), phb1 = elem(FL(phb), ok1), ph0 = elem(FL(ph_2), o), ph1 = elem(FL(ph_2), ok1);
  pin1(u);
  pin1(v);
  pin1(w);
  pin1(al);
  pin1(alb);
  pin1(phb0);
  pin1(phb1);
  pin1(ph0);
  pin1(ph1);
  // ru / rv / rw
  elem(FL(ru), o) = (j <= jde - 1 && k <= kde - 1) ? u * (c1h * muu + c2h) / msfuy : 0.f;
  elem(FL(rv), o) = (i <= ide - 1 && k <= kde - 1) ? v * (c1h * muv + c2h) * msfvxi : 0.f;
  elem(FL(rw), o) = (i <= ide - 1 && j <= jde - 1) ? w * (c1f * mut + c2f) / msfty : 0.f;
  // alt / php on mass points
  const bool mass = (i <= ide - 1 && j <= jde - 1 && k <= kde - 1);
  elem(FL(alt), o) = mass ? al + alb : 0.f;
  if (i == ide || j == jde) elem(FL(ww), o) = 0.f;
  elem(FL(php), o) = mass ? 0.5f * (phb0 + phb1 + ph0 + ph1) : 0.f;
  // cqu / cqv / cqw; neighbour i-1 / j-1 at the low edge wraps to the last
  // stored index like the reference's negative indexing
  float cu = 0.f, cv = 0.f, cw = 0.f;
  if (k <= kde - 1) {
    const int im = (i == 1) ? ide : i - 1;
    const int jm = (j == 1) ? jde : j - 1;
    const unsigned om = o, oi = XI3(im, k, j), oj = XI3(i, k, jm), ok = XI3(i, imax_(k - 1, 1), j);
    float su = 0.f, sv = 0.f, sw = 0.f;
    // species in groups of 4, every load of a group issued first (species
    // past the last load the last one again); sums in species order
    const int nm = x->nmoist;
    for (int s0 = 0; s0 < nm; s0 += 4) {
      float q0[4], qi[4], qj[4], qk[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const float *qp = SP(moist, imin_(s0 + q, nm - 1));
        q0[q] = elem(qp, om);
        qi[q] = elem(qp, oi);
        qj[q] = elem(qp, oj);
        qk[q] = elem(qp, ok);
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (s0 + q < nm) {
          su = su + q0[q] + qi[q];
          sv = sv + q0[q] + qj[q];
          if (k > 1) sw = sw + q0[q] + qk[q];
        }
      }
    }
    if (j < jde) cu = 1.f / (1.f + 0.5f * su);
    if (i < ide) cv = 1.f / (1.f + 0.5f * sv);
    if (i < ide && j < jde && k > 1) cw = 0.5f * sw;
  }
  elem(FL(cqu), o) = cu;
  elem(FL(cqv), o) = cv;
  elem(FL(cqw), o) = cw;
}

// calc_ww_cp (module_big_step_utilities_em.py:655-803): column kernel over
// mass columns.  Omega is the bottom-up integral of the coupled horizontal
// mass divergence minus its column share.
// One thread per mass column (block 64): the divergence of every level is
// formed once, 8 levels of loads at a time, and kept in LDS (dynamic,
// (kde - 1) x 64 floats); the column sum and the integration then run in the
// reference's level order, and the integration's stores never hold up a
// load.
__global__ void __launch_bounds__(64) k_ww_cp(const DevCtx xv, R2 r) {
  const DevCtx *__restrict__ x = &xv;
  constexpr int kMaxLev = 152;  // kde <= 151 (wrfdc_create checks nz <= 150)
  __shared__ float s_c1h[kMaxLev], s_c2h[kMaxLev], s_dnw[kMaxLev];
  extern __shared__ float s_div_[];  // (kde - 1) x 64: the divergence of every level
#define s_div(kk) s_div_[(kk) * 64 + lane]
  const int kde = x->kde, ktf = kde - 1;
  for (int q = (int)threadIdx.x; q < kde; q += 64) {
    s_c1h[q] = FL(c1h)[q];
    s_c2h[q] = FL(c2h)[q];
    s_dnw[q] = FL(dnw)[q];
  }
  __syncthreads();
  GET_IJ(r);
  const float *mup = FL(mu_2), *mub = FL(mub), *u = FL(u_2), *v = FL(v_2);
  float *ww = FL(ww);
  const int lane = (int)threadIdx.x;
  const int im = (i - 1 < 1) ? i : i - 1, jm = (j - 1 < 1) ? j : j - 1;
  const float muu_i = 0.5f * (A2(mup, i, j) + A2(mub, i, j) + A2(mup, im, j) + A2(mub, im, j));
  const float muu_ip = 0.5f * (A2(mup, i + 1, j) + A2(mub, i + 1, j) + A2(mup, i, j) + A2(mub, i, j));
  const float muv_j = 0.5f * (A2(mup, i, j) + A2(mub, i, j) + A2(mup, i, jm) + A2(mub, i, jm));
  const float muv_jp = 0.5f * (A2(mup, i, j + 1) + A2(mub, i, j + 1) + A2(mup, i, j) + A2(mub, i, j));
  const float fu1 = A2(FL(msfuy), i + 1, j), fu0 = A2(FL(msfuy), i, j), fv1 = A2(FL(msfvx_inv), i, j + 1), fv0 = A2(FL(msfvx_inv), i, j);
  const float mtx = A2(FL(msftx), i, j);
  const unsigned sk = (unsigned)x->ip, sj = (unsigned)(kde * x->ip), o1 = XI3(i, 1, j);
  float dmdt = 0.f;
  for (int k0 = 1; k0 <= ktf; k0 += 8) {
    float ue[8], uw[8], vn[8], vs[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const unsigned o = o1 + (unsigned)(imin_(k0 + q, ktf) - 1) * sk;
      ue[q] = elem(u, o + 1u);
      uw[q] = elem(u, o);
      vn[q] = elem(v, o + sj);
      vs[q] = elem(v, o);
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = k0 + q;
      if (k > ktf) break;
      // ww_divv with the loaded values (same arithmetic)
      const float c1 = s_c1h[k - 1], c2 = s_c2h[k - 1];
      const float xf = (c1 * muu_ip + c2) * ue[q] / fu1 - (c1 * muu_i + c2) * uw[q] / fu0;
      const float yf = (c1 * muv_jp + c2) * vn[q] * fv1 - (c1 * muv_j + c2) * vs[q] * fv0;
      const float dv = mtx * s_dnw[k - 1] * (x->rdx * xf + x->rdy * yf);
      s_div(k - 1) = dv;
      dmdt += dv;
    }
  }
  A3(ww, i, 1, j) = 0.f;
  float prev = 0.f;
  for (int k = 2; k <= ktf; k++) {
    const float cur = prev - s_dnw[k - 2] * s_c1h[k - 2] * dmdt - s_div(k - 2);
    elem(ww, o1 + (unsigned)(k - 1) * sk) = cur;
    prev = cur;
  }
  A3(ww, i, x->kde, j) = 0.f;
#undef s_div
}

// rhs_ph (module_big_step_utilities_em.py:2386-2794): vertical advection of
// phi, the w g term and horizontal phi advection (6th/4th/2nd order with the
// reference's specified-boundary column pattern).  k in [2, kde].
// the phi tendency at (i,k,j), k in [2, kde], starting from `out`
__device__ __forceinline__ float rhs_ph_val(const DevCtx *__restrict__ x, int i, int k, int j, float out) {
  const int ide = x->ide, jde = x->jde, kde = x->kde, kte = kde;
  const float *c1f = FL(c1f), *c2f = FL(c2f), *fnm = FL(fnm), *fnp = FL(fnp), *rdnw = FL(rdnw);
  const float rdx = x->rdx, rdy = x->rdy, g = x->g, cfn = x->cfn, cfn1 = x->cfn1;
  const unsigned nb = (unsigned)x->n3, o = XI3(i, k, j), o2 = XI2(i, j), sk = (unsigned)x->ip, sj = (unsigned)(kde * x->ip);
  // every load of the point first (buffer loads: the levels k-2 .. k+1 and
  // the stencil windows may fall outside the field where they are not used)
  const BufF bu = bufv(FL(u_2), nb), bv = bufv(FL(v_2), nb), bww = bufv(FL(ww), nb), bph = bufv(FL(ph_2), nb), bpb = bufv(FL(phb), nb);
  const float wwm = bld(bww, o - sk), ww0 = bld(bww, o), wwp = bld(bww, o + sk);
  const float phm = bld(bph, o - sk), php1 = bld(bph, o + sk), pbm = bld(bpb, o - sk), pbp1 = bld(bpb, o + sk);
  float phy[7], pby[7], phx[7], pbx[7];  // ph, phb at j-3 .. j+3 and i-3 .. i+3 (slot 3: the point)
#pragma unroll
  for (int n = 0; n < 7; n++) {
    phy[n] = bld(bph, o + (unsigned)(n - 3) * sj);
    pby[n] = bld(bpb, o + (unsigned)(n - 3) * sj);
    phx[n] = bld(bph, o + (unsigned)(n - 3));
    pbx[n] = bld(bpb, o + (unsigned)(n - 3));
  }
  float vn[3], vs[3], ue[3], uw[3];  // v at rows j+1, j and u at columns i+1, i; levels k, k-1, k-2
#pragma unroll
  for (int n = 0; n < 3; n++) {
    vn[n] = bld(bv, o + sj - (unsigned)n * sk);
    vs[n] = bld(bv, o - (unsigned)n * sk);
    ue[n] = bld(bu, o + 1u - (unsigned)n * sk);
    uw[n] = bld(bu, o - (unsigned)n * sk);
  }
  const float wk = elem(FL(w_2), o), mut = elem(FL(mut), o2), msfty = elem(FL(msfty), o2);
  // column coefficients (the top level's rdnw index clamped: not used there)
  const float rdnw0 = A1(rdnw, k - 1), rdnw1 = A1(rdnw, imin_(k, kde - 1)), fnm0 = A1(fnm, k), fnp0 = A1(fnp, k);
  const float c1fk = A1(c1f, k), c2fk = A1(c2f, k);
  const float muv1 = A2(FL(muv), i, j + 1), muv0 = elem(FL(muv), o2), msfvy1 = A2(FL(msfvy), i, j + 1), msfvy0 = elem(FL(msfvy), o2);
  const float muu1 = A2(FL(muu), i + 1, j), muu0 = elem(FL(muu), o2), msfux1 = A2(FL(msfux), i + 1, j), msfux0 = elem(FL(msfux), o2);
  pin1(wk);
  pin1(mut);
  pin1(msfty);
  pin1(muv1);
  pin1(muv0);
  pin1(msfvy1);
  pin1(msfvy0);
  pin1(muu1);
  pin1(muu0);
  pin1(msfux1);
  pin1(msfux0);
  pin1(wwm);
  pin1(ww0);
  pin1(wwp);
  pin1(phm);
  pin1(php1);
  pin1(pbm);
  pin1(pbp1);
#pragma unroll
  for (int n = 0; n < 7; n++) {
    pin1(phy[n]);
    pin1(pby[n]);
    pin1(phx[n]);
    pin1(pbx[n]);
  }
#pragma unroll
  for (int n = 0; n < 3; n++) {
    pin1(vn[n]);
    pin1(vs[n]);
    pin1(ue[n]);
    pin1(uw[n]);
  }
  const float ph0 = phy[3], pb0 = pby[3];
  if (k <= kte - 1) {
    float wdk = 0.5f * (ww0 + wwm) * rdnw0 * (ph0 - phm + pb0 - pbm);
    float wdk1 = 0.5f * (wwp + ww0) * rdnw1 * (php1 - ph0 + pbp1 - pb0);
    out += -(fnm0 * wdk1 + fnp0 * wdk);
  } else {
    out = 0.f;
  }
  out += (c1fk * mut + c2fk) * g * wk / msfty;
  // y direction
  {
    int mode = 0;  // 0 none, 2 second, 4 fourth, 6 sixth
    if (j >= 4 && j <= jde - 4) mode = 6;
    else if (j == 3 || j == jde - 3) mode = 4;
    else if (j == 2 || j == jde - 2) mode = 2;
    if (mode) {
      float va1, va0, fac;
      if (k < kte) {
        va1 = (c1fk * muv1 + c2fk) * (vn[0] + vn[1]) * msfvy1;
        va0 = (c1fk * muv0 + c2fk) * (vs[0] + vs[1]) * msfvy0;
        fac = 0.25f;
      } else {
        va1 = (c1fk * muv1 + c2fk) * (cfn * vn[1] + cfn1 * vn[2]) * msfvy1;
        va0 = (c1fk * muv0 + c2fk) * (cfn * vs[1] + cfn1 * vs[2]) * msfvy0;
        fac = 0.5f;
      }
      if (mode == 2) {
        float term = va1 * (pby[4] - pb0 + phy[4] - ph0) + va0 * (pb0 - pby[2] + ph0 - phy[2]);
        out += -(fac * rdy / msfty) * term;
      } else {
        float ca = mode == 6 ? 45.f : 8.f, cb = mode == 6 ? -9.f : -1.f, cc = mode == 6 ? 1.f : 0.f;
        float scale = mode == 6 ? (1.f / 60.f) : (1.f / 12.f);
        const int mo = mode == 6 ? 3 : 2;
        float grad = ca * (phy[4] - phy[2]) + cb * (phy[5] - phy[1]) + cc * (phy[3 + mo] - phy[3 - mo]) + ca * (pby[4] - pby[2]) +
                     cb * (pby[5] - pby[1]) + cc * (pby[3 + mo] - pby[3 - mo]);
        out += -(fac * rdy / msfty) * (va1 + va0) * scale * grad;
      }
    }
  }
  // x direction: columns ids+2 / ide-3 get no advection for specified
  // boundaries (the reference applies their 4th-order term only for open ones)
  {
    int mode = 0;
    if (i >= 4 && i <= ide - 4) mode = 6;
    else if (i == 2 || i == ide - 2) mode = 2;
    if (mode) {
      float ua1, ua0, fac;
      if (k < kte) {
        ua1 = (c1fk * muu1 + c2fk) * (ue[0] + ue[1]) * msfux1;
        ua0 = (c1fk * muu0 + c2fk) * (uw[0] + uw[1]) * msfux0;
        fac = 0.25f;
      } else {
        ua1 = (c1fk * muu1 + c2fk) * (cfn * ue[1] + cfn1 * ue[2]) * msfux1;
        ua0 = (c1fk * muu0 + c2fk) * (cfn * uw[1] + cfn1 * uw[2]) * msfux0;
        fac = 0.5f;
      }
      if (mode == 2) {
        float term = ua1 * (pbx[4] - pb0 + phx[4] - ph0) + ua0 * (pb0 - pbx[2] + ph0 - phx[2]);
        out += -(fac * rdx / msfty) * term;
      } else {
        float grad = 45.f * (phx[4] - phx[2]) - 9.f * (phx[5] - phx[1]) + 1.f * (phx[6] - phx[0]) + 45.f * (pbx[4] - pbx[2]) -
                     9.f * (pbx[5] - pbx[1]) + 1.f * (pbx[6] - pbx[0]);
        out += -(fac * rdx / msfty) * (ua1 + ua0) * (1.f / 60.f) * grad;
      }
    }
  }
  return out;
}
__global__ void k_rhs_ph(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  A3(FL(ph_tend), i, k, j) = rhs_ph_val(x, i, k, j, A3(FL(ph_tend), i, k, j));
}

// horizontal_pressure_gradient (module_big_step_utilities_em.py:1452-1671).
// axis 0: u points i [2, ide-1], j [1, jde-1]; axis 1: v points i [1, ide-1], j [2, jde-1].
// the amount the pressure gradient subtracts from the u (axis 0) or v
// (axis 1) tendency at (i,k,j): cq * dp
__device__ __forceinline__ float hpg_term(const DevCtx *__restrict__ x, int axis, int i, int k, int j) {
  const float *ph = FL(ph_2), *alt = FL(alt), *p = FL(p), *pb = FL(pb), *al = FL(al), *php = FL(php), *mu = FL(mu_2);
  const float *c1h = FL(c1h), *c2h = FL(c2h), *rdnw = FL(rdnw);
  const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
  const float rd = axis == 0 ? x->rdx : x->rdy;
  const float rr = axis == 0 ? A2(FL(msfux), i, j) / A2(FL(msfuy), i, j) : A2(FL(msfvy), i, j) / A2(FL(msfvx), i, j);
  const float m = axis == 0 ? A2(FL(muu), i, j) : A2(FL(muv), i, j);
  float dp = rr * 0.5f * rd * (A1(c1h, k) * m + A1(c2h, k)) *
             ((A3(ph, i, k + 1, j) - A3(ph, i - di, k + 1, j - dj) + A3(ph, i, k, j) - A3(ph, i - di, k, j - dj)) +
              (A3(alt, i, k, j) + A3(alt, i - di, k, j - dj)) * (A3(p, i, k, j) - A3(p, i - di, k, j - dj)) +
              (A3(al, i, k, j) + A3(al, i - di, k, j - dj)) * (A3(pb, i, k, j) - A3(pb, i - di, k, j - dj)));
  // level-interpolated pressure sums of both columns at w levels k+1 and k
  // (none at and above kde; level 1 extrapolated from levels 1-3), from p at
  // levels k, k+1 and k-1 (level 3 at k = 1), every form evaluated and the
  // level's selected
  const int kde = x->kde, kp = k + 1;
  const BufF bp = bufv(p, (unsigned)x->n3);
  const unsigned sk = (unsigned)x->ip, oa = XI3(i - di, k, j - dj), ob = XI3(i, k, j);
  const unsigned ol = k == 1 ? 2u * sk : 0u - sk;
  const float a0 = bld(bp, oa), a1 = bld(bp, oa + sk), a2 = bld(bp, oa + ol);
  const float b0 = bld(bp, ob), b1 = bld(bp, ob + sk), b2 = bld(bp, ob + ol);
  const float dpn1 = kp >= kde ? 0.f : 0.5f * (A1(FL(fnm), kp) * (a1 + b1) + A1(FL(fnp), kp) * (a0 + b0));
  const float dpn0 = k == 1 ? 0.5f * (x->cf1 * (a0 + b0) + x->cf2 * (a1 + b1) + x->cf3 * (a2 + b2))
                            : (k >= kde ? 0.f : 0.5f * (A1(FL(fnm), k) * (a0 + b0) + A1(FL(fnp), k) * (a2 + b2)));
  dp = dp + rr * rd * (A3(php, i, k, j) - A3(php, i - di, k, j - dj)) *
                (A1(rdnw, k) * (dpn1 - dpn0) - 0.5f * (A1(c1h, k) * A2(mu, i - di, j - dj) + A1(c1h, k) * A2(mu, i, j)));
  return (axis == 0 ? A3(FL(cqu), i, k, j) : A3(FL(cqv), i, k, j)) * dp;
}
__global__ void k_hpg(const DevCtx xv, int axis, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  if (axis == 0) A3(FL(ru_tend), i, k, j) = A3(FL(ru_tend), i, k, j) - hpg_term(x, 0, i, k, j);
  else A3(FL(rv_tend), i, k, j) = A3(FL(rv_tend), i, k, j) - hpg_term(x, 1, i, k, j);
}

// pg_buoy_w (module_big_step_utilities_em.py:1674-1800) at one w level
// k in [2, kde], added to rwt; reads cqw as it is before the column kernel
// below replaces it by 1 / (1 + cqw)
// (cq1_out: 1 / (1 + cqw) at level min(k, kde-1), the value pg_buoy_w then
// stores in cqw)
__device__ __forceinline__ float buoy_w_term(const DevCtx *__restrict__ x, int i, int k, int j, float *cq1_out = nullptr) {
  const float *p = FL(p), *mu = FL(mu_2), *mub = FL(mub), *c1f = FL(c1f), *c2f = FL(c2f);
  const int kde = x->kde, kc = imin_(k, kde - 1);  // cqw and p level: k, kde-1 at the top
  const float gm = (1.f / A2(FL(msfty), i, j)) * x->g;
  const float cq = A3(FL(cqw), i, kc, j);
  const float cq1 = 1.f / (1.f + cq), cq2 = cq * cq1;
  if (cq1_out) *cq1_out = cq1;
  const float pk = A3(p, i, kc, j), pkm = A3(p, i, k - 1, j);
  const float top = gm * (cq1 * 2.f * A1(FL(rdnw), kde - 1) * (-pk) - (A1(c1f, kde) * A2(mu, i, j)) -
                          cq2 * (A1(c1f, kde) * A2(mub, i, j) + A1(c2f, kde)));
  const float mid = gm * (cq1 * A1(FL(rdn), kc) * (pk - pkm) - (A1(c1f, k) * A2(mu, i, j)) - cq2 * (A1(c1f, k) * A2(mub, i, j) + A1(c2f, k)));
  return k == kde ? top : mid;
}
// cqw := 1 / (1 + cqw) (the second half of pg_buoy_w): levels [2, kde-1] in
// the stage API; in the fused step only level kde-1 (the w kernel's top
// level reads it, so the fused w kernel replaces levels [2, kde-2] itself)
__global__ void k_cqw_inv(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  float *cqw = FL(cqw);
  A3(cqw, i, k, j) = 1.f / (1.f + A3(cqw, i, k, j));
}
// pg_buoy_w (module_big_step_utilities_em.py:1674-1800): column kernel
__global__ void k_pg_buoy_w(const DevCtx xv, R2 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IJ(r);
  float *rwt = FL(rw_tend), *cqw = FL(cqw);
  const float *p = FL(p), *mu = FL(mu_2), *mub = FL(mub), *c1f = FL(c1f), *c2f = FL(c2f), *rdnw = FL(rdnw), *rdn = FL(rdn);
  const int kde = x->kde;
  const float gm = (1.f / A2(FL(msfty), i, j)) * x->g;
  float cq = A3(cqw, i, kde - 1, j);
  float cq1 = 1.f / (1.f + cq), cq2 = cq * cq1;
  A3(rwt, i, kde, j) += gm * (cq1 * 2.f * A1(rdnw, kde - 1) * (-A3(p, i, kde - 1, j)) - (A1(c1f, kde) * A2(mu, i, j)) -
                              cq2 * (A1(c1f, kde) * A2(mub, i, j) + A1(c2f, kde)));
  for (int k = 2; k <= kde - 1; k++) {
    cq = A3(cqw, i, k, j);
    cq1 = 1.f / (1.f + cq);
    cq2 = cq * cq1;
    A3(rwt, i, k, j) += gm * (cq1 * A1(rdn, k) * (A3(p, i, k, j) - A3(p, i, k - 1, j)) - (A1(c1f, k) * A2(mu, i, j)) -
                              cq2 * (A1(c1f, k) * A2(mub, i, j) + A1(c2f, k)));
    A3(cqw, i, k, j) = cq1;
  }
}

// coriolis + curvature (module_big_step_utilities_em.py:978-1120, 1214-1449)
// for specified boundaries.  comp 0: u points; 1: v points; 2: w points.
__device__ __forceinline__ float vxgm(const DevCtx *__restrict__ x, int i, int k, int j) {
  const float *u = FL(u_2), *v = FL(v_2);
  return 0.5f * (A3(u, i, k, j) + A3(u, i + 1, k, j)) * (A2(FL(msfvx), i, j + 1) - A2(FL(msfvx), i, j)) * x->rdy -
         0.5f * (A3(v, i, k, j) + A3(v, i, k, j + 1)) * (A2(FL(msfuy), i + 1, j) - A2(FL(msfuy), i, j)) * x->rdx;
}
// The terms of each component as values of their own (a kernel can issue
// their loads with its others) and their accumulation in the reference's
// order.
__device__ __forceinline__ float4 cor_u_terms(const DevCtx *__restrict__ x, int i, int k, int j) {
  const float *rv = FL(rv), *rw = FL(rw), *u = FL(u_2);
  const float *f = FL(f), *e = FL(e), *cosa = FL(cosa);
  const float rer = x->reradius;
  float4 c;
  c.x = (A2(FL(msfux), i, j) / A2(FL(msfuy), i, j)) * 0.5f * (A2(f, i, j) + A2(f, i - 1, j)) * 0.25f *
        (A3(rv, i - 1, k, j + 1) + A3(rv, i, k, j + 1) + A3(rv, i - 1, k, j) + A3(rv, i, k, j));
  c.y = 0.5f * (A2(e, i, j) + A2(e, i - 1, j)) * 0.5f * (A2(cosa, i, j) + A2(cosa, i - 1, j)) * 0.25f *
        (A3(rw, i - 1, k + 1, j) + A3(rw, i - 1, k, j) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
  c.z = 0.5f * (vxgm(x, i, k, j) + vxgm(x, i - 1, k, j)) * 0.25f * (A3(rv, i - 1, k, j + 1) + A3(rv, i, k, j + 1) + A3(rv, i - 1, k, j) + A3(rv, i, k, j));
  c.w = A3(u, i, k, j) * rer * 0.25f * (A3(rw, i - 1, k + 1, j) + A3(rw, i - 1, k, j) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
  return c;
}
__device__ __forceinline__ float cor_u_add(float val, float4 c) {
  val = val + c.x;
  val = val - c.y;
  return val + c.z - c.w;
}
__device__ __forceinline__ float cor_curv_u(const DevCtx *__restrict__ x, int i, int k, int j, float val) { return cor_u_add(val, cor_u_terms(x, i, k, j)); }
__device__ __forceinline__ float4 cor_v_terms(const DevCtx *__restrict__ x, int i, int k, int j) {
  const float *ru = FL(ru), *rw = FL(rw), *v = FL(v_2);
  const float *f = FL(f), *e = FL(e), *sina = FL(sina);
  const float rer = x->reradius;
  const float r = A2(FL(msfvy), i, j) / A2(FL(msfvx), i, j);
  float4 c;
  c.x = r * 0.5f * (A2(f, i, j) + A2(f, i, j - 1)) * 0.25f * (A3(ru, i, k, j) + A3(ru, i + 1, k, j) + A3(ru, i, k, j - 1) + A3(ru, i + 1, k, j - 1));
  c.y = r * 0.5f * (A2(e, i, j) + A2(e, i, j - 1)) * 0.5f * (A2(sina, i, j) + A2(sina, i, j - 1)) * 0.25f *
        (A3(rw, i, k + 1, j - 1) + A3(rw, i, k, j - 1) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
  c.z = 0.5f * (vxgm(x, i, k, j) + vxgm(x, i, k, j - 1)) * 0.25f * (A3(ru, i, k, j) + A3(ru, i + 1, k, j) + A3(ru, i, k, j - 1) + A3(ru, i + 1, k, j - 1));
  c.w = r * A3(v, i, k, j) * rer * 0.25f * (A3(rw, i, k + 1, j - 1) + A3(rw, i, k, j - 1) + A3(rw, i, k + 1, j) + A3(rw, i, k, j));
  return c;
}
__device__ __forceinline__ float cor_v_add(float val, float4 c) {
  val = val - c.x;
  val = val + c.y;
  return val - c.z - c.w;
}
__device__ __forceinline__ float cor_curv_v(const DevCtx *__restrict__ x, int i, int k, int j, float val) { return cor_v_add(val, cor_v_terms(x, i, k, j)); }
__device__ __forceinline__ float2 cor_w_terms(const DevCtx *__restrict__ x, int i, int k, int j) {
  const float *ru = FL(ru), *rv = FL(rv), *u = FL(u_2), *v = FL(v_2);
  const float *e = FL(e), *sina = FL(sina), *cosa = FL(cosa), *fzm = FL(fnm), *fzp = FL(fnp);
  const float rer = x->reradius;
  const float mr = A2(FL(msftx), i, j) / A2(FL(msfty), i, j);
  float2 c;
  c.x = A2(e, i, j) * (A2(cosa, i, j) * 0.5f * (A1(fzm, k) * (A3(ru, i, k, j) + A3(ru, i + 1, k, j)) + A1(fzp, k) * (A3(ru, i, k - 1, j) + A3(ru, i + 1, k - 1, j))) -
                       mr * A2(sina, i, j) * 0.5f * (A1(fzm, k) * (A3(rv, i, k, j) + A3(rv, i, k, j + 1)) + A1(fzp, k) * (A3(rv, i, k - 1, j) + A3(rv, i, k - 1, j + 1))));
  c.y = rer * (0.5f * (A1(fzm, k) * (A3(ru, i, k, j) + A3(ru, i + 1, k, j)) + A1(fzp, k) * (A3(ru, i, k - 1, j) + A3(ru, i + 1, k - 1, j))) *
                   0.5f * (A1(fzm, k) * (A3(u, i, k, j) + A3(u, i + 1, k, j)) + A1(fzp, k) * (A3(u, i, k - 1, j) + A3(u, i + 1, k - 1, j))) +
               mr * 0.5f * (A1(fzm, k) * (A3(rv, i, k, j) + A3(rv, i, k, j + 1)) + A1(fzp, k) * (A3(rv, i, k - 1, j) + A3(rv, i, k - 1, j + 1))) *
                   0.5f * (A1(fzm, k) * (A3(v, i, k, j) + A3(v, i, k, j + 1)) + A1(fzp, k) * (A3(v, i, k - 1, j) + A3(v, i, k - 1, j + 1))));
  return c;
}
__device__ __forceinline__ float cor_w_add(float val, float2 c) {
  val = val + c.x;
  return val + c.y;
}
__device__ __forceinline__ float cor_curv_w(const DevCtx *__restrict__ x, int i, int k, int j, float val) { return cor_w_add(val, cor_w_terms(x, i, k, j)); }
__global__ void k_cor_curv(const DevCtx xv, int comp, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  if (comp == 0) A3(FL(ru_tend), i, k, j) = cor_curv_u(x, i, k, j, A3(FL(ru_tend), i, k, j));
  else if (comp == 1) A3(FL(rv_tend), i, k, j) = cor_curv_v(x, i, k, j, A3(FL(rv_tend), i, k, j));
  else A3(FL(rw_tend), i, k, j) = cor_curv_w(x, i, k, j, A3(FL(rw_tend), i, k, j));
}

// rk_addtend_dry (module_em.py:959-1148), one component at one point: the
// physics / boundary tendency (scaled by the map factor) added to the RK
// tendency t, which is returned; at the first stage the saved tendency is
// folded into the physics tendency first (written back).
// In two parts, so a kernel can issue the loads with its others:
// tendf_load reads the physics tendency and, at the first stage, the saved
// one; tendf_fold adds the saved tendency times its weight at the first
// stage (written back; only the thread owning the point may call it) and
// returns the physics tendency.
__device__ __forceinline__ float2 tendf_load(const float *tf, const float *sv, unsigned o, int rk) {
  return make_float2(elem(tf, o), rk == 1 ? elem(sv, o) : 0.f);
}
__device__ __forceinline__ float tendf_fold(float *tf, unsigned o, float2 ld, float w, int rk) {
  float v = ld.x;
  if (rk == 1) {
    v = ld.x + ld.y * w;
    elem(tf, o) = v;
  }
  return v;
}
__device__ __forceinline__ float addtend_u(const DevCtx *__restrict__ x, int i, int k, int j, int rk, float t) {
  const unsigned o = XI3(i, k, j);
  const float m = A2(FL(msfuy), i, j);
  return t + tendf_fold(FL(ru_tendf), o, tendf_load(FL(ru_tendf), FL(u_save), o, rk), m, rk) / m;
}
__device__ __forceinline__ float addtend_v(const DevCtx *__restrict__ x, int i, int k, int j, int rk, float t) {
  const unsigned o = XI3(i, k, j);
  return t + tendf_fold(FL(rv_tendf), o, tendf_load(FL(rv_tendf), FL(v_save), o, rk), A2(FL(msfvx), i, j), rk) * A2(FL(msfvx_inv), i, j);
}
__device__ __forceinline__ float addtend_w(const DevCtx *__restrict__ x, int i, int k, int j, int rk, float t) {
  const unsigned o = XI3(i, k, j);
  const float m = A2(FL(msfty), i, j);
  return t + tendf_fold(FL(rw_tendf), o, tendf_load(FL(rw_tendf), FL(w_save), o, rk), m, rk) / m;
}
// (the saved phi and theta tendencies carry no weight: x * 1 is x)
__device__ __forceinline__ float addtend_ph(const DevCtx *__restrict__ x, int i, int k, int j, int rk, float t) {
  const unsigned o = XI3(i, k, j);
  const float m = A2(FL(msfty), i, j);
  return t + tendf_fold(FL(ph_tendf), o, tendf_load(FL(ph_tendf), FL(ph_save), o, rk), 1.f, rk) / m;
}
// theta: also the diabatic heating, and the column's mu tendency at k = 1
// mu_zero: mu_tend is known to be 0 here (a later stage of the fused step,
// which does not clear it) and is formed as 0 + mu_tendf
__device__ __forceinline__ float addtend_t(const DevCtx *__restrict__ x, int i, int k, int j, int rk, float t, bool mu_zero = false) {
  const unsigned o = XI3(i, k, j);
  const float m = A2(FL(msfty), i, j);
  const float mass = A1(FL(c1h), k) * A2(FL(mut), i, j) + A1(FL(c2h), k);
  const float tf = tendf_fold(FL(t_tendf), o, tendf_load(FL(t_tendf), FL(t_save), o, rk), 1.f, rk);
  const float heating = mass * elem(FL(h_diabatic), o) / m;
  if (k == 1) A2(FL(mu_tend), i, j) = (mu_zero ? 0.f : A2(FL(mu_tend), i, j)) + A2(FL(mu_tendf), i, j);
  return t + tf / m + heating;
}
// comp 0 u, 1 v, 2 w/ph, 3 theta
__global__ void k_addtend(const DevCtx xv, int comp, int rk, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const unsigned o = XI3(i, k, j);
  if (comp == 0) {
    elem(FL(ru_tend), o) = addtend_u(x, i, k, j, rk, elem(FL(ru_tend), o));
  } else if (comp == 1) {
    elem(FL(rv_tend), o) = addtend_v(x, i, k, j, rk, elem(FL(rv_tend), o));
  } else if (comp == 2) {
    elem(FL(rw_tend), o) = addtend_w(x, i, k, j, rk, elem(FL(rw_tend), o));
    elem(FL(ph_tend), o) = addtend_ph(x, i, k, j, rk, elem(FL(ph_tend), o));
  } else {
    elem(FL(t_tend), o) = addtend_t(x, i, k, j, rk, elem(FL(t_tend), o));
  }
}

// _calc_p_rho_phi_nonhydro_opt2 (runtime.py:12369-12537): hypsometric al, EOS p
__global__ void k_calc_p_rho_phi(const DevCtx xv, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const float *muts = FL(muts), *ph = FL(ph_2), *phb = FL(phb), *alb = FL(alb), *pb = FL(pb), *t = FL(t_2);
  const float m = A2(muts, i, j), ptop = x->ptop;
  const float pfu = A1(FL(c3f), k + 1) * m + A1(FL(c4f), k + 1) + ptop;
  const float pfd = A1(FL(c3f), k) * m + A1(FL(c4f), k) + ptop;
  const float phm = A1(FL(c3h), k) * m + A1(FL(c4h), k) + ptop;
  // every load first (the stores below could alias them as far as the
  // compiler knows)
  const unsigned o = XI3(i, k, j), o1 = o + (unsigned)x->ip;
  const float ph1 = elem(ph, o1), ph0 = elem(ph, o), phb1 = elem(phb, o1), phb0 = elem(phb, o), albv = elem(alb, o);
  const float qv = elem(SP(moist, 0), o), tv = elem(t, o), pbv = elem(pb, o);
  // log and pow rounded once to FP32 from an FP64 evaluation (f64_log,
  // f64_pow: the same FP32 values as the C library's FP64 log and pow for
  // every FP32 input in [1, 2), and in [1/16, 2) with exponent 1.4f).
  // p and al are small differences from the base state (p =
  // p0*temp^cpovcv - pb), which magnify an ulp of the total about a
  // thousandfold: the device library's FP32 logf / powf (within 1-2 ulp) put
  // p and al 4.5e-5 and 2.8e-5 (relative L2) away from the correctly rounded
  // result after one call.
  const float lg = (float)f64_log((double)(pfd / pfu));
  const float alv = (ph1 - ph0 + phb1 - phb0) / phm / lg - albv;
  elem(FL(al), o) = alv;
  const float qvf = 1.f + x->rvovrd * qv;
  const float temp = (x->r_d * (x->t0 + tv) * qvf) / (x->p0 * (alv + albv));
  elem(FL(p), o) = (float)f64_pow((double)temp, (double)x->cpovcv) * x->p0 - pbv;
}
