// Fused acoustic substep (one iteration of the small-step loop,
// src/wrf_jax/core/direct_step.py:3224-3390 / solve_em.py:560-640):
//
//   k_uv_substep    advance_uv for u and v in one pass over the (u|v) grid plus
//                   spec_bdyupdate of u and v on the specified zone;
//   k_acoustic_col  per column: advance_mu_t, spec_bdyupdate of t/mu/muts,
//                   advance_w (implicit w-phi solve), sumflux,
//                   spec_bdyupdate_ph, calc_p_rho and the physical-boundary
//                   copies of ph/al/p/mu_2/muts/mudf (the owner of column
//                   ide-1 / row jde-1 also writes its copy);
//   k_zero_grad_w   zero_grad_bdy of w, rows and columns in one launch.
//
// The column kernel runs 64 columns (one j row, lanes = i) per block of four
// waves.  Work that is pointwise in k is spread over the waves with loads
// coalesced along i; the two k recurrences (column sum + omega, tridiagonal
// sweeps) run in wave 0 over LDS-resident columns, in the same order as the
// reference's loops, so every rounded operation matches it.
#pragma once
#include "wrfdc_common.h"
#include "k_small.h"

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_uv_substep(const DevCtx *__restrict__ x, float dts, R3 r) {
  GET_IKJ(r);
  const int ide = x->ide, jde = x->jde;
  const size_t o = XI3(i, k, j);
  if (j <= jde - 1) {  // u point
    if (i >= 2 && i <= ide - 1 && j >= 2 && j <= jde - 2) {
      const float *p = FL(p), *pb = FL(pb), *ph = FL(ph_2), *alt = FL(alt), *al = FL(al), *php = FL(php);
      const float *c1h = FL(c1h), *c2h = FL(c2h), *mu = FL(mu_2), *mudf = FL(mudf);
      const float spacing = x->rdx, grid_length = 1.f / spacing;
      const float adv = FL(u_2)[o] + dts * FL(ru_tend)[o];
      float damping = -1.f * x->emdiv * grid_length * (A2(mudf, i, j) - A2(mudf, i - 1, j));
      damping = damping / A2(FL(msfuy), i, j);
      const float phi = (A3(ph, i, k + 1, j) - A3(ph, i - 1, k + 1, j)) + (A3(ph, i, k, j) - A3(ph, i - 1, k, j));
      const float pres = (A3(alt, i, k, j) + A3(alt, i - 1, k, j)) * (A3(p, i, k, j) - A3(p, i - 1, k, j));
      const float base = (A3(al, i, k, j) + A3(al, i - 1, k, j)) * (A3(pb, i, k, j) - A3(pb, i - 1, k, j));
      const float term = (phi + pres) + base;
      const float ratio = A2(FL(msfux), i, j) / A2(FL(msfuy), i, j);
      const float m = A1(c1h, k) * A2(FL(muu), i, j) + A1(c2h, k);
      float grad = (ratio * 0.5f * spacing * m) * term;
      const float vert = A1(FL(rdnw), k) * (uv_iface(x, p, i, j, 1, 0, k + 1) - uv_iface(x, p, i, j, 1, 0, k));
      const float muv_ = 0.5f * (A1(c1h, k) * A2(mu, i - 1, j) + A1(c1h, k) * A2(mu, i, j));
      grad = grad + (ratio * spacing * (A3(php, i, k, j) - A3(php, i - 1, k, j))) * (vert - muv_);
      FL(u_2)[o] = (adv - dts * FL(cqu)[o] * grad) + A1(c1h, k) * damping;
    } else {
      FL(u_2)[o] = FL(u_2)[o] + dts * FL(ru_tend)[o];
    }
  }
  if (i <= ide - 1) {  // v point
    if (i >= 2 && i <= ide - 2 && j >= 2 && j <= jde - 1) {
      const float *p = FL(p), *pb = FL(pb), *ph = FL(ph_2), *alt = FL(alt), *al = FL(al), *php = FL(php);
      const float *c1h = FL(c1h), *c2h = FL(c2h), *mu = FL(mu_2), *mudf = FL(mudf);
      const float spacing = x->rdy, grid_length = 1.f / spacing;
      const float adv = FL(v_2)[o] + dts * FL(rv_tend)[o];
      float damping = -1.f * x->emdiv * grid_length * (A2(mudf, i, j) - A2(mudf, i, j - 1));
      damping = damping * A2(FL(msfvx_inv), i, j);
      const float phi = (A3(ph, i, k + 1, j) - A3(ph, i, k + 1, j - 1)) + (A3(ph, i, k, j) - A3(ph, i, k, j - 1));
      const float pres = (A3(alt, i, k, j) + A3(alt, i, k, j - 1)) * (A3(p, i, k, j) - A3(p, i, k, j - 1));
      const float base = (A3(al, i, k, j) + A3(al, i, k, j - 1)) * (A3(pb, i, k, j) - A3(pb, i, k, j - 1));
      const float term = (phi + pres) + base;
      const float ratio = A2(FL(msfvy), i, j) / A2(FL(msfvx), i, j);
      const float m = A1(c1h, k) * A2(FL(muv), i, j) + A1(c2h, k);
      float grad = (ratio * 0.5f * spacing * m) * term;
      const float vert = A1(FL(rdnw), k) * (uv_iface(x, p, i, j, 0, 1, k + 1) - uv_iface(x, p, i, j, 0, 1, k));
      const float muv_ = 0.5f * (A1(c1h, k) * A2(mu, i, j - 1) + A1(c1h, k) * A2(mu, i, j));
      grad = grad + (ratio * spacing * (A3(php, i, k, j) - A3(php, i, k, j - 1))) * (vert - muv_);
      FL(v_2)[o] = (adv - dts * FL(cqv)[o] * grad) + A1(c1h, k) * damping;
    } else {
      FL(v_2)[o] = FL(v_2)[o] + dts * FL(rv_tend)[o];
    }
  }
}

// zero_grad_bdy of w: every destination (the zone ring of the mass grid)
// copies an interior point, so rows and columns go in one race-free pass.
__global__ void k_zero_grad_w(const DevCtx *__restrict__ x) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = 1 + (int)blockIdx.y;
  const int ibe = x->ide - 1, jbe = x->jde - 1;
  int i, j, e;
  if (k > x->kde || !zone_point(t, ibe, jbe, i, j, e)) return;
  float *w = FL(w_2);
  if (e < 2) {
    const int ii = imin_(imax_(i, 2), ibe - 1);
    A3(w, i, k, j) = A3(w, ii, k, e == 0 ? 2 : jbe - 1);
  } else {
    const int jj = imin_(imax_(j, 2), jbe - 1);
    A3(w, e == 2 ? 1 : ibe, k, j) = A3(w, e == 2 ? 2 : ibe - 1, k, jj);
  }
}

// ---------------------------------------------------------------------------
// bytes of dynamic LDS of k_acoustic_col for kde levels
__host__ __device__ inline size_t acoustic_col_lds(int kde) { return (size_t)4 * (kde + 1) * 64 * sizeof(float); }

__global__ void __launch_bounds__(256) k_acoustic_col(const DevCtx *__restrict__ x, float dts, int first, int fin, float ns) {
  extern __shared__ float sm[];
  __shared__ float c_mts[64], c_mav[64], c_dmdt[64];
  const int lane = (int)threadIdx.x, wv = (int)threadIdx.y;
  const int i = (int)blockIdx.x * 64 + lane + 1, j = (int)blockIdx.y + 1;
  const int ide = x->ide, jde = x->jde, kde = x->kde, kend = kde - 1;
  const int S = (kde + 1) * 64;
  float *B0 = sm, *B1 = sm + S, *B2 = sm + 2 * S, *B3 = sm + 3 * S;
#define LS(buf, kk) (buf)[(kk) * 64 + lane]
  const bool mass = (i <= ide - 1 && j <= jde - 1);
  const bool interior = (i >= 2 && i <= ide - 2 && j >= 2 && j <= jde - 2);
  const bool zone = mass && !interior;
  const float g = x->g, t0 = x->t0, epssm = x->epssm;
  const float ope = 1.f + epssm, ome = 1.f - epssm;
  const float *c1h = FL(c1h), *c2h = FL(c2h), *c1f = FL(c1f), *c2f = FL(c2f), *dnw = FL(dnw), *fnm = FL(fnm), *fnp = FL(fnp);
  const float *rdnw = FL(rdnw), *rdn = FL(rdn);
  const float *t_1 = FL(t_save);
  float *ww = FL(ww);

  // ---- advance_mu_t: horizontal mass divergence per level
  if (interior)
    for (int k = 1 + wv; k <= kend; k += 4) LS(B0, k) = mut_dvdxi(x, i, k, j);
  __syncthreads();
  // ---- column sum, column-mass update and the omega recurrence (wave 0)
  if (wv == 0 && interior) {
    float dmdt = 0.f;
    for (int k = 1; k <= kend; k++) dmdt = dmdt + A1(dnw, k) * LS(B0, k);
    float *mu = FL(mu_2);
    const float old = A2(mu, i, j);
    const float mtend = A2(FL(mu_tend), i, j);
    const float tend = dmdt + mtend;
    const float nmu = old + dts * tend;
    const float mts = A2(FL(mut), i, j) + nmu;
    const float mav = 0.5f * ((1.f + epssm) * nmu + (1.f - epssm) * old);
    A2(mu, i, j) = nmu;
    A2(FL(mudf), i, j) = tend;
    A2(FL(muts), i, j) = mts;
    A2(FL(muave), i, j) = mav;
    c_mts[lane] = mts;
    c_mav[lane] = mav;
    c_dmdt[lane] = dmdt;
    const float msfty = A2(FL(msfty), i, j);
    float prev = A3(ww, i, 1, j);
    LS(B1, 1) = prev;
    for (int k = 2; k <= kend; k++) {
      float integ = A1(c1h, k - 1) * dmdt + LS(B0, k - 1);
      integ = integ + A1(c1h, k - 1) * mtend;
      const float cur = prev - (A1(dnw, k - 1) * integ) / msfty;
      LS(B1, k) = cur;
      prev = cur;
    }
  }
  if (wv == 0 && zone) {
    const float mt = A2(FL(mu_tend), i, j);
    A2(FL(mu_2), i, j) = A2(FL(mu_2), i, j) + dts * mt;
    const float mts = A2(FL(muts), i, j) + dts * mt;
    A2(FL(muts), i, j) = mts;
    c_mts[lane] = mts;
  }
  __syncthreads();
  // ---- omega minus the stage-start omega; ww(kde) kept for the phi advection
  if (interior) {
    for (int k = 1 + wv; k <= kend; k += 4) {
      const float v = LS(B1, k) - A3(FL(ww_1), i, k, j);
      A3(ww, i, k, j) = v;
      LS(B1, k) = v;
    }
    if (wv == 0) LS(B1, kde) = A3(ww, i, kde, j);
  }
  __syncthreads();
  // ---- theta (advance_mu_t) and its acoustic average (advance_w); zone theta
  if (interior) {
    const float msfty = A2(FL(msfty), i, j), msftx = A2(FL(msftx), i, j);
    const float mts = c_mts[lane], mav = c_mav[lane];
    const float *u = FL(u_2), *v = FL(v_2), *ft = FL(t_tend);
    float *t = FL(t_2);
    for (int k = 1 + wv; k <= kend; k += 4) {
      const float wlo = (k >= 2) ? LS(B1, k) * (A1(fnm, k) * A3(t_1, i, k, j) + A1(fnp, k) * A3(t_1, i, k - 1, j)) : 0.f;
      const float whi = (k + 1 <= kend) ? LS(B1, k + 1) * (A1(fnm, k + 1) * A3(t_1, i, k + 1, j) + A1(fnp, k + 1) * A3(t_1, i, k, j)) : 0.f;
      const float tk = A3(t, i, k, j);
      const float ta = tk + (msfty * dts) * A3(ft, i, k, j);
      const float vnorth = A3(v, i, k, j + 1) * (A3(t_1, i, k, j + 1) + A3(t_1, i, k, j));
      const float vsouth = A3(v, i, k, j) * (A3(t_1, i, k, j) + A3(t_1, i, k, j - 1));
      const float vft = (0.5f * x->rdy) * (vnorth - vsouth);
      const float ueast = A3(u, i + 1, k, j) * (A3(t_1, i + 1, k, j) + A3(t_1, i, k, j));
      const float uwest = A3(u, i, k, j) * (A3(t_1, i, k, j) + A3(t_1, i - 1, k, j));
      const float uft = (0.5f * x->rdx) * (ueast - uwest);
      const float horiz = msftx * (vft + uft);
      const float vert = A1(rdnw, k) * (whi - wlo);
      const float tnew = ta - (dts * msfty) * (horiz + vert);
      A3(t, i, k, j) = tnew;
      float tav = 0.5f * (ope * tnew + ome * tk);
      tav = (tav + (A1(c1h, k) * mav) * t0) / ((A1(c1h, k) * mts + A1(c2h, k)) * (t0 + A3(t_1, i, k, j)));
      A3(FL(t_2ave), i, k, j) = tav;
      LS(B2, k) = tav;
    }
  } else if (zone) {
    float *t = FL(t_2);
    for (int k = 1 + wv; k <= kend; k += 4) A3(t, i, k, j) = A3(t, i, k, j) + dts * A3(FL(t_tend), i, k, j);
  }
  __syncthreads();
  // ---- advance_w: right-hand side of the phi equation, levels 2..kde
  const float *ph = FL(ph_2);
  const float m = mass ? A2(FL(mut), i, j) : 0.f;
  const float msfty = mass ? A2(FL(msfty), i, j) : 1.f;
  if (interior) {
    const float *ph_1 = FL(ph_save), *phb = FL(phb), *ph_tend = FL(ph_tend), *w = FL(w_2);
#define WDWN(kk) (((0.5f * (LS(B1, kk) + LS(B1, (kk) - 1))) * A1(rdnw, (kk) - 1)) * \
                  (A3(ph_1, i, kk, j) - A3(ph_1, i, (kk) - 1, j) + A3(phb, i, kk, j) - A3(phb, i, (kk) - 1, j)))
    for (int k = 1 + wv; k <= kde; k += 4) {
      if (k == 1) {
        LS(B0, 1) = 0.f;
        continue;
      }
      float rk = dts * (A3(ph_tend, i, k, j) + ((0.5f * g) * ome) * A3(w, i, k, j));
      if (k <= kend) rk = rk - dts * (A1(fnm, k) * WDWN(k + 1) + A1(fnp, k) * WDWN(k));
      rk = A3(ph, i, k, j) + (msfty * rk) / (A1(c1f, k) * m + A1(c2f, k));
      LS(B0, k) = rk;
    }
#undef WDWN
  }
  __syncthreads();
  // ---- explicit w (raw right-hand side of the tridiagonal system)
  if (interior) {
    const float *w = FL(w_2), *rw_tend = FL(rw_tend), *c2a = FL(c2a), *cqw = FL(cqw), *alt = FL(alt);
    const float mav = c_mav[lane], msft_inv = 1.f / msfty;
    for (int k = 1 + wv; k <= kde; k += 4) {
      float wl;
      if (k == 1) {
        const float *ht = FL(ht), *u = FL(u_2), *v = FL(v_2);
        const float cf1 = x->cf1, cf2 = x->cf2, cf3 = x->cf3;
        const float lv = (A2(ht, i, j + 1) - A2(ht, i, j)) * (cf1 * A3(v, i, 1, j + 1) + cf2 * A3(v, i, 2, j + 1) + cf3 * A3(v, i, 3, j + 1)) +
                         (A2(ht, i, j) - A2(ht, i, j - 1)) * (cf1 * A3(v, i, 1, j) + cf2 * A3(v, i, 2, j) + cf3 * A3(v, i, 3, j));
        const float lu = (A2(ht, i + 1, j) - A2(ht, i, j)) * (cf1 * A3(u, i + 1, 1, j) + cf2 * A3(u, i + 1, 2, j) + cf3 * A3(u, i + 1, 3, j)) +
                         (A2(ht, i, j) - A2(ht, i - 1, j)) * (cf1 * A3(u, i, 1, j) + cf2 * A3(u, i, 2, j) + cf3 * A3(u, i, 3, j));
        wl = ((msfty * 0.5f) * x->rdy) * lv + ((A2(FL(msftx), i, j) * 0.5f) * x->rdx) * lu;
      } else if (k <= kend) {
        const float rkp = LS(B0, k + 1), rk0 = LS(B0, k), rkm = LS(B0, k - 1);
        const float pu = ((A3(c2a, i, k, j) * A1(rdnw, k)) / (A1(c1h, k) * m + A1(c2h, k))) *
                         (ope * (rkp - rk0) + ome * (A3(ph, i, k + 1, j) - A3(ph, i, k, j)));
        const float pl = ((A3(c2a, i, k - 1, j) * A1(rdnw, k - 1)) / (A1(c1h, k - 1) * m + A1(c2h, k - 1))) *
                         (ope * (rk0 - rkm) + ome * (A3(ph, i, k, j) - A3(ph, i, k - 1, j)));
        const float ppart = (((0.5f * dts) * g) * A1(rdn, k)) * (pu - pl);
        const float tu = (A3(c2a, i, k, j) * A3(alt, i, k, j)) * LS(B2, k);
        const float tl = (A3(c2a, i, k - 1, j) * A3(alt, i, k - 1, j)) * LS(B2, k - 1);
        const float bpart = A1(rdn, k) * (tu - tl) - A1(c1f, k) * mav;
        wl = ((A3(w, i, k, j) + dts * A3(rw_tend, i, k, j)) + (msft_inv * A3(cqw, i, k, j)) * ppart) + ((dts * g) * msft_inv) * bpart;
      } else {
        float fac = ((-0.5f * dts) * g) / (A1(c1h, k - 1) * m + A1(c2h, k - 1));
        fac = fac * (A1(rdnw, k - 1) * A1(rdnw, k - 1));
        fac = fac * 2.f;
        fac = fac * A3(c2a, i, k - 1, j);
        const float ptop = fac * (ope * (LS(B0, k) - LS(B0, k - 1)) + ome * (A3(ph, i, k, j) - A3(ph, i, k - 1, j)));
        const float btop =
            (-dts * g) * (((2.f * A1(rdnw, k - 1)) * A3(c2a, i, k - 1, j)) * A3(alt, i, k - 1, j) * LS(B2, k - 1) + A1(c1f, k) * mav);
        wl = A3(w, i, k, j) + dts * A3(rw_tend, i, k, j) + (1.f / msfty) * (ptop + btop);
        if (x->top_lid) wl = 0.f;
      }
      LS(B3, k) = wl;
    }
  }
  __syncthreads();
  // ---- tridiagonal sweeps (wave 0, one column per lane)
  if (wv == 0 && interior) {
    const float *a = FL(a), *alpha = FL(alpha), *gamma = FL(gamma);
    float wprev = LS(B3, 1);
#pragma unroll 4
    for (int k = 2; k <= kde; k++) {
      const float wl = (LS(B3, k) - A3(a, i, k, j) * wprev) * A3(alpha, i, k, j);
      LS(B3, k) = wl;
      wprev = wl;
    }
#pragma unroll 4
    for (int k = kend; k >= 2; k--) {
      const float wl = LS(B3, k) - A3(gamma, i, k, j) * wprev;
      LS(B3, k) = wl;
      wprev = wl;
    }
  }
  __syncthreads();
  // ---- new w and phi (interior); spec_bdyupdate_ph (zone); sumflux (all)
  float *php2 = FL(ph_2);
  if (interior) {
    const float mts = c_mts[lane];
    const float pcoef = (((msfty * 0.5f) * dts) * g) * ope;
    float *w = FL(w_2);
    for (int k = 1 + wv; k <= kde; k += 4) {
      const float wl = LS(B3, k);
      A3(w, i, k, j) = wl;
      float phn;
      if (k == 1) phn = A3(php2, i, 1, j);
      else {
        phn = LS(B0, k) + pcoef * wl / (A1(c1f, k) * mts + A1(c2f, k));
        A3(php2, i, k, j) = phn;
      }
      LS(B2, k) = phn;
    }
  } else if (zone) {
    const float mts = c_mts[lane];
    const float mtend = A2(FL(mu_tend), i, j);
    const float mu_old = mts - dts * mtend;
    for (int k = 1 + wv; k <= kde; k += 4) {
      const float den = A1(c1f, k) * mts + A1(c2f, k);
      const float ratio = (A1(c1f, k) * mu_old + A1(c2f, k)) / den;
      const float phn = A3(php2, i, k, j) * ratio + dts * A3(FL(ph_tend), i, k, j) / den + A3(FL(ph_save), i, k, j) * (ratio - 1.f);
      A3(php2, i, k, j) = phn;
      LS(B2, k) = phn;
    }
  }
  {
    // sumflux over the column's u, v and w points (first iteration clears
    // the whole arrays, as the reference's fresh accumulators)
    float *ru_m = FL(ru_m), *rv_m = FL(rv_m), *ww_m = FL(ww_m);
    if (i <= ide && j <= jde)
      for (int k = 1 + wv; k <= kde; k += 4) {
        const size_t o = XI3(i, k, j);
        if (j <= jde - 1 && k <= kend) {
          float acc = (first ? 0.f : ru_m[o]) + FL(u_2)[o];
          if (fin) acc = acc / ns + ((A1(c1h, k) * A2(FL(muu), i, j) + A1(c2h, k)) * FL(u_save)[o]) / A2(FL(msfuy), i, j);
          ru_m[o] = acc;
        } else if (first) {
          ru_m[o] = 0.f;
        }
        if (i <= ide - 1 && k <= kend) {
          float acc = (first ? 0.f : rv_m[o]) + FL(v_2)[o];
          if (fin) acc = acc / ns + ((A1(c1h, k) * A2(FL(muv), i, j) + A1(c2h, k)) * FL(v_save)[o]) * A2(FL(msfvx_inv), i, j);
          rv_m[o] = acc;
        } else if (first) {
          rv_m[o] = 0.f;
        }
        if (mass) {
          float acc = (first ? 0.f : ww_m[o]) + ww[o];
          if (fin) acc = acc / ns + FL(ww_1)[o];
          ww_m[o] = acc;
        } else if (first) {
          ww_m[o] = 0.f;
        }
      }
  }
  __syncthreads();
  // ---- calc_p_rho (iteration >= 1) and the physical-boundary copies
  if (mass) {
    const float mts = c_mts[lane], mu2 = A2(FL(mu_2), i, j);
    const bool gx = (i == ide - 1), gy = (j == jde - 1);
    const float *alt = FL(alt), *t_2 = FL(t_2), *c2a = FL(c2a);
    float *al = FL(al), *p = FL(p), *pm1 = FL(pm1);
    for (int k = 1 + wv; k <= kend; k += 4) {
      const size_t o = XI3(i, k, j);
      const float mass_ = A1(c1h, k) * mts + A1(c2h, k);
      const float dry = A1(c1h, k) * mu2;
      const float altk = alt[o], t1 = t_1[o];
      const float bracket = altk * dry + A1(rdnw, k) * (LS(B2, k + 1) - LS(B2, k));
      const float alv = (-1.f / mass_) * bracket;
      const float tdel = t_2[o] - dry * t1;
      const float frac = (altk * tdel) / (mass_ * (t0 + t1));
      const float pr = c2a[o] * (frac - alv);
      const float pn = pr + x->smdiv * (pr - pm1[o]);
      al[o] = alv;
      p[o] = pn;
      pm1[o] = pr;
      if (gx) {
        A3(al, ide, k, j) = alv;
        A3(p, ide, k, j) = pn;
      }
      if (gy) {
        A3(al, i, k, jde) = alv;
        A3(p, i, k, jde) = pn;
      }
      if (gx && gy) {
        A3(al, ide, k, jde) = alv;
        A3(p, ide, k, jde) = pn;
      }
    }
    if (gx || gy) {
      for (int k = 1 + wv; k <= kde; k += 4) {
        const float v = LS(B2, k);
        if (gx) A3(php2, ide, k, j) = v;
        if (gy) A3(php2, i, k, jde) = v;
        if (gx && gy) A3(php2, ide, k, jde) = v;
      }
      if (wv == 0) {
        const float a2 = A2(FL(muts), i, j), b2 = A2(FL(mu_2), i, j), c2 = A2(FL(mudf), i, j);
        const int is[3] = {ide, i, ide}, js[3] = {j, jde, jde};
        const bool on[3] = {gx, gy, gx && gy};
        for (int q = 0; q < 3; q++)
          if (on[q]) {
            A2(FL(muts), is[q], js[q]) = a2;
            A2(FL(mu_2), is[q], js[q]) = b2;
            A2(FL(mudf), is[q], js[q]) = c2;
          }
      }
    }
  }
#undef LS
}
