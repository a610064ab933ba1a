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
// The column kernel runs 64 columns (one j row, lanes = i) per block of
// eight waves.  Work that is pointwise in k is spread over the waves (wave w
// owns levels w+1, w+1+8, ...) with loads coalesced along i; every phase first
// loads the inputs of a group of its levels into registers, then computes and
// stores, so a wave keeps a group's worth of loads in flight.  The two k
// recurrences (column sum + omega, tridiagonal sweeps) run in wave 0 over
// LDS-resident columns, in the same order as the reference's loops, so every
// rounded operation matches it.
#pragma once
#include "wrfdc_common.h"
#include "k_small.h"

// ---------------------------------------------------------------------------
// vertical pressure difference of advance_uv, rdnw(k) * (P(k+1) - P(k)) with
// P the interface pressure average of uv_iface, for levels k in [1, kde-1],
// without branches around its loads: p is read at levels k-1 .. k+2 (clamped
// into the column) and the surface / interior / top expressions of uv_iface
// are selected on the level
__device__ __forceinline__ float uv_vert(const DevCtx *__restrict__ x, const float *p, int i, int j, int di, int dj, int k) {
  const int kde = x->kde;
  const int km = imax_(k - 1, 1), kp = imin_(k + 1, kde), kq = imin_(k + 2, kde);
  const float ppm = A3(p, i, km, j) + A3(p, i - di, km, j - dj);
  const float pp0 = A3(p, i, k, j) + A3(p, i - di, k, j - dj);
  const float ppp = A3(p, i, kp, j) + A3(p, i - di, kp, j - dj);
  const float ppq = A3(p, i, kq, j) + A3(p, i - di, kq, j - dj);
  // interface k+1 (levels 2 .. kde): interior expression, 0 at the top
  const int lu = imin_(k + 1, kde - 1);
  const float up = (k + 1 <= kde - 1) ? 0.5f * (A1(FL(fnm), lu) * ppp + A1(FL(fnp), lu) * pp0) : 0.f;
  // interface k (levels 1 .. kde-1): surface extrapolation at level 1
  const int ld = imax_(k, 2);
  const float dn = (k == 1) ? 0.5f * (x->cf1 * pp0 + x->cf2 * ppp + x->cf3 * ppq)
                            : 0.5f * (A1(FL(fnm), ld) * pp0 + A1(FL(fnp), ld) * ppm);
  return A1(FL(rdnw), k) * (up - dn);
}
// advance_uv at an interior u point (axis 0: i [2, ide-1], j [2, jde-2]) or v
// point (axis 1: i [2, ide-2], j [2, jde-1]); adv = u + dts * ru_tend (v +
// dts * rv_tend) of that point, formed by the caller
__device__ __forceinline__ float uv_interior(const DevCtx *__restrict__ x, int axis, int i, int k, int j, float dts, float adv) {
  const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
  const unsigned o = XI3(i, k, j);
  const float *p = FL(p), *pb = FL(pb), *ph = FL(ph_2), *alt = FL(alt), *al = FL(al), *php = FL(php);
  const float *c1h = FL(c1h), *c2h = FL(c2h), *mu = FL(mu_2), *mudf = FL(mudf);
  const float spacing = axis == 0 ? x->rdx : x->rdy, grid_length = 1.f / spacing;
  float damping = -1.f * x->emdiv * grid_length * (A2(mudf, i, j) - A2(mudf, i - di, j - dj));
  damping = axis == 0 ? damping / A2(FL(msfuy), i, j) : damping * A2(FL(msfvx_inv), i, j);
  const float phi = (A3(ph, i, k + 1, j) - A3(ph, i - di, k + 1, j - dj)) + (A3(ph, i, k, j) - A3(ph, i - di, k, j - dj));
  const float pres = (A3(alt, i, k, j) + A3(alt, i - di, k, j - dj)) * (A3(p, i, k, j) - A3(p, i - di, k, j - dj));
  const float base = (A3(al, i, k, j) + A3(al, i - di, k, j - dj)) * (A3(pb, i, k, j) - A3(pb, i - di, k, j - dj));
  const float term = (phi + pres) + base;
  const float ratio = axis == 0 ? A2(FL(msfux), i, j) / A2(FL(msfuy), i, j) : A2(FL(msfvy), i, j) / A2(FL(msfvx), i, j);
  const float m = A1(c1h, k) * (axis == 0 ? A2(FL(muu), i, j) : A2(FL(muv), i, j)) + A1(c2h, k);
  float grad = (ratio * 0.5f * spacing * m) * term;
  const float vert = uv_vert(x, p, i, j, di, dj, k);
  const float muv_ = 0.5f * (A1(c1h, k) * A2(mu, i - di, j - dj) + A1(c1h, k) * A2(mu, i, j));
  grad = grad + (ratio * spacing * (A3(php, i, k, j) - A3(php, i - di, k, j - dj))) * (vert - muv_);
  const float cq = axis == 0 ? elem(FL(cqu), o) : elem(FL(cqv), o);
  return (adv - dts * cq * grad) + A1(c1h, k) * damping;
}
// u and v of one acoustic substep (advance_uv plus the specified-zone
// update).  Every lane evaluates the interior formulas at an in-range point
// (its own where they apply) and the zone formula, with no branch around the
// loads, so all of them are in flight together; the results are selected
// and stored at the end.
__global__ void __launch_bounds__(256) k_uv_substep(const DevCtx xv, float dts, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const int ide = x->ide, jde = x->jde;
  const unsigned o = XI3(i, k, j);
  const bool upt = (j <= jde - 1), vpt = (i <= ide - 1);
  const bool uin = upt && i >= 2 && i <= ide - 1 && j >= 2 && j <= jde - 2;
  const bool vin = vpt && i >= 2 && i <= ide - 2 && j >= 2 && j <= jde - 1;
  // u + dts * ru_tend at the point itself: the zone value, and the start of
  // the interior formula (whose result is kept only where the point is
  // interior, i.e. where the clamped point is the point itself)
  const float un_z = elem(FL(u_2), o) + dts * elem(FL(ru_tend), o);
  const float vn_z = elem(FL(v_2), o) + dts * elem(FL(rv_tend), o);
  const float un_in = uv_interior(x, 0, imin_(imax_(i, 2), ide - 1), k, imin_(imax_(j, 2), jde - 2), dts, un_z);
  const float vn_in = uv_interior(x, 1, imin_(imax_(i, 2), ide - 2), k, imin_(imax_(j, 2), jde - 1), dts, vn_z);
  if (upt) elem(FL(u_2), o) = uin ? un_in : un_z;
  if (vpt) elem(FL(v_2), o) = vin ? vn_in : vn_z;
}

// zero_grad_bdy of w: every destination (the zone ring of the mass grid)
// copies an interior point, so rows and columns go in one race-free pass.
__global__ void k_zero_grad_w(const DevCtx xv) {
  const DevCtx *__restrict__ x = &xv;
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
// waves per block of k_acoustic_col, and levels per register group of each
// phase (the loads of one group are issued together).  The group sizes trade
// memory round trips against registers: phases with few loads per level take
// larger groups (phase 3: 2 loads per level, all 8 levels of a wave at once;
// phases 5, 6, 10: 4 levels), phases 1, 4 and 9 stay at 2 so the kernel keeps
// 6 waves per SIMD (at most 80 VGPRs, the LDS limit being 3 blocks per CU).
#ifndef WRFDC_COL_WAVES
#define WRFDC_COL_WAVES 8
#endif
#ifndef WRFDC_COL_GROUP
#define WRFDC_COL_GROUP 2
#endif
constexpr int kColG1 = WRFDC_COL_GROUP, kColG3 = 4 * WRFDC_COL_GROUP, kColG4 = WRFDC_COL_GROUP, kColG5 = 2 * WRFDC_COL_GROUP,
              kColG6 = 2 * WRFDC_COL_GROUP, kColG9 = WRFDC_COL_GROUP, kColG10 = 2 * WRFDC_COL_GROUP;
constexpr int kColWaves = WRFDC_COL_WAVES;

constexpr int kColMaxLev = 152;  // levels the column kernel holds (nz <= 150)
// bytes of dynamic LDS of k_acoustic_col for kde levels
__host__ __device__ inline size_t acoustic_col_lds(int kde) { return (size_t)3 * (kde + 1) * 64 * sizeof(float); }
// levels per wave the kernel is instantiated for (kde <= kColWaves * NPW)
inline int acoustic_col_npw(int kde) {
  const int need = (kde + kColWaves - 1) / kColWaves;
  return need <= 4 ? 4 : need <= 8 ? 8 : need <= 12 ? 12 : need <= 16 ? 16 : 20;
}

template <int NPW>
__global__ void __launch_bounds__(64 * kColWaves) k_acoustic_col(const DevCtx xv, float dts, int first, int fin, float ns, int store_tav) {
  const DevCtx *__restrict__ x = &xv;
  constexpr int NW = kColWaves;
  constexpr int G1 = kColG1, G3 = kColG3, G4 = kColG4, G5 = kColG5, G6 = kColG6, G9 = kColG9, G10 = kColG10;
  extern __shared__ float sm[];
  __shared__ float c_mts[64], c_mav[64];
  // dnw and c1h in LDS for the serial column sum and omega recurrence (their
  // global loads would otherwise sit on the recurrence's critical path)
  __shared__ float c_dnw[kColMaxLev], c_c1h[kColMaxLev];
  const int lane = (int)threadIdx.x, wv = (int)threadIdx.y;
  const unsigned gx = (unsigned)(x->ide + 63) / 64u, t = xcd_tile(blockIdx.x, gridDim.x);
  const int i = (int)(t % gx) * 64 + lane + 1, j = (int)(t / gx) + 1;
  const int ide = x->ide, jde = x->jde, kde = x->kde, kend = kde - 1;
  const int S = (kde + 1) * 64;
  // B0: divergence, then the phi right-hand side; B1: omega, then w (the
  // tridiagonal unknowns); B2: the acoustic theta average, then final phi
  float *B0 = sm, *B1 = sm + S, *B2 = sm + 2 * S, *B3 = B1;
#define LS(buf, kk) (buf)[(kk) * 64 + lane]
#define LEV(n) (1 + wv + NW * (n))
  const bool mass = (i <= ide - 1 && j <= jde - 1);
  const bool interior = (i >= 2 && i <= ide - 2 && j >= 2 && j <= jde - 2);
  const bool zone = mass && !interior;
  const float g = x->g, t0 = x->t0, epssm = x->epssm;
  const float ope = 1.f + epssm, ome = 1.f - epssm;
  const float *c1h = FL(c1h), *c2h = FL(c2h), *c1f = FL(c1f), *c2f = FL(c2f), *dnw = FL(dnw), *fnm = FL(fnm), *fnp = FL(fnp);
  const float *rdnw = FL(rdnw), *rdn = FL(rdn);
  const float *t_1 = FL(t_save), *u = FL(u_2), *v = FL(v_2);
  float *ww = FL(ww);
  const int ic = mass ? i : 1, jc = mass ? j : 1;  // clamped column for address arithmetic
  const float msfty = A2(FL(msfty), ic, jc), msftx = A2(FL(msftx), ic, jc);

  float *ru_m = FL(ru_m), *rv_m = FL(rv_m), *ww_m = FL(ww_m);
  for (int q = (int)(threadIdx.y * 64 + threadIdx.x); q < kde; q += 64 * NW) {
    c_dnw[q] = dnw[q];
    c_c1h[q] = c1h[q];
  }
  // ---- 1. advance_mu_t: horizontal mass divergence per level -> B0; with
  // the same loads, sumflux of the column's u and v points (levels 1..kde-1)
  if (interior) {
    const float *us = FL(u_save), *vs = FL(v_save), *muu = FL(muu), *muv = FL(muv);
    const float muu0 = A2(muu, i, j), muu1 = A2(muu, i + 1, j), muv0 = A2(muv, i, j), muv1 = A2(muv, i, j + 1);
    const float iuy0 = A2(FL(msfuy), i, j), iuy1 = A2(FL(msfuy), i + 1, j);
    const float vx0 = A2(FL(msfvx_inv), i, j), vx1 = A2(FL(msfvx_inv), i, j + 1);
    const float mm = msftx * msfty;
#pragma unroll
    for (int gq = 0; gq < NPW; gq += G1) {
      float r_ue[G1], r_uw[G1], r_use[G1], r_usw[G1], r_vn[G1], r_vs[G1], r_vsn[G1], r_vss[G1], r_rum[G1], r_rvm[G1];
#pragma unroll
      for (int q = 0; q < G1; q++) {
        const int k = imin_(LEV(gq + q), kend);
        r_rum[q] = first ? 0.f : A3(ru_m, i, k, j);
        r_rvm[q] = first ? 0.f : A3(rv_m, i, k, j);
        r_ue[q] = A3(u, i + 1, k, j);
        r_uw[q] = A3(u, i, k, j);
        r_use[q] = A3(us, i + 1, k, j);
        r_usw[q] = A3(us, i, k, j);
        r_vn[q] = A3(v, i, k, j + 1);
        r_vs[q] = A3(v, i, k, j);
        r_vsn[q] = A3(vs, i, k, j + 1);
        r_vss[q] = A3(vs, i, k, j);
      }
#pragma unroll
      for (int q = 0; q < G1; q++) {
        const int k = LEV(gq + q);
        if (k > kend) break;
        const float c1 = A1(c1h, k), c2 = A1(c2h, k);
        const float vn = r_vn[q] + ((c1 * muv1 + c2) * r_vsn[q]) * vx1;
        const float vsv = r_vs[q] + ((c1 * muv0 + c2) * r_vss[q]) * vx0;
        const float ue = r_ue[q] + ((c1 * muu1 + c2) * r_use[q]) / iuy1;
        const float uw = r_uw[q] + ((c1 * muu0 + c2) * r_usw[q]) / iuy0;
        const float hf = x->rdy * (vn - vsv) + x->rdx * (ue - uw);
        LS(B0, k) = mm * hf;
        float au = r_rum[q] + r_uw[q];
        if (fin) au = au / ns + ((c1 * muu0 + c2) * r_usw[q]) / iuy0;
        A3(ru_m, i, k, j) = au;
        float av = r_rvm[q] + r_vs[q];
        if (fin) av = av / ns + ((c1 * muv0 + c2) * r_vss[q]) * vx0;
        A3(rv_m, i, k, j) = av;
      }
    }
  }
  __syncthreads();
  // ---- 2. column sum, column-mass update and the omega recurrence (wave 0)
  if (wv == 0 && interior) {
    // column mass tendency: FP32 level terms summed in FP64, rounded once
    double dmdt_acc = 0.0;
    for (int k = 1; k <= kend; k++) dmdt_acc += (double)(c_dnw[k - 1] * LS(B0, k));
    const float dmdt = (float)dmdt_acc;
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
    float prev = A3(ww, i, 1, j);
    LS(B1, 1) = prev;
    for (int k = 2; k <= kend; k++) {
      float integ = c_c1h[k - 2] * dmdt + LS(B0, k - 1);
      integ = integ + c_c1h[k - 2] * mtend;
      const float cur = prev - (c_dnw[k - 2] * integ) / msfty;
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
  // ---- 3. omega minus the stage-start omega (B1 = final omega, k = 1..kde),
  // and sumflux of the column's w points
  if (interior) {
    const float *ww_1 = FL(ww_1);
#pragma unroll
    for (int gq = 0; gq < NPW; gq += G3) {
      float r[G3], r_wwm[G3];
#pragma unroll
      for (int q = 0; q < G3; q++) {
        const int k = imin_(LEV(gq + q), kend);
        r[q] = A3(ww_1, i, k, j);
        r_wwm[q] = first ? 0.f : A3(ww_m, i, k, j);
      }
#pragma unroll
      for (int q = 0; q < G3; q++) {
        const int k = LEV(gq + q);
        if (k > kend) break;
        const float vv = LS(B1, k) - r[q];
        A3(ww, i, k, j) = vv;
        LS(B1, k) = vv;
        float aw = r_wwm[q] + vv;
        if (fin) aw = aw / ns + r[q];
        A3(ww_m, i, k, j) = aw;
      }
    }
    if (wv == NW - 1) {
      const float wtop = A3(ww, i, kde, j);
      LS(B1, kde) = wtop;
      float aw = (first ? 0.f : A3(ww_m, i, kde, j)) + wtop;
      if (fin) aw = aw / ns + A3(ww_1, i, kde, j);
      A3(ww_m, i, kde, j) = aw;
    }
  }
  __syncthreads();
  // ---- 4. theta (advance_mu_t) and its acoustic average (advance_w) -> B2
  // (phase 10 reads the new theta back: the same thread wrote it)
  if (interior) {
    const float mts = c_mts[lane], mav = c_mav[lane];
    const float *ft = FL(t_tend);
    float *t = FL(t_2), *tav_out = FL(t_2ave);
#pragma unroll
    for (int gq = 0; gq < NPW; gq += G4) {
      float r_t1[G4], r_t1m[G4], r_t1p[G4], r_t1e[G4], r_t1w[G4], r_t1n[G4], r_t1s[G4], r_ue[G4], r_uw[G4], r_vn[G4], r_vs[G4], r_t[G4], r_ft[G4];
#pragma unroll
      for (int q = 0; q < G4; q++) {
        const int k = imin_(LEV(gq + q), kend);
        r_t1[q] = A3(t_1, i, k, j);
        r_t1m[q] = A3(t_1, i, imax_(k - 1, 1), j);
        r_t1p[q] = A3(t_1, i, imin_(k + 1, kend), j);
        r_t1e[q] = A3(t_1, i + 1, k, j);
        r_t1w[q] = A3(t_1, i - 1, k, j);
        r_t1n[q] = A3(t_1, i, k, j + 1);
        r_t1s[q] = A3(t_1, i, k, j - 1);
        r_ue[q] = A3(u, i + 1, k, j);
        r_uw[q] = A3(u, i, k, j);
        r_vn[q] = A3(v, i, k, j + 1);
        r_vs[q] = A3(v, i, k, j);
        r_t[q] = A3(t, i, k, j);
        r_ft[q] = A3(ft, i, k, j);
      }
#pragma unroll
      for (int q = 0; q < G4; q++) {
        const int k = LEV(gq + q);
        if (k > kend) break;
        const float wlo = (k >= 2) ? LS(B1, k) * (A1(fnm, k) * r_t1[q] + A1(fnp, k) * r_t1m[q]) : 0.f;
        const float whi = (k + 1 <= kend) ? LS(B1, k + 1) * (A1(fnm, k + 1) * r_t1p[q] + A1(fnp, k + 1) * r_t1[q]) : 0.f;
        const float tk = r_t[q];
        const float ta = tk + (msfty * dts) * r_ft[q];
        const float vnorth = r_vn[q] * (r_t1n[q] + r_t1[q]);
        const float vsouth = r_vs[q] * (r_t1[q] + r_t1s[q]);
        const float vft = (0.5f * x->rdy) * (vnorth - vsouth);
        const float ueast = r_ue[q] * (r_t1e[q] + r_t1[q]);
        const float uwest = r_uw[q] * (r_t1[q] + r_t1w[q]);
        const float uft = (0.5f * x->rdx) * (ueast - uwest);
        const float horiz = msftx * (vft + uft);
        const float vert = A1(rdnw, k) * (whi - wlo);
        const float tnew = ta - (dts * msfty) * (horiz + vert);
        A3(t, i, k, j) = tnew;
        float tav = 0.5f * (ope * tnew + ome * tk);
        tav = (tav + (A1(c1h, k) * mav) * t0) / ((A1(c1h, k) * mts + A1(c2h, k)) * (t0 + r_t1[q]));
        if (store_tav) A3(tav_out, i, k, j) = tav;
        LS(B2, k) = tav;
      }
    }
  } else if (zone) {
    float *t = FL(t_2);
    const float *ft = FL(t_tend);
    for (int k = 1 + wv; k <= kend; k += NW) A3(t, i, k, j) = A3(t, i, k, j) + dts * A3(ft, i, k, j);
  }
  // (no barrier: phase 5 reads omega (final since the barrier above), phi and
  // w, and writes B0; phase 4 writes theta, its average and B2 only)
  // ---- 5. advance_w: right-hand side of the phi equation (levels 2..kde) -> B0
  const float *ph = FL(ph_2);
  const float m = A2(FL(mut), ic, jc);
  if (interior) {
    const float *ph_1 = FL(ph_save), *phb = FL(phb), *ph_tend = FL(ph_tend), *w = FL(w_2);
#pragma unroll
    for (int gq = 0; gq < NPW; gq += G5) {
      float r_pt[G5], r_w[G5], r_p1[G5], r_p1m[G5], r_p1p[G5], r_pb[G5], r_pbm[G5], r_pbp[G5], r_ph[G5];
#pragma unroll
      for (int q = 0; q < G5; q++) {
        const int k = imax_(imin_(LEV(gq + q), kde), 2);
        r_pt[q] = A3(ph_tend, i, k, j);
        r_w[q] = A3(w, i, k, j);
        r_p1[q] = A3(ph_1, i, k, j);
        r_p1m[q] = A3(ph_1, i, k - 1, j);
        r_p1p[q] = A3(ph_1, i, imin_(k + 1, kde), j);
        r_pb[q] = A3(phb, i, k, j);
        r_pbm[q] = A3(phb, i, k - 1, j);
        r_pbp[q] = A3(phb, i, imin_(k + 1, kde), j);
        r_ph[q] = A3(ph, i, k, j);
      }
#pragma unroll
      for (int q = 0; q < G5; q++) {
        const int k = LEV(gq + q);
        if (k > kde) break;
        if (k == 1) {
          LS(B0, 1) = 0.f;
          continue;
        }
        float rk = dts * (r_pt[q] + ((0.5f * g) * ome) * r_w[q]);
        if (k <= kend) {
          const float wdn_k = ((0.5f * (LS(B1, k) + LS(B1, k - 1))) * A1(rdnw, k - 1)) * (r_p1[q] - r_p1m[q] + r_pb[q] - r_pbm[q]);
          const float wdn_kp = ((0.5f * (LS(B1, k + 1) + LS(B1, k))) * A1(rdnw, k)) * (r_p1p[q] - r_p1[q] + r_pbp[q] - r_pb[q]);
          rk = rk - dts * (A1(fnm, k) * wdn_kp + A1(fnp, k) * wdn_k);
        }
        rk = r_ph[q] + (msfty * rk) / (A1(c1f, k) * m + A1(c2f, k));
        LS(B0, k) = rk;
      }
    }
  }
  __syncthreads();
  // ---- 6. explicit w (raw right-hand side of the tridiagonal system) -> B3
  if (interior) {
    const float *w = FL(w_2), *rw_tend = FL(rw_tend), *c2a = FL(c2a), *cqw = FL(cqw), *alt = FL(alt);
    const float mav = c_mav[lane], msft_inv = 1.f / msfty;
#pragma unroll
    for (int gq = 0; gq < NPW; gq += G6) {
      float r_c2a[G6], r_c2am[G6], r_alt[G6], r_altm[G6], r_ph[G6], r_phm[G6], r_php[G6], r_w[G6], r_rwt[G6], r_cqw[G6];
#pragma unroll
      for (int q = 0; q < G6; q++) {
        const int k = imax_(imin_(LEV(gq + q), kde), 2);
        r_c2a[q] = A3(c2a, i, imin_(k, kend), j);
        r_c2am[q] = A3(c2a, i, k - 1, j);
        r_alt[q] = A3(alt, i, imin_(k, kend), j);
        r_altm[q] = A3(alt, i, k - 1, j);
        r_ph[q] = A3(ph, i, k, j);
        r_phm[q] = A3(ph, i, k - 1, j);
        r_php[q] = A3(ph, i, imin_(k + 1, kde), j);
        r_w[q] = A3(w, i, k, j);
        r_rwt[q] = A3(rw_tend, i, k, j);
        r_cqw[q] = A3(cqw, i, k, j);
      }
#pragma unroll
      for (int q = 0; q < G6; q++) {
        const int k = LEV(gq + q);
        if (k > kde) break;
        float wl;
        if (k == 1) {
          const float *ht = FL(ht);
          const float cf1 = x->cf1, cf2 = x->cf2, cf3 = x->cf3;
          const float lv = (A2(ht, i, j + 1) - A2(ht, i, j)) * (cf1 * A3(v, i, 1, j + 1) + cf2 * A3(v, i, 2, j + 1) + cf3 * A3(v, i, 3, j + 1)) +
                           (A2(ht, i, j) - A2(ht, i, j - 1)) * (cf1 * A3(v, i, 1, j) + cf2 * A3(v, i, 2, j) + cf3 * A3(v, i, 3, j));
          const float lu = (A2(ht, i + 1, j) - A2(ht, i, j)) * (cf1 * A3(u, i + 1, 1, j) + cf2 * A3(u, i + 1, 2, j) + cf3 * A3(u, i + 1, 3, j)) +
                           (A2(ht, i, j) - A2(ht, i - 1, j)) * (cf1 * A3(u, i, 1, j) + cf2 * A3(u, i, 2, j) + cf3 * A3(u, i, 3, j));
          wl = ((msfty * 0.5f) * x->rdy) * lv + ((msftx * 0.5f) * x->rdx) * lu;
        } else if (k <= kend) {
          const float rkp = LS(B0, k + 1), rk0 = LS(B0, k), rkm = LS(B0, k - 1);
          const float pu = ((r_c2a[q] * A1(rdnw, k)) / (A1(c1h, k) * m + A1(c2h, k))) * (ope * (rkp - rk0) + ome * (r_php[q] - r_ph[q]));
          const float pl = ((r_c2am[q] * A1(rdnw, k - 1)) / (A1(c1h, k - 1) * m + A1(c2h, k - 1))) * (ope * (rk0 - rkm) + ome * (r_ph[q] - r_phm[q]));
          const float ppart = (((0.5f * dts) * g) * A1(rdn, k)) * (pu - pl);
          const float tu = (r_c2a[q] * r_alt[q]) * LS(B2, k);
          const float tl = (r_c2am[q] * r_altm[q]) * LS(B2, k - 1);
          const float bpart = A1(rdn, k) * (tu - tl) - A1(c1f, k) * mav;
          wl = ((r_w[q] + dts * r_rwt[q]) + (msft_inv * r_cqw[q]) * ppart) + ((dts * g) * msft_inv) * bpart;
        } else {
          float fac = ((-0.5f * dts) * g) / (A1(c1h, k - 1) * m + A1(c2h, k - 1));
          fac = fac * (A1(rdnw, k - 1) * A1(rdnw, k - 1));
          fac = fac * 2.f;
          fac = fac * r_c2am[q];
          const float ptop = fac * (ope * (LS(B0, k) - LS(B0, k - 1)) + ome * (r_ph[q] - r_phm[q]));
          const float btop = (-dts * g) * (((2.f * A1(rdnw, k - 1)) * r_c2am[q]) * r_altm[q] * LS(B2, k - 1) + A1(c1f, k) * mav);
          wl = r_w[q] + dts * r_rwt[q] + (1.f / msfty) * (ptop + btop);
          if (x->top_lid) wl = 0.f;
        }
        LS(B3, k) = wl;
      }
    }
  }
  __syncthreads();
  // ---- 7. tridiagonal sweeps (wave 0, one column per lane).  Forward: the
  // LU factors a, alpha come from global memory in chunks of C levels (the
  // factors of chunk c+1 are requested before the recurrence of chunk c runs,
  // so their latency overlaps it).  Meanwhile the other waves copy gamma of
  // every level into B2 (free between phase 6 and phase 9), so the backward
  // sweep reads it from LDS.
  const int sk = x->ip;  // element stride between levels
  const unsigned o1 = XI3(ic, 1, jc);
  float wprev = 0.f;
  if (wv == 0 && interior) {
    constexpr int C = 8;
    const float *a = FL(a), *alpha = FL(alpha);
    wprev = LS(B3, 1);
    // forward: levels 2..kde
    {
      const int n = kde - 1;
      float ca[C], cl[C], na[C], nl[C];
#pragma unroll
      for (int q = 0; q < C; q++) {
        const int k = imin_(2 + q, kde);
        ca[q] = a[o1 + (unsigned)((k - 1) * sk)];
        cl[q] = alpha[o1 + (unsigned)((k - 1) * sk)];
      }
      for (int c0 = 0; c0 < n; c0 += C) {
#pragma unroll
        for (int q = 0; q < C; q++) {
          const int k = imin_(2 + c0 + C + q, kde);
          na[q] = a[o1 + (unsigned)((k - 1) * sk)];
          nl[q] = alpha[o1 + (unsigned)((k - 1) * sk)];
        }
#pragma unroll
        for (int q = 0; q < C; q++) {
          const int k = 2 + c0 + q;
          if (k <= kde) {
            const float wl = (LS(B3, k) - ca[q] * wprev) * cl[q];
            LS(B3, k) = wl;
            wprev = wl;
          }
        }
#pragma unroll
        for (int q = 0; q < C; q++) {
          ca[q] = na[q];
          cl[q] = nl[q];
        }
      }
    }
  } else if (interior) {
    const float *gamma = FL(gamma);
    for (int k = 1 + wv; k <= kend; k += NW - 1) LS(B2, k) = gamma[o1 + (unsigned)((k - 1) * sk)];
  }
  __syncthreads();
  // backward: levels kend..2
  if (wv == 0 && interior) {
#pragma unroll 8
    for (int k = kend; k >= 2; k--) {
      const float wl = LS(B3, k) - LS(B2, k) * wprev;
      LS(B3, k) = wl;
      wprev = wl;
    }
  }
  __syncthreads();
  // ---- 9. new w and phi (interior); spec_bdyupdate_ph (zone) -> B2 = final phi
  float *php2 = FL(ph_2);
  if (interior) {
    const float mts = c_mts[lane];
    const float pcoef = (((msfty * 0.5f) * dts) * g) * ope;
    float *w = FL(w_2);
    for (int k = 1 + wv; k <= kde; k += NW) {
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
    const float *ph_tend = FL(ph_tend), *ph_save = FL(ph_save);
#pragma unroll
    for (int gq = 0; gq < NPW; gq += G9) {
      float r_ph[G9], r_pt[G9], r_ps[G9];
#pragma unroll
      for (int q = 0; q < G9; q++) {
        const int k = imin_(LEV(gq + q), kde);
        r_ph[q] = A3(php2, i, k, j);
        r_pt[q] = A3(ph_tend, i, k, j);
        r_ps[q] = A3(ph_save, i, k, j);
      }
#pragma unroll
      for (int q = 0; q < G9; q++) {
        const int k = LEV(gq + q);
        if (k > kde) break;
        const float den = A1(c1f, k) * mts + A1(c2f, k);
        const float ratio = (A1(c1f, k) * mu_old + A1(c2f, k)) / den;
        const float phn = r_ph[q] * ratio + dts * r_pt[q] / den + r_ps[q] * (ratio - 1.f);
        A3(php2, i, k, j) = phn;
        LS(B2, k) = phn;
      }
    }
  }
  // ---- sumflux over the u, v and w points of the columns phases 1 and 3
  // did not cover (the first iteration clears the whole arrays, as the
  // reference's fresh accumulators)
  if (interior) {
    if (first && wv == 0) {
      A3(ru_m, i, kde, j) = 0.f;
      A3(rv_m, i, kde, j) = 0.f;
    }
  } else if (i <= ide && j <= jde) {
    const bool upt = (j <= jde - 1), vpt = (i <= ide - 1);
    const float cmu_u = upt ? A2(FL(muu), i, j) : 0.f, cmsf_u = upt ? A2(FL(msfuy), i, j) : 1.f;
    const float cmu_v = vpt ? A2(FL(muv), i, j) : 0.f, cmsf_v = vpt ? A2(FL(msfvx_inv), i, j) : 1.f;
#pragma unroll
    for (int gq = 0; gq < NPW; gq += G9) {
      float r_rum[G9], r_rvm[G9], r_wwm[G9], r_u[G9], r_v[G9], r_ww[G9], r_us[G9], r_vs[G9], r_ww1[G9];
#pragma unroll
      for (int q = 0; q < G9; q++) {
        const int k = imin_(LEV(gq + q), kde);
        const size_t o = XI3(i, k, j);
        r_rum[q] = first ? 0.f : ru_m[o];
        r_rvm[q] = first ? 0.f : rv_m[o];
        r_wwm[q] = first ? 0.f : ww_m[o];
        r_u[q] = upt ? u[o] : 0.f;
        r_v[q] = vpt ? v[o] : 0.f;
        r_ww[q] = mass ? ww[o] : 0.f;
        r_us[q] = (fin && upt) ? FL(u_save)[o] : 0.f;
        r_vs[q] = (fin && vpt) ? FL(v_save)[o] : 0.f;
        r_ww1[q] = (fin && mass) ? FL(ww_1)[o] : 0.f;
      }
#pragma unroll
      for (int q = 0; q < G9; q++) {
        const int k = LEV(gq + q);
        if (k > kde) break;
        const size_t o = XI3(i, k, j);
        if (upt && k <= kend) {
          float acc = r_rum[q] + r_u[q];
          if (fin) acc = acc / ns + ((A1(c1h, k) * cmu_u + A1(c2h, k)) * r_us[q]) / cmsf_u;
          ru_m[o] = acc;
        } else if (first) {
          ru_m[o] = 0.f;
        }
        if (vpt && k <= kend) {
          float acc = r_rvm[q] + r_v[q];
          if (fin) acc = acc / ns + ((A1(c1h, k) * cmu_v + A1(c2h, k)) * r_vs[q]) * cmsf_v;
          rv_m[o] = acc;
        } else if (first) {
          rv_m[o] = 0.f;
        }
        if (mass) {
          float acc = r_wwm[q] + r_ww[q];
          if (fin) acc = acc / ns + r_ww1[q];
          ww_m[o] = acc;
        } else if (first) {
          ww_m[o] = 0.f;
        }
      }
    }
  }
  __syncthreads();
  // ---- 10. calc_p_rho (iteration >= 1) and the physical-boundary copies
  if (mass) {
    const float mts = c_mts[lane], mu2 = A2(FL(mu_2), i, j);
    const bool gx = (i == ide - 1), gy = (j == jde - 1);
    const float *alt = FL(alt), *t_2 = FL(t_2), *c2a = FL(c2a);
    float *al = FL(al), *p = FL(p), *pm1 = FL(pm1);
#pragma unroll
    for (int gq = 0; gq < NPW; gq += G10) {
      float r_alt[G10], r_t1[G10], r_t2[G10], r_c2a[G10], r_pm1[G10];
#pragma unroll
      for (int q = 0; q < G10; q++) {
        const size_t o = XI3(i, imin_(LEV(gq + q), kend), j);
        r_alt[q] = alt[o];
        r_t1[q] = t_1[o];
        r_t2[q] = t_2[o];
        r_c2a[q] = c2a[o];
        r_pm1[q] = pm1[o];
      }
#pragma unroll
      for (int q = 0; q < G10; q++) {
        const int k = LEV(gq + q);
        if (k > kend) break;
        const size_t o = XI3(i, k, j);
        const float mass_ = A1(c1h, k) * mts + A1(c2h, k);
        const float dry = A1(c1h, k) * mu2;
        const float bracket = r_alt[q] * dry + A1(rdnw, k) * (LS(B2, k + 1) - LS(B2, k));
        const float alv = (-1.f / mass_) * bracket;
        const float tdel = r_t2[q] - dry * r_t1[q];
        const float frac = (r_alt[q] * tdel) / (mass_ * (t0 + r_t1[q]));
        const float pr = r_c2a[q] * (frac - alv);
        const float pn = pr + x->smdiv * (pr - r_pm1[q]);
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
    }
    if (gx || gy) {
      for (int k = 1 + wv; k <= kde; k += NW) {
        const float vv = LS(B2, k);
        if (gx) A3(php2, ide, k, j) = vv;
        if (gy) A3(php2, i, k, jde) = vv;
        if (gx && gy) A3(php2, ide, k, jde) = vv;
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
#undef LEV
}

// host launcher: instantiation by levels per wave
// store_tav: write the acoustic theta average to t_2ave (only its last
// substep's value is read outside the kernel)
inline void launch_acoustic_col(const DevCtx &d, hipStream_t s, float dts, int first, int fin, float ns, int store_tav) {
  const dim3 grid((unsigned)((d.ide + 63) / 64) * (unsigned)d.jde), block(64, kColWaves);
  const size_t lds = acoustic_col_lds(d.kde);
  switch (acoustic_col_npw(d.kde)) {
    case 4: k_acoustic_col<4><<<grid, block, lds, s>>>(d, dts, first, fin, ns, store_tav); break;
    case 8: k_acoustic_col<8><<<grid, block, lds, s>>>(d, dts, first, fin, ns, store_tav); break;
    case 12: k_acoustic_col<12><<<grid, block, lds, s>>>(d, dts, first, fin, ns, store_tav); break;
    case 16: k_acoustic_col<16><<<grid, block, lds, s>>>(d, dts, first, fin, ns, store_tav); break;
    default: k_acoustic_col<20><<<grid, block, lds, s>>>(d, dts, first, fin, ns, store_tav); break;
  }
}
// raise the dynamic-LDS limit of the instantiation kde needs (beyond 64 KB)
inline hipError_t acoustic_col_allow_lds(int kde) {
  const int lds = (int)acoustic_col_lds(kde);
  if (lds <= 65536) return hipSuccess;
  switch (acoustic_col_npw(kde)) {
    case 4: return hipFuncSetAttribute((const void *)k_acoustic_col<4>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    case 8: return hipFuncSetAttribute((const void *)k_acoustic_col<8>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    case 12: return hipFuncSetAttribute((const void *)k_acoustic_col<12>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    case 16: return hipFuncSetAttribute((const void *)k_acoustic_col<16>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    default: return hipFuncSetAttribute((const void *)k_acoustic_col<20>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  }
}
