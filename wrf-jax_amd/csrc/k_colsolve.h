// Two ways to solve advance_w's tridiagonal w system on the GPU
// (module_small_step_em.py:2949-2999: forward w_k = (r_k - a_k w_{k-1})
// alpha_k for k = 2..kde, backward w_k -= gamma_k w_{k+1} for k = kde-1..2,
// with the LU factors a, alpha, gamma of calc_coef_w), over the interior
// columns i [2, ide-2], j [2, jde-2], for measuring the layouts against each
// other on the same right-hand side r (rhs and the solution may be one
// array: every value of r is read before its level is written):
//
//   k_tridiag_lanes_columns  one column per lane: 64 columns of a row per
//                            wave, both sweeps serial in k (the layout of the
//                            column kernel's sweeps), factors and r read
//                            coalesced along i with 8 levels in flight;
//   k_tridiag_lanes_levels   one column per wavefront: lanes = levels (kde
//                            <= 64), each sweep an affine recurrence
//                            x_k = c_k + d_k x_{k-1} evaluated as an
//                            inclusive scan of affine maps over the lanes in
//                            log2(64) = 6 shuffle steps.  A block stages its
//                            64 columns' factors and r in LDS with coalesced
//                            loads (levels x 65-float stride, no bank
//                            conflicts on the column reads), each of its 8
//                            waves solves 8 columns, and the solution leaves
//                            through LDS with coalesced stores.
//
// The scan reassociates the recurrence, so its result differs from the
// serial sweeps in the last bits (a few ulp); the column kernel keeps the
// serial order, which the reference's rounding requires.
#pragma once
#include "wrfdc_common.h"

__global__ void __launch_bounds__(64) k_tridiag_lanes_columns(const DevCtx xv, const float *rhs, float *wout) {
  const DevCtx *__restrict__ x = &xv;
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const unsigned gx = (unsigned)(ide - 3 + 63) / 64u;
  const int i = 2 + (int)(blockIdx.x % gx) * 64 + (int)threadIdx.x, j = 2 + (int)(blockIdx.x / gx);
  if (i > ide - 2 || j > jde - 2) return;
  const float *a = FL(a), *alpha = FL(alpha), *gamma = FL(gamma);
  constexpr int C = 8;
  float wprev = A3(rhs, i, 1, j);
  A3(wout, i, 1, j) = wprev;
  for (int k0 = 2; k0 <= kde; k0 += C) {
    float ca[C], cl[C], cr[C];
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int k = imin_(k0 + q, kde);
      ca[q] = A3(a, i, k, j);
      cl[q] = A3(alpha, i, k, j);
      cr[q] = A3(rhs, i, k, j);
    }
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int k = k0 + q;
      if (k <= kde) {
        wprev = (cr[q] - ca[q] * wprev) * cl[q];
        A3(wout, i, k, j) = wprev;
      }
    }
  }
  for (int k1 = kde - 1; k1 >= 2; k1 -= C) {
    float cg[C], cw[C];
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int k = imax_(k1 - q, 2);
      cg[q] = A3(gamma, i, k, j);
      cw[q] = A3(wout, i, k, j);
    }
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int k = k1 - q;
      if (k >= 2) {
        wprev = cw[q] - cg[q] * wprev;
        A3(wout, i, k, j) = wprev;
      }
    }
  }
}

// inclusive scan of affine maps x -> c + d x over the lanes, lane 0 first
// (up) or lane 63 first (down): after it, lane l holds the composition of
// the maps of lanes 0..l (or l..63), the later lane's map applied last
__device__ __forceinline__ void affine_scan(float &c, float &d, int lane, bool up) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const float cp = up ? __shfl_up(c, off) : __shfl_down(c, off);
    const float dp = up ? __shfl_up(d, off) : __shfl_down(d, off);
    const bool has = up ? lane >= off : lane + off < 64;
    if (has) {
      c = c + d * cp;
      d = d * dp;
    }
  }
}

constexpr int kTriStride = 65;  // LDS floats per level row (64 columns + 1)
__host__ __device__ inline size_t tridiag_levels_lds(int kde) { return (size_t)4 * kde * kTriStride * sizeof(float); }

__global__ void __launch_bounds__(512) k_tridiag_lanes_levels(const DevCtx xv, const float *rhs, float *wout) {
  const DevCtx *__restrict__ x = &xv;
  extern __shared__ float sm[];
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const int lane = (int)threadIdx.x, wv = (int)threadIdx.y;
  const unsigned gx = (unsigned)(ide - 3 + 63) / 64u;
  const int i0 = 2 + (int)(blockIdx.x % gx) * 64, j = 2 + (int)(blockIdx.x / gx);
  float *SA = sm, *SL = sm + kde * kTriStride, *SG = sm + 2 * kde * kTriStride, *SR = sm + 3 * kde * kTriStride;
  const int ic = imin_(i0 + lane, ide - 2);
  // stage: level k of the 64 columns (coalesced along i), waves over levels
  for (int k = 1 + wv; k <= kde; k += 8) {
    SA[(k - 1) * kTriStride + lane] = A3(FL(a), ic, k, j);
    SL[(k - 1) * kTriStride + lane] = A3(FL(alpha), ic, k, j);
    SG[(k - 1) * kTriStride + lane] = A3(FL(gamma), ic, k, j);
    SR[(k - 1) * kTriStride + lane] = A3(rhs, ic, k, j);
  }
  __syncthreads();
  // solve: wave wv takes columns wv, wv + 8, ...; lane = level - 1
  const int k = lane + 1;
  const bool lev = k <= kde;
  for (int col = wv; col < 64; col += 8) {
    const size_t o = (size_t)(k - 1) * kTriStride + col;
    // forward: w_1 = r_1; w_k = alpha_k r_k - alpha_k a_k w_{k-1}
    float c = 0.f, d = 0.f;
    if (lev) {
      const float al = SL[o], r = SR[o];
      c = k == 1 ? r : r * al;
      d = k == 1 ? 0.f : -SA[o] * al;
    }
    affine_scan(c, d, lane, true);
    // backward: w_k = f_k - gamma_k w_{k+1} for 2 <= k <= kde-1
    float e = c, gq = 0.f;
    if (lev && k >= 2 && k <= kde - 1) gq = -SG[o];
    if (!lev) {
      e = 0.f;
      gq = 0.f;
    }
    affine_scan(e, gq, lane, false);
    if (lev) SR[o] = e;
  }
  __syncthreads();
  if (i0 + lane <= ide - 2)
    for (int kk = 1 + wv; kk <= kde; kk += 8) A3(wout, i0 + lane, kk, j) = SR[(kk - 1) * kTriStride + lane];
}
