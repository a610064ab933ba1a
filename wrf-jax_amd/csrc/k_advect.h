// Flux-form advection kernels: 5th-order horizontal / 3rd-order vertical with
// the reference's boundary-degraded orders for specified boundaries
// (src/wrf_jax/dyn_em/module_advect_em.py: advect_u/v/w/scalar/scalar_pd).
// Each thread owns one output point and evaluates the fluxes of its faces.
#pragma once
#include "wrfdc_common.h"
#include "k_bigstep.h"

__device__ __forceinline__ float flux3(float qm2, float qm1, float q0, float qp1, float ua) {
  return div_const(7.f * (q0 + qm1) - (qp1 + qm2), 12.f, 1.f / 12.f) +
         div_const(sgn1(ua) * ((qp1 - qm2) - 3.f * (q0 - qm1)), 12.f, 1.f / 12.f);
}
__device__ __forceinline__ float flux5(float qm3, float qm2, float qm1, float q0, float qp1, float qp2, float ua) {
  return div_const(37.f * (q0 + qm1) - 8.f * (qp1 + qm2) + (qp2 + qm3), 60.f, 1.f / 60.f) -
         div_const(sgn1(ua) * ((qp2 - qm3) - 5.f * (qp1 - qm2) + 10.f * (q0 - qm1)), 60.f, 1.f / 60.f);
}
__device__ __forceinline__ float flux3pd(float qm2, float qm1, float q0, float qp1, float ua) {
  float f4 = (7.f / 12.f) * (q0 + qm1) - (1.f / 12.f) * (qp1 + qm2);
  return f4 + sgn1(ua) * (1.f / 12.f) * ((qp1 - qm2) - 3.f * (q0 - qm1));
}
__device__ __forceinline__ float flux5pd(float qm3, float qm2, float qm1, float q0, float qp1, float qp2, float ua) {
  float f6 = (37.f / 60.f) * (q0 + qm1) - (2.f / 15.f) * (qp1 + qm2) + (1.f / 60.f) * (qp2 + qm3);
  return f6 - sgn1(ua) * (1.f / 60.f) * ((qp2 - qm3) - 5.f * (qp1 - qm2) + 10.f * (q0 - qm1));
}
__device__ __forceinline__ float flux_up(float qm1, float q0, float cr) {
  return 0.5f * fminf(1.f, cr + fabsf(cr)) * qm1 + 0.5f * fmaxf(-1.f, cr - fabsf(cr)) * q0;
}

// ---------------- face-sharing blocks ---------------------------------------
// value of v in the next lane of the wave (lane 63 receives lane 0's)
__device__ __forceinline__ float lane_next(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(((lane + 1) & 63) << 2, __float_as_int(v)));
}
// Blocks of the flux-form kernels: 64 lanes along i, of which the first 63
// own a point (lane 63 only forms the west face its left neighbour needs as
// east face), 4 waves on 4 consecutive rows j0..j0+3 (each wave forms the
// south face of its row, the row below reads it from LDS as north face; the
// top wave also forms its own north face), one level k.  Each x face flux is
// evaluated once and each y face 5/4 times instead of twice; every thread
// still owns one point, so the grid keeps one wave per 63 points.
// Tile order: i fastest, then 4 consecutive levels, then rows, then groups
// of 4 levels, XCD-contiguous: the neighbour rows of a y stencil and the
// neighbour levels of a z stencil are read by the same XCD shortly before.
struct FaceTiles {
  unsigned ni, nj, nkc;  // tiles along i, row groups, level groups of 4
  __host__ __device__ unsigned count() const { return ni * nj * nkc * 4u; }
};
__host__ __device__ inline FaceTiles face_tiles(const R3 &r) {
  return FaceTiles{(unsigned)(r.i1 - r.i0 + 63) / 63u, (unsigned)(r.j1 - r.j0 + 4) / 4u, (unsigned)(r.k1 - r.k0 + 4) / 4u};
}
// thread -> lane, wave (row in the block), column i, level k, row j and the
// block's first row j0; own: the lane holds a point of r; ic: i clamped into
// r (in-range loads for every lane).  Blocks past the last level return as a
// whole; waves past the last row (jrow false) still take part in the block's
// barrier and form their south face.
#define GET_FACE_BLOCK(r)                                                                \
  const FaceTiles T_ = face_tiles(r);                                                    \
  const unsigned t_ = xcd_tile(blockIdx.x, gridDim.x);                                   \
  const int lane = (int)threadIdx.x, wv = (int)wave_y();                                 \
  const unsigned u_ = t_ / T_.ni;                                                        \
  const int i = (r).i0 + (int)(t_ % T_.ni) * 63 + lane;                                  \
  const int k = (r).k0 + (int)((u_ / 4u) / T_.nj) * 4 + (int)(u_ % 4u);                  \
  const int j0 = (r).j0 + (int)((u_ / 4u) % T_.nj) * 4;                                  \
  const int j = j0 + wv;                                                                 \
  if (k > (r).k1) return; /* whole block */                                              \
  const bool jrow = (j <= (r).j1);                                                       \
  const bool own = (lane < 63 && i <= (r).i1 && jrow);                                   \
  const int ic = i <= (r).i1 ? i : (r).i1;                                               \
  const int jc = jrow ? j : (r).j1

// along-j / along-i stencils of a field at fixed (i,k) / (k,j)
#define QJ(q, i, k, j) A3(q, i, k, j)
#define QI(q, i, k, j) A3(q, i, k, j)

// ---------------- scalar (theta, moist, scalars) ----------------------------
__device__ __forceinline__ float fy_s(const DevCtx *__restrict__ x, const float *q, const float *rv, int i, int k, int j) {
  const int jde = x->jde;
  float vel = A3(rv, i, k, j);
  if (j >= 4 && j <= jde - 3)
    return vel * flux5(A3(q, i, k, j - 3), A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), A3(q, i, k, j + 2), vel);
  if (j == 2 || j == jde - 1) return 0.5f * vel * (A3(q, i, k, j) + A3(q, i, k, j - 1));
  return vel * flux3(A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), vel);
}
__device__ __forceinline__ float fx_s(const DevCtx *__restrict__ x, const float *q, const float *ru, int i, int k, int j) {
  const int ide = x->ide;
  float vel = A3(ru, i, k, j);
  if (i >= 4 && i <= ide - 3)
    return vel * flux5(A3(q, i - 3, k, j), A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), A3(q, i + 2, k, j), vel);
  if (i == 2 || i == ide - 1) return 0.5f * vel * (A3(q, i, k, j) + A3(q, i - 1, k, j));
  return vel * flux3(A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), vel);
}
__device__ __forceinline__ float fz_s(const DevCtx *__restrict__ x, const float *q, const float *rom, int i, int k, int j) {
  const int ktf = x->kde - 1;
  if (k <= 1 || k >= x->kde) return 0.f;
  const float *fzm = FL(fnm), *fzp = FL(fnp);
  float vel = A3(rom, i, k, j);
  if (k == 2 || k == ktf) return vel * (A1(fzm, k) * A3(q, i, k, j) + A1(fzp, k) * A3(q, i, k - 1, j));
  return vel * flux3(A3(q, i, k - 2, j), A3(q, i, k - 1, j), A3(q, i, k, j), A3(q, i, k + 1, j), -vel);
}

// advect_scalar (module_advect_em.py:90-369): tend -= div(flux), mass points
// (face-sharing blocks; r = mass points; y faces j in [2, jde-1]).
// TEND: the theta tendency of the RK step, from zero (the cleared array) plus
// rk_addtend_dry, written once.
template <bool TEND>
__global__ void __launch_bounds__(256) k_advect_scalar(const DevCtx xv, const float *q, float *tend, const float *ru, const float *rv,
                                                       const float *rom, R3 r, int rk) {
  const DevCtx *__restrict__ x = &xv;
  __shared__ float ys[4][64];
  GET_FACE_BLOCK(r);
  const float *msftx = FL(msftx), *rdzw = FL(rdnw);
  const int ide = x->ide, jde = x->jde;
  const float fs = (j >= 2 && j <= jde - 1) ? fy_s(x, q, rv, ic, k, j) : 0.f;
  ys[wv][lane] = fs;
  __syncthreads();
  if (!jrow) return;
  float fn;
  if (wv < 3) fn = ys[wv + 1][lane];
  else fn = (j + 1 <= jde - 1) ? fy_s(x, q, rv, ic, k, j + 1) : 0.f;
  const float fw = (i >= 2 && i <= ide - 1) ? fx_s(x, q, ru, i, k, j) : 0.f;
  const float fe = lane_next(fw, lane);
  if (!own) return;
  float t = TEND ? 0.f : A3(tend, ic, k, j);
  if (j >= 2 && j <= jde - 2) t -= A2(msftx, ic, j) * x->rdy * (fn - fs);
  if (ic >= 2 && ic <= ide - 2) t -= A2(msftx, ic, j) * x->rdx * (fe - fw);
  t -= A1(rdzw, k) * (fz_s(x, q, rom, ic, k + 1, j) - fz_s(x, q, rom, ic, k, j));
  if (TEND) t = addtend_t(x, ic, k, j, rk, t);
  A3(tend, ic, k, j) = t;
}

// ---------------- u momentum (u points i in [2, ide-1]) ---------------------
__device__ __forceinline__ float fy_u(const DevCtx *__restrict__ x, const float *u, const float *rv, int i, int k, int j) {
  const int jde = x->jde;
  if (j >= 4 && j <= jde - 3) {
    float vel = 0.5f * (A3(rv, i, k, j) + A3(rv, i - 1, k, j));
    return vel * flux5(A3(u, i, k, j - 3), A3(u, i, k, j - 2), A3(u, i, k, j - 1), A3(u, i, k, j), A3(u, i, k, j + 1), A3(u, i, k, j + 2), vel);
  }
  if (j == 2 || j == jde - 1) return 0.25f * (A3(rv, i, k, j) + A3(rv, i - 1, k, j)) * (A3(u, i, k, j) + A3(u, i, k, j - 1));
  float vel = 0.5f * (A3(rv, i, k, j) + A3(rv, i - 1, k, j));
  return vel * flux3(A3(u, i, k, j - 2), A3(u, i, k, j - 1), A3(u, i, k, j), A3(u, i, k, j + 1), vel);
}
__device__ __forceinline__ float fx_u(const DevCtx *__restrict__ x, const float *u, const float *ru, int i, int k, int j) {
  const int ide = x->ide;
  if (i >= 4 && i <= ide - 2) {
    float vel = 0.5f * (A3(ru, i, k, j) + A3(ru, i - 1, k, j));
    return vel * flux5(A3(u, i - 3, k, j), A3(u, i - 2, k, j), A3(u, i - 1, k, j), A3(u, i, k, j), A3(u, i + 1, k, j), A3(u, i + 2, k, j), vel);
  }
  if (i == 2) {
    float ub = A3(u, i - 1, k, j);
    if (A3(u, i, k, j) < 0.f) ub = A3(u, i, k, j);
    return 0.25f * (A3(ru, i, k, j) + A3(ru, i - 1, k, j)) * (A3(u, i, k, j) + ub);
  }
  if (i == ide) {
    float ub = A3(u, i, k, j);
    if (A3(u, i - 1, k, j) > 0.f) ub = A3(u, i - 1, k, j);
    return 0.25f * (A3(ru, i, k, j) + A3(ru, i - 1, k, j)) * (A3(u, i - 1, k, j) + ub);
  }
  float vel = 0.5f * (A3(ru, i, k, j) + A3(ru, i - 1, k, j));  // i == 3 or ide-1
  return vel * flux3(A3(u, i - 2, k, j), A3(u, i - 1, k, j), A3(u, i, k, j), A3(u, i + 1, k, j), vel);
}
__device__ __forceinline__ float fz_u(const DevCtx *__restrict__ x, const float *u, const float *rom, int i, int k, int j) {
  const int ktf = x->kde - 1;
  if (k <= 1 || k >= x->kde) return 0.f;
  const float *fzm = FL(fnm), *fzp = FL(fnp);
  if (k == 2 || k == ktf)
    return 0.5f * (A3(rom, i, k, j) + A3(rom, i - 1, k, j)) * (A1(fzm, k) * A3(u, i, k, j) + A1(fzp, k) * A3(u, i, k - 1, j));
  float vel = 0.5f * (A3(rom, i - 1, k, j) + A3(rom, i, k, j));
  return vel * flux3(A3(u, i, k - 2, j), A3(u, i, k - 1, j), A3(u, i, k, j), A3(u, i, k + 1, j), -vel);
}
// advect_u (module_advect_em.py:762-1034); r covers i [2, ide-1], j [1, jde-1]
// (face-sharing blocks: x faces i in [2, ide], y faces j in [2, jde-1]).
// TEND: the u tendency of the RK step over i [1, ide]: from zero (the cleared
// array), advection, horizontal pressure gradient and Coriolis/curvature at
// i [2, ide-1] in the reference's order, then rk_addtend_dry; written once.
template <bool TEND>
__global__ void __launch_bounds__(256) k_advect_u(const DevCtx xv, const float *u, float *tend, const float *ru, const float *rv,
                                                  const float *rom, R3 r, int rk) {
  const DevCtx *__restrict__ x = &xv;
  __shared__ float ys[4][64];
  GET_FACE_BLOCK(r);
  const float *msfux = FL(msfux), *rdzw = FL(rdnw);
  const int ide = x->ide, jde = x->jde;
  const bool adv = !TEND || (ic >= 2 && ic <= ide - 1);
  const float fs = (adv && j >= 2 && j <= jde - 1) ? fy_u(x, u, rv, ic, k, j) : 0.f;
  ys[wv][lane] = fs;
  __syncthreads();
  if (!jrow) return;
  float fn;
  if (wv < 3) fn = ys[wv + 1][lane];
  else fn = (adv && j + 1 <= jde - 1) ? fy_u(x, u, rv, ic, k, j + 1) : 0.f;
  const float fw = (i >= 2 && i <= ide) ? fx_u(x, u, ru, i, k, j) : 0.f;
  const float fe = lane_next(fw, lane);
  if (!own) return;
  float t = TEND ? 0.f : A3(tend, ic, k, j);
  if (adv) {
    if (j >= 2 && j <= jde - 2) t -= A2(msfux, ic, j) * x->rdy * (fn - fs);
    t -= A2(msfux, ic, j) * x->rdx * (fe - fw);
    t -= A1(rdzw, k) * (fz_u(x, u, rom, ic, k + 1, j) - fz_u(x, u, rom, ic, k, j));
    if (TEND) {
      t = t - hpg_term(x, 0, ic, k, j);
      t = cor_curv_u(x, ic, k, j, t);
    }
  }
  if (TEND) t = addtend_u(x, ic, k, j, rk, t);
  A3(tend, ic, k, j) = t;
}
// ---------------- v momentum (v rows j in [2, jde-1]) -----------------------
__device__ __forceinline__ float fy_v(const DevCtx *__restrict__ x, const float *v, const float *rv, int i, int k, int j) {
  const int jde = x->jde;
  if (j >= 4 && j <= jde - 2) {
    float vel = 0.5f * (A3(rv, i, k, j) + A3(rv, i, k, j - 1));
    return vel * flux5(A3(v, i, k, j - 3), A3(v, i, k, j - 2), A3(v, i, k, j - 1), A3(v, i, k, j), A3(v, i, k, j + 1), A3(v, i, k, j + 2), vel);
  }
  if (j == 2) {
    float vb = A3(v, i, k, j - 1);
    if (A3(v, i, k, j) < 0.f) vb = A3(v, i, k, j);
    return 0.25f * (A3(rv, i, k, j) + A3(rv, i, k, j - 1)) * (A3(v, i, k, j) + vb);
  }
  if (j == jde) {
    float vb = A3(v, i, k, j);
    if (A3(v, i, k, j - 1) > 0.f) vb = A3(v, i, k, j - 1);
    return 0.25f * (A3(rv, i, k, j) + A3(rv, i, k, j - 1)) * (vb + A3(v, i, k, j - 1));
  }
  float vel = 0.5f * (A3(rv, i, k, j) + A3(rv, i, k, j - 1));  // j == 3 or jde-1
  return vel * flux3(A3(v, i, k, j - 2), A3(v, i, k, j - 1), A3(v, i, k, j), A3(v, i, k, j + 1), vel);
}
__device__ __forceinline__ float fx_v(const DevCtx *__restrict__ x, const float *v, const float *ru, int i, int k, int j) {
  const int ide = x->ide;
  if (i >= 4 && i <= ide - 3) {
    float vel = 0.5f * (A3(ru, i, k, j) + A3(ru, i, k, j - 1));
    return vel * flux5(A3(v, i - 3, k, j), A3(v, i - 2, k, j), A3(v, i - 1, k, j), A3(v, i, k, j), A3(v, i + 1, k, j), A3(v, i + 2, k, j), vel);
  }
  if (i == 2 || i == ide - 1) return 0.25f * (A3(ru, i, k, j) + A3(ru, i, k, j - 1)) * (A3(v, i, k, j) + A3(v, i - 1, k, j));
  float vel = 0.5f * (A3(ru, i, k, j) + A3(ru, i, k, j - 1));  // i == 3 or ide-2
  return vel * flux3(A3(v, i - 2, k, j), A3(v, i - 1, k, j), A3(v, i, k, j), A3(v, i + 1, k, j), vel);
}
__device__ __forceinline__ float fz_v(const DevCtx *__restrict__ x, const float *v, const float *rom, int i, int k, int j) {
  const int ktf = x->kde - 1;
  if (k <= 1 || k >= x->kde) return 0.f;
  const float *fzm = FL(fnm), *fzp = FL(fnp);
  if (k == 2 || k == ktf)
    return 0.5f * (A3(rom, i, k, j) + A3(rom, i, k, j - 1)) * (A1(fzm, k) * A3(v, i, k, j) + A1(fzp, k) * A3(v, i, k - 1, j));
  float vel = 0.5f * (A3(rom, i, k, j) + A3(rom, i, k, j - 1));
  return vel * flux3(A3(v, i, k - 2, j), A3(v, i, k - 1, j), A3(v, i, k, j), A3(v, i, k + 1, j), -vel);
}
// advect_v (module_advect_em.py:1037-1354); r covers i [1, ide-1], j [2, jde-1]
// (face-sharing blocks: x faces i in [2, ide-1], y faces j in [2, jde]).
// TEND: the v tendency of the RK step over j [1, jde] (advection, pressure
// gradient, Coriolis/curvature at j [2, jde-1], then rk_addtend_dry).
template <bool TEND>
__global__ void __launch_bounds__(256) k_advect_v(const DevCtx xv, const float *v, float *tend, const float *ru, const float *rv,
                                                  const float *rom, R3 r, int rk) {
  const DevCtx *__restrict__ x = &xv;
  __shared__ float ys[4][64];
  GET_FACE_BLOCK(r);
  const float *msfvy = FL(msfvy), *msfvx = FL(msfvx), *rdzw = FL(rdnw);
  const int ide = x->ide, jde = x->jde;
  const bool adv = !TEND || (j >= 2 && j <= jde - 1);  // whole wave
  const float fs = (j >= 2 && j <= jde) ? fy_v(x, v, rv, ic, k, j) : 0.f;
  ys[wv][lane] = fs;
  __syncthreads();
  if (!jrow) return;
  float fn;
  if (wv < 3) fn = ys[wv + 1][lane];
  else fn = (j + 1 >= 2 && j + 1 <= jde) ? fy_v(x, v, rv, ic, k, j + 1) : 0.f;
  const float fw = (adv && i >= 2 && i <= ide - 1) ? fx_v(x, v, ru, i, k, j) : 0.f;
  const float fe = lane_next(fw, lane);
  if (!own) return;
  float t = TEND ? 0.f : A3(tend, ic, k, j);
  if (adv) {
    t -= A2(msfvy, ic, j) * x->rdy * (fn - fs);
    if (ic >= 2 && ic <= ide - 2) t -= A2(msfvy, ic, j) * x->rdx * (fe - fw);
    t -= (A2(msfvy, ic, j) / A2(msfvx, ic, j)) * A1(rdzw, k) * (fz_v(x, v, rom, ic, k + 1, j) - fz_v(x, v, rom, ic, k, j));
    if (TEND) {
      t = t - hpg_term(x, 1, ic, k, j);
      t = cor_curv_v(x, ic, k, j, t);
    }
  }
  if (TEND) t = addtend_v(x, ic, k, j, rk, t);
  A3(tend, ic, k, j) = t;
}
// ---------------- w momentum (levels k in [2, kde]) -------------------------
__device__ __forceinline__ float rvw(const DevCtx *__restrict__ x, const float *rv, int i, int k, int j) {
  const float *fzm = FL(fnm), *fzp = FL(fnp);
  if (k == x->kde) return (2.f - A1(fzm, k - 1)) * A3(rv, i, k - 1, j) - A1(fzp, k - 1) * A3(rv, i, k - 2, j);
  return A1(fzm, k) * A3(rv, i, k, j) + A1(fzp, k) * A3(rv, i, k - 1, j);
}
__device__ __forceinline__ float fy_w(const DevCtx *__restrict__ x, const float *w, const float *rv, int i, int k, int j) {
  const int jde = x->jde;
  float vel = rvw(x, rv, i, k, j);
  if (j >= 4 && j <= jde - 3)
    return vel * flux5(A3(w, i, k, j - 3), A3(w, i, k, j - 2), A3(w, i, k, j - 1), A3(w, i, k, j), A3(w, i, k, j + 1), A3(w, i, k, j + 2), vel);
  if (j == 2 || j == jde - 1) return 0.5f * vel * (A3(w, i, k, j) + A3(w, i, k, j - 1));
  return vel * flux3(A3(w, i, k, j - 2), A3(w, i, k, j - 1), A3(w, i, k, j), A3(w, i, k, j + 1), vel);
}
__device__ __forceinline__ float fx_w(const DevCtx *__restrict__ x, const float *w, const float *ru, int i, int k, int j) {
  const int ide = x->ide;
  float vel = rvw(x, ru, i, k, j);
  if (i >= 4 && i <= ide - 3)
    return vel * flux5(A3(w, i - 3, k, j), A3(w, i - 2, k, j), A3(w, i - 1, k, j), A3(w, i, k, j), A3(w, i + 1, k, j), A3(w, i + 2, k, j), vel);
  if (i == 2 || i == ide - 1) return 0.5f * vel * (A3(w, i, k, j) + A3(w, i - 1, k, j));
  return vel * flux3(A3(w, i - 2, k, j), A3(w, i - 1, k, j), A3(w, i, k, j), A3(w, i + 1, k, j), vel);
}
__device__ __forceinline__ float fz_w(const DevCtx *__restrict__ x, const float *w, const float *rom, int i, int k, int j) {
  if (k <= 1 || k > x->kde) return 0.f;
  if (k == 2 || k == x->kde) return 0.25f * (A3(rom, i, k, j) + A3(rom, i, k - 1, j)) * (A3(w, i, k, j) + A3(w, i, k - 1, j));
  float vel = 0.5f * (A3(rom, i, k, j) + A3(rom, i, k - 1, j));
  return vel * flux3(A3(w, i, k - 2, j), A3(w, i, k - 1, j), A3(w, i, k, j), A3(w, i, k + 1, j), -vel);
}
// advect_w (module_advect_em.py:1357-1640); r covers i [1, ide-1], k [2, kde], j [1, jde-1]
// (face-sharing blocks: x faces i in [2, ide-1], y faces j in [2, jde-1]).
// TEND: the w tendency of the RK step over k [1, kde]: advection, buoyancy
// and pressure (pg_buoy_w, reading cqw before k_cqw_inv replaces it) and, below
// the top, Coriolis/curvature at k [2, kde]; then rk_addtend_dry.
template <bool TEND>
__global__ void __launch_bounds__(256) k_advect_w(const DevCtx xv, const float *w, float *tend, const float *ru, const float *rv,
                                                  const float *rom, R3 r, int rk) {
  const DevCtx *__restrict__ x = &xv;
  __shared__ float ys[4][64];
  GET_FACE_BLOCK(r);
  const float *msftx = FL(msftx), *rdzu = FL(rdn);
  const int ide = x->ide, jde = x->jde;
  const bool adv = !TEND || k >= 2;  // whole block
  const float fs = (adv && j >= 2 && j <= jde - 1) ? fy_w(x, w, rv, ic, k, j) : 0.f;
  ys[wv][lane] = fs;
  __syncthreads();
  if (!jrow) return;
  float fn;
  if (wv < 3) fn = ys[wv + 1][lane];
  else fn = (adv && j + 1 <= jde - 1) ? fy_w(x, w, rv, ic, k, j + 1) : 0.f;
  const float fw = (adv && i >= 2 && i <= ide - 1) ? fx_w(x, w, ru, i, k, j) : 0.f;
  const float fe = lane_next(fw, lane);
  if (!own) return;
  float t = TEND ? 0.f : A3(tend, ic, k, j);
  if (adv) {
    if (j >= 2 && j <= jde - 2) t -= A2(msftx, ic, j) * x->rdy * (fn - fs);
    if (ic >= 2 && ic <= ide - 2) t -= A2(msftx, ic, j) * x->rdx * (fe - fw);
    if (k <= x->kde - 1) t -= A1(rdzu, k) * (fz_w(x, w, rom, ic, k + 1, j) - fz_w(x, w, rom, ic, k, j));
    else t += 2.f * A1(rdzu, k - 1) * fz_w(x, w, rom, ic, k, j);
    if (TEND) {
      t = buoy_w(x, ic, k, j, t);
      if (k <= x->kde - 1) t = cor_curv_w(x, ic, k, j, t);
    }
  }
  if (TEND) t = addtend_w(x, ic, k, j, rk, t);
  A3(tend, ic, k, j) = t;
}

// the phi tendency of the RK step over k [1, kde]: rhs_ph at k [2, kde] from
// zero (the cleared array), then rk_addtend_dry; r = i [1, ide-1], k [1, kde],
// j [1, jde-1]
__global__ void __launch_bounds__(256) k_tend_ph(const DevCtx xv, R3 r, int rk) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  float t = 0.f;
  if (k >= 2) t = rhs_ph_val(x, i, k, j, t);
  t = addtend_ph(x, i, k, j, rk, t);
  A3(FL(ph_tend), i, k, j) = t;
}
// ---------------- positive-definite scalar advection ------------------------
// pass 1: high-order minus upwind (antidiffusive) and upwind fluxes
// (module_advect_em.py:537-660) on all faces the limiter and tendency read.
__global__ void k_pd_fluxes(const DevCtx xv, const float *q, const float *q_old, const float *ru, const float *rv,
                            const float *rom, const float *mut, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *c1 = FL(c1h), *c2 = FL(c2h), *msftx = FL(msftx), *msfty = FL(msfty), *rdzw = FL(rdnw);
  const float *fzm = FL(fnm), *fzp = FL(fnp);
  float *fqx = SCR(S_FQX), *fqy = SCR(S_FQY), *fqz = SCR(S_FQZ), *fqxl = SCR(S_FQXL), *fqyl = SCR(S_FQYL), *fqzl = SCR(S_FQZL);
  const int ide = x->ide, jde = x->jde, ktf = x->kde - 1;
  const float rdx = x->rdx, rdy = x->rdy;
  // y face j (rows 2..jde-1)
  if (j >= 2 && j <= jde - 1 && i <= ide - 1 && k <= ktf) {
    float dy = 2.f / (A2(msftx, i, j) + A2(msftx, i, j - 1)) / rdy;
    float mu = 0.5f * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i, j - 1) + A1(c2, k)));
    float vel = A3(rv, i, k, j);
    float cr = vel * dt / dy / mu;
    float low = mu * (dy / dt) * flux_up(A3(q_old, i, k, j - 1), A3(q_old, i, k, j), cr);
    float high;
    if (j >= 4 && j <= jde - 3)
      high = vel * flux5pd(A3(q, i, k, j - 3), A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), A3(q, i, k, j + 2), vel);
    else if (j == 2 || j == jde - 1)
      high = 0.5f * vel * (A3(q, i, k, j) + A3(q, i, k, j - 1));
    else
      high = vel * flux3pd(A3(q, i, k, j - 2), A3(q, i, k, j - 1), A3(q, i, k, j), A3(q, i, k, j + 1), vel);
    A3(fqyl, i, k, j) = low;
    A3(fqy, i, k, j) = high - low;
  } else {
    A3(fqyl, i, k, j) = 0.f;
    A3(fqy, i, k, j) = 0.f;
  }
  // x face i (2..ide-1)
  if (i >= 2 && i <= ide - 1 && j <= jde - 1 && k <= ktf) {
    float dx = 2.f / (A2(msfty, i, j) + A2(msfty, i - 1, j)) / rdx;
    float mu = 0.5f * ((A1(c1, k) * A2(mut, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mut, i - 1, j) + A1(c2, k)));
    float vel = A3(ru, i, k, j);
    float cr = vel * dt / dx / mu;
    float low = mu * (dx / dt) * flux_up(A3(q_old, i - 1, k, j), A3(q_old, i, k, j), cr);
    float high;
    if (i >= 4 && i <= ide - 3)
      high = vel * flux5pd(A3(q, i - 3, k, j), A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), A3(q, i + 2, k, j), vel);
    else if (i == 2 || i == ide - 1)
      high = 0.5f * vel * (A3(q, i, k, j) + A3(q, i - 1, k, j));
    else
      high = vel * flux3pd(A3(q, i - 2, k, j), A3(q, i - 1, k, j), A3(q, i, k, j), A3(q, i + 1, k, j), vel);
    A3(fqxl, i, k, j) = low;
    A3(fqx, i, k, j) = high - low;
  } else {
    A3(fqxl, i, k, j) = 0.f;
    A3(fqx, i, k, j) = 0.f;
  }
  // z face k (2..ktf); zero at k = 1 and k = kde
  if (k >= 2 && k <= ktf && i <= ide - 1 && j <= jde - 1) {
    float dz = 2.f / (A1(rdzw, k) + A1(rdzw, k - 1));
    float mu = A1(c1, k) * A2(mut, i, j) + A1(c2, k);
    float vel = A3(rom, i, k, j);
    float cr = vel * dt / dz / mu;
    float low = mu * (dz / dt) * flux_up(A3(q_old, i, k - 1, j), A3(q_old, i, k, j), cr);
    float high;
    if (k == 2 || k == ktf) high = vel * (A1(fzm, k) * A3(q, i, k, j) + A1(fzp, k) * A3(q, i, k - 1, j));
    else high = vel * flux3pd(A3(q, i, k - 2, j), A3(q, i, k - 1, j), A3(q, i, k, j), A3(q, i, k + 1, j), -vel);
    A3(fqzl, i, k, j) = low;
    A3(fqz, i, k, j) = high - low;
  } else {
    A3(fqzl, i, k, j) = 0.f;
    A3(fqz, i, k, j) = 0.f;
  }
}
// pass 2: renormalisation factor of each limited cell (1 = not limited)
// (module_advect_em.py:662-720); cells i in [2, ide-2], j in [2, jde-2].
__global__ void k_pd_scale(const DevCtx xv, const float *q_old, const float *mub, const float *mu_old, float dt, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *c1 = FL(c1h), *c2 = FL(c2h), *msftx = FL(msftx), *msfty = FL(msfty), *rdzw = FL(rdnw);
  const float *fqx = SCR(S_FQX), *fqy = SCR(S_FQY), *fqz = SCR(S_FQZ), *fqxl = SCR(S_FQXL), *fqyl = SCR(S_FQYL), *fqzl = SCR(S_FQZL);
  float *sc = SCR(S_SCALE);
  float scale = 1.f;
  if (i >= 2 && i <= x->ide - 2 && j >= 2 && j <= x->jde - 2 && k <= x->kde - 1) {
    const float rdx = x->rdx, rdy = x->rdy;
    float mm = A2(msftx, i, j) * A2(msfty, i, j);
    float ph_low = ((A1(c1, k) * A2(mub, i, j) + A1(c2, k)) + (A1(c1, k) * A2(mu_old, i, j))) * A3(q_old, i, k, j) -
                   dt * (mm * (rdx * (A3(fqxl, i + 1, k, j) - A3(fqxl, i, k, j)) + rdy * (A3(fqyl, i, k, j + 1) - A3(fqyl, i, k, j))) +
                         A2(msfty, i, j) * A1(rdzw, k) * (A3(fqzl, i, k + 1, j) - A3(fqzl, i, k, j)));
    float flux_out = dt * (mm * (rdx * (fmaxf(0.f, A3(fqx, i + 1, k, j)) - fminf(0.f, A3(fqx, i, k, j))) +
                                 rdy * (fmaxf(0.f, A3(fqy, i, k, j + 1)) - fminf(0.f, A3(fqy, i, k, j)))) +
                           A2(msfty, i, j) * A1(rdzw, k) * (fminf(0.f, A3(fqz, i, k + 1, j)) - fmaxf(0.f, A3(fqz, i, k, j))));
    if (flux_out > ph_low) scale = fmaxf(0.f, ph_low / (flux_out + 1.0e-20f));
  }
  A3(sc, i, k, j) = scale;
}
// pass 3: tendency from the renormalised faces; a face is scaled by the
// cell its antidiffusive flux leaves (the only cell that may limit it).
__global__ void k_pd_tend(const DevCtx xv, float *tend, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const float *msftx = FL(msftx), *rdzw = FL(rdnw);
  const float *fqx = SCR(S_FQX), *fqy = SCR(S_FQY), *fqz = SCR(S_FQZ), *fqxl = SCR(S_FQXL), *fqyl = SCR(S_FQYL), *fqzl = SCR(S_FQZL);
  const float *sc = SCR(S_SCALE);
  const float rdx = x->rdx, rdy = x->rdy;
  // face fluxes after limiting
  float zt = A3(fqz, i, k + 1, j), zb = A3(fqz, i, k, j);
  if (zt < 0.f) zt *= A3(sc, i, k, j);          // outflow upward from this cell
  else if (zt > 0.f && k + 1 <= x->kde - 1) zt *= A3(sc, i, k + 1, j);
  if (zb > 0.f) zb *= A3(sc, i, k, j);
  else if (zb < 0.f && k - 1 >= 1) zb *= A3(sc, i, k - 1, j);
  float t = A3(tend, i, k, j);
  t += -A1(rdzw, k) * (zt - zb + A3(fqzl, i, k + 1, j) - A3(fqzl, i, k, j));
  if (i >= 2 && i <= x->ide - 2) {
    float xe = A3(fqx, i + 1, k, j), xw = A3(fqx, i, k, j);
    if (xe > 0.f) xe *= A3(sc, i, k, j);
    else if (xe < 0.f) xe *= A3(sc, i + 1, k, j);
    if (xw < 0.f) xw *= A3(sc, i, k, j);
    else if (xw > 0.f) xw *= A3(sc, i - 1, k, j);
    t += -A2(msftx, i, j) * (rdx * (xe - xw + A3(fqxl, i + 1, k, j) - A3(fqxl, i, k, j)));
  }
  if (j >= 2 && j <= x->jde - 2) {
    float yn = A3(fqy, i, k, j + 1), ys = A3(fqy, i, k, j);
    if (yn > 0.f) yn *= A3(sc, i, k, j);
    else if (yn < 0.f) yn *= A3(sc, i, k, j + 1);
    if (ys < 0.f) ys *= A3(sc, i, k, j);
    else if (ys > 0.f) ys *= A3(sc, i, k, j - 1);
    t += -A2(msftx, i, j) * (rdy * (yn - ys + A3(fqyl, i, k, j + 1) - A3(fqyl, i, k, j)));
  }
  A3(tend, i, k, j) = t;
}
