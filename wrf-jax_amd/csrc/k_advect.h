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
// Tile order: i fastest, then LG consecutive levels, then rows, then groups
// of LG levels, XCD-contiguous: the neighbour rows of a y stencil and the
// neighbour levels of a z stencil are read by the same XCD shortly before.
// LG is a launch setting of the domain (DevCtx::face_lg, wrfdc_set_tile):
// by default 4 for most kernels and 8 for the PD species passes and
// advect_w, which measured fastest at 400x400x60 (each kernel timed with LG
// 2, 4, 8 and 16); the ensemble's autotuner tries others per layout.
#ifndef WRFDC_FACE_LG
#define WRFDC_FACE_LG 4
#endif
#ifndef WRFDC_FACE_ROWS
#define WRFDC_FACE_ROWS 4
#endif
constexpr unsigned kFaceLG = WRFDC_FACE_LG;  // default levels per level group of the tile order
// rows (waves) per face block: 4 measured fastest for every face-block kernel
// at 400x400x60 (2 rows: step +8 %; 8 rows: step +2 %, k_species_pd_update +11 %)
constexpr int kFaceRows = WRFDC_FACE_ROWS;
constexpr unsigned kFaceLGPd = 8;            // default for k_species_pd_scale, k_species_pd_update
constexpr unsigned kFaceLGW = 8;             // default for k_advect_w
struct FaceTiles {
  unsigned ni, nj, nkc, lg;  // tiles along i, row groups, level groups, levels per group
  __host__ __device__ unsigned count() const { return ni * nj * nkc * lg; }
};
__host__ __device__ inline FaceTiles face_tiles(const R3 &r, unsigned lg = kFaceLG) {
  return FaceTiles{(unsigned)(r.i1 - r.i0 + 63) / 63u, (unsigned)(r.j1 - r.j0 + kFaceRows) / (unsigned)kFaceRows,
                   (unsigned)(r.k1 - r.k0 + (int)lg) / lg, lg};
}
// thread -> lane, wave (row in the block), column i, level k, row j and the
// block's first row j0; own: the lane holds a point of r; ic: i clamped into
// r (in-range loads for every lane).  Blocks past the last level return as a
// whole; waves past the last row (jrow false) still take part in the block's
// barrier and form their south face.  The launch's grid is
// face_tiles(r, LG).count() blocks with the kernel's LG.
#define GET_FACE_BLOCK_LG(r, LG)                                                         \
  const FaceTiles T_ = face_tiles(r, LG);                                                \
  const unsigned t_ = xcd_tile(blockIdx.x, gridDim.x);                                   \
  const int lane = (int)threadIdx.x, wv = (int)wave_y();                                 \
  const unsigned u_ = t_ / T_.ni;                                                        \
  const int i = (r).i0 + (int)(t_ % T_.ni) * 63 + lane;                                  \
  const int k = (r).k0 + (int)((u_ / (LG)) / T_.nj) * (int)(LG) + (int)(u_ % (LG));      \
  const int j0 = (r).j0 + (int)((u_ / (LG)) % T_.nj) * kFaceRows;                        \
  const int j = j0 + wv;                                                                 \
  if (k > (r).k1) return; /* whole block */                                              \
  const bool jrow = (j <= (r).j1);                                                       \
  const bool own = (lane < 63 && i <= (r).i1 && jrow);                                   \
  const int ic = i <= (r).i1 ? i : (r).i1
#define GET_FACE_BLOCK(r) GET_FACE_BLOCK_LG(r, (unsigned)x->face_lg[0])

// ---------------- face fluxes from loaded stencils ---------------------------
// Every kernel below issues all loads of a point first (buffer loads,
// unclamped: the slots a boundary-degraded order does not use may fall
// outside the field, where they read 0), then evaluates every order of a
// face and selects the face's own, so no branch separates the loads from
// their use.  Each selection keeps the arithmetic of the reference's
// expression for that order.

// the 6 values around the face at element offset of along stride st
// (slots -3 .. +2: w[3] is the cell on the face's high side)
__device__ __forceinline__ void win6b(const BufF &b, unsigned of, unsigned st, float w[6]) {
#pragma unroll
  for (int n = 0; n < 6; n++) w[n] = bld(b, of + (unsigned)(n - 3) * st);
}
// the column window of levels k-2 .. k+2 at element offset o
__device__ __forceinline__ void win5z(const BufF &b, unsigned o, unsigned sk, float w[5]) {
#pragma unroll
  for (int n = 0; n < 5; n++) w[n] = bld(b, o + (unsigned)(n - 2) * sk);
}
// Holds loaded values in registers at this point (an empty asm statement
// that reads them): their loads complete here, so the compiler cannot sink
// them into the flux-order branch that uses them.  Placed after the last
// load of a group (the statement also keeps later loads after it).
__device__ __forceinline__ void pin6(const float w[6]) { asm volatile("" : : "v"(w[0]), "v"(w[1]), "v"(w[2]), "v"(w[3]), "v"(w[4]), "v"(w[5])); }
__device__ __forceinline__ void pin5(const float w[5]) { asm volatile("" : : "v"(w[0]), "v"(w[1]), "v"(w[2]), "v"(w[3]), "v"(w[4])); }
// Two ways to pick a face's flux order once its stencil is in registers:
// *_sel evaluates the 5th-order flux and, unless `fast` (a wave-uniform
// flag: every lane's face takes the 5th order) says it is the face's own for
// the whole wave, every other order and selects (no divergent branch:
// kernels with many loads in flight keep their registers); *_br branches on
// the order (only the face's own order is evaluated: the species and scalar
// kernels, whose loads are pinned first).  The order of a face in j or k is
// uniform over a wave; in i it differs only in the waves at the domain edges.
__device__ __forceinline__ float hflux_br(const float w[6], float vel, int f, int fe) {
  if (f >= 4 && f <= fe - 3) return vel * flux5(w[0], w[1], w[2], w[3], w[4], w[5], vel);
  if (f == 2 || f == fe - 1) return 0.5f * vel * (w[3] + w[2]);
  return vel * flux3(w[1], w[2], w[3], w[4], vel);
}
__device__ __forceinline__ float zflux_br(float qm2, float qm1, float q0, float qp1, float vel, int kk, int kde, float fm, float fp) {
  if (kk <= 1 || kk >= kde) return 0.f;
  if (kk == 2 || kk == kde - 1) return vel * (fm * q0 + fp * qm1);
  return vel * flux3(qm2, qm1, q0, qp1, -vel);
}
// scalar and w faces along x / y (fx_s, fy_s, fx_w, fy_w): face velocity
// vel; 5th order on [4, fe-3], centred at 2 and fe-1, 3rd order between
__device__ __forceinline__ float hflux_sel(const float w[6], float vel, int f, int fe, bool fast) {
  const float h5 = vel * flux5(w[0], w[1], w[2], w[3], w[4], w[5], vel);
  if (fast) return h5;
  const float h2 = 0.5f * vel * (w[3] + w[2]);
  const float h3 = vel * flux3(w[1], w[2], w[3], w[4], vel);
  return (f >= 4 && f <= fe - 3) ? h5 : ((f == 2 || f == fe - 1) ? h2 : h3);
}
// fy_u: rvs = rv(i) + rv(i-1) at the face
__device__ __forceinline__ float fy_u_sel(const float w[6], float rvs, int j, int jde, bool fast) {
  const float vel = 0.5f * rvs;
  const float h5 = vel * flux5(w[0], w[1], w[2], w[3], w[4], w[5], vel);
  if (fast) return h5;
  const float h2 = 0.25f * rvs * (w[3] + w[2]);
  const float h3 = vel * flux3(w[1], w[2], w[3], w[4], vel);
  return (j >= 4 && j <= jde - 3) ? h5 : ((j == 2 || j == jde - 1) ? h2 : h3);
}
// fx_u: rus = ru(i) + ru(i-1); 5th order on [4, ide-2], upwind-biased
// second order at the boundary faces 2 and ide
__device__ __forceinline__ float fx_u_sel(const float w[6], float rus, int i, int ide, bool fast) {
  const float vel = 0.5f * rus;
  const float h5 = vel * flux5(w[0], w[1], w[2], w[3], w[4], w[5], vel);
  if (fast) return h5;
  const float h3 = vel * flux3(w[1], w[2], w[3], w[4], vel);
  const float ub_lo = w[3] < 0.f ? w[3] : w[2];
  const float h_lo = 0.25f * rus * (w[3] + ub_lo);
  const float ub_hi = w[2] > 0.f ? w[2] : w[3];
  const float h_hi = 0.25f * rus * (w[2] + ub_hi);
  return (i >= 4 && i <= ide - 2) ? h5 : (i == 2 ? h_lo : (i == ide ? h_hi : h3));
}
// fy_v: rvs = rv(j) + rv(j-1); 5th order on [4, jde-2], upwind-biased second
// order at the boundary faces 2 and jde
__device__ __forceinline__ float fy_v_sel(const float w[6], float rvs, int j, int jde, bool fast) {
  const float vel = 0.5f * rvs;
  const float h5 = vel * flux5(w[0], w[1], w[2], w[3], w[4], w[5], vel);
  if (fast) return h5;
  const float h3 = vel * flux3(w[1], w[2], w[3], w[4], vel);
  const float vb_lo = w[3] < 0.f ? w[3] : w[2];
  const float h_lo = 0.25f * rvs * (w[3] + vb_lo);
  const float vb_hi = w[2] > 0.f ? w[2] : w[3];
  const float h_hi = 0.25f * rvs * (vb_hi + w[2]);
  return (j >= 4 && j <= jde - 2) ? h5 : (j == 2 ? h_lo : (j == jde ? h_hi : h3));
}
// fx_v: rus = ru(j) + ru(j-1); 5th order on [4, ide-3], centred at 2 and ide-1
__device__ __forceinline__ float fx_v_sel(const float w[6], float rus, int i, int ide, bool fast) {
  const float vel = 0.5f * rus;
  const float h5 = vel * flux5(w[0], w[1], w[2], w[3], w[4], w[5], vel);
  if (fast) return h5;
  const float h2 = 0.25f * rus * (w[3] + w[2]);
  const float h3 = vel * flux3(w[1], w[2], w[3], w[4], vel);
  return (i >= 4 && i <= ide - 3) ? h5 : ((i == 2 || i == ide - 1) ? h2 : h3);
}
// vertical faces of mass-level fields (fz_s, fz_u, fz_v) at level kk from
// q(kk-2 .. kk+1), the face velocity and the weights fnm(kk), fnp(kk): none
// at kk <= 1 and kk >= kde, weighted centred at 2 and kde-1, 3rd order between
__device__ __forceinline__ float zflux_sel(float qm2, float qm1, float q0, float qp1, float vel, int kk, int kde, float fm, float fp) {
  const float h2 = vel * (fm * q0 + fp * qm1);
  const float h3 = vel * flux3(qm2, qm1, q0, qp1, -vel);
  return (kk <= 1 || kk >= kde) ? 0.f : ((kk == 2 || kk == kde - 1) ? h2 : h3);
}
// fz_w at w level kk: roms = rom(kk) + rom(kk-1); none at kk <= 1 and
// kk > kde, centred at 2 and kde, 3rd order between
__device__ __forceinline__ float zflux_w_sel(float qm2, float qm1, float q0, float qp1, float roms, int kk, int kde) {
  const float h2 = 0.25f * roms * (q0 + qm1);
  const float vel = 0.5f * roms;
  const float h3 = vel * flux3(qm2, qm1, q0, qp1, -vel);
  return (kk <= 1 || kk > kde) ? 0.f : ((kk == 2 || kk == kde) ? h2 : h3);
}
// mass flux at a w level from the values at levels k, k-1, k-2 of the column
// (extrapolated at the top level kde)
__device__ __forceinline__ float rvw_sel(const DevCtx *__restrict__ x, float r0, float rm1, float rm2, int k) {
  const int kde = x->kde, kc = imin_(k, kde - 1);
  const float top = (2.f - A1(FL(fnm), kde - 1)) * rm1 - A1(FL(fnp), kde - 1) * rm2;
  const float mid = A1(FL(fnm), kc) * r0 + A1(FL(fnp), kc) * rm1;
  return k == kde ? top : mid;
}

// ---------------- scalar (theta, moist, scalars) ----------------------------
// advect_scalar (module_advect_em.py:90-369): tend -= div(flux), mass points
// (face-sharing blocks; r = mass points; y faces j in [2, jde-1]).
// TEND: the theta tendency of the RK step, from zero (the cleared array) plus
// rk_addtend_dry, written once.
template <bool TEND>
__global__ void __launch_bounds__(64 * kFaceRows) k_advect_scalar(const DevCtx xv, const float *q, float *tend, const float *ru, const float *rv,
                                                       const float *rom, R3 r, int rk) {
  const DevCtx *__restrict__ x = &xv;
  __shared__ float ys[kFaceRows][64];
  GET_FACE_BLOCK(r);
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const unsigned nb = (unsigned)x->n3, sj = (unsigned)(kde * x->ip), sk = (unsigned)x->ip;
  const BufF bq = bufv(q, nb), bv = bufv(rv, nb);
  const unsigned o = XI3(ic, k, j);
  float fs;
  {
    float w[6];
    win6b(bq, o, sj, w);
    const float vs = bld(bv, o);
    pin6(w);
    const float f = hflux_br(w, vs, j, jde);
    fs = (j >= 2 && j <= jde - 1) ? f : 0.f;
  }
  ys[wv][lane] = fs;
  __syncthreads();
  if (!jrow) return;
  // every load of the point: x stencil of the west face, column window,
  // north face stencil (top wave)
  const BufF bu = bufv(ru, nb), bw = bufv(rom, nb);
  const unsigned ox = XI3(i, k, j);
  float wx[6], wz[5], wn[6];
  win6b(bq, ox, 1u, wx);
  win5z(bq, o, sk, wz);
  const float uw = bld(bu, ox), wt = bld(bw, o + sk), wb = bld(bw, o);
  float vn = 0.f;
  if (wv == kFaceRows - 1) {
    win6b(bq, o + sj, sj, wn);
    vn = bld(bv, o + sj);
  }
  const int k1 = imin_(k + 1, kde);
  const float fm0 = A1(FL(fnm), k), fp0 = A1(FL(fnp), k), fm1 = A1(FL(fnm), k1), fp1 = A1(FL(fnp), k1);
  const float msf = A2(FL(msftx), ic, j), rdz = A1(FL(rdnw), k);
  const float t0 = TEND ? 0.f : elem(tend, o);
  pin6(wx);
  pin5(wz);
  if (wv == kFaceRows - 1) pin6(wn);
  const float fw0 = hflux_br(wx, uw, i, ide);
  const float fw = (i >= 2 && i <= ide - 1) ? fw0 : 0.f;
  const float fe = lane_next(fw, lane);
  if (!own) return;
  float fn;
  if (wv < kFaceRows - 1) fn = ys[wv + 1][lane];
  else fn = (j + 1 <= jde - 1) ? hflux_br(wn, vn, j + 1, jde) : 0.f;
  float t = t0;
  if (j >= 2 && j <= jde - 2) t -= msf * x->rdy * (fn - fs);
  if (ic >= 2 && ic <= ide - 2) t -= msf * x->rdx * (fe - fw);
  t -= rdz * (zflux_br(wz[1], wz[2], wz[3], wz[4], wt, k + 1, kde, fm1, fp1) - zflux_br(wz[0], wz[1], wz[2], wz[3], wb, k, kde, fm0, fp0));
  if (TEND) t = addtend_t(x, ic, k, j, rk, t, rk > 1);
  elem(tend, o) = t;
}

// ---------------- u momentum (u points i in [2, ide-1]) ---------------------
// advect_u (module_advect_em.py:762-1034); r covers i [2, ide-1], j [1, jde-1]
// (face-sharing blocks: x faces i in [2, ide], y faces j in [2, jde-1]).
// TEND: the u tendency of the RK step over i [1, ide]: from zero (the cleared
// array), advection, horizontal pressure gradient and Coriolis/curvature at
// i [2, ide-1] in the reference's order, then rk_addtend_dry; written once.
// The pressure-gradient and Coriolis terms are evaluated at every point (at
// a column clamped into [2, ide-1], so their loads stay in the field) and
// used where they apply.
template <bool TEND>
__global__ void __launch_bounds__(64 * kFaceRows) k_advect_u(const DevCtx xv, const float *u, float *tend, const float *ru, const float *rv,
                                                  const float *rom, R3 r, int rk) {
  const DevCtx *__restrict__ x = &xv;
  __shared__ float ys[kFaceRows][64];
  GET_FACE_BLOCK(r);
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const unsigned nb = (unsigned)x->n3, sj = (unsigned)(kde * x->ip), sk = (unsigned)x->ip;
  const BufF bq = bufv(u, nb), bv = bufv(rv, nb);
  const bool adv = !TEND || (ic >= 2 && ic <= ide - 1);
  const unsigned o = XI3(ic, k, j);
  float fs;
  {
    float w[6];
    win6b(bq, o, sj, w);
    const float f = fy_u_sel(w, bld(bv, o) + bld(bv, o - 1u), j, jde, j >= 4 && j <= jde - 3);
    fs = (adv && j >= 2 && j <= jde - 1) ? f : 0.f;
  }
  ys[wv][lane] = fs;
  __syncthreads();
  if (!jrow) return;
  const BufF bu = bufv(ru, nb), bw = bufv(rom, nb);
  const unsigned ox = XI3(i, k, j);
  float wx[6], wz[5], wn[6];
  win6b(bq, ox, 1u, wx);
  win5z(bq, o, sk, wz);
  const float rus = bld(bu, ox) + bld(bu, ox - 1u);
  const float romt = bld(bw, o + sk) + bld(bw, o + sk - 1u), romb = bld(bw, o) + bld(bw, o - 1u);
  float rvn = 0.f;
  if (wv == kFaceRows - 1) {
    win6b(bq, o + sj, sj, wn);
    rvn = bld(bv, o + sj) + bld(bv, o + sj - 1u);
  }
  const int k1 = imin_(k + 1, kde);
  const float fm0 = A1(FL(fnm), k), fp0 = A1(FL(fnp), k), fm1 = A1(FL(fnm), k1), fp1 = A1(FL(fnp), k1);
  const float msf = A2(FL(msfux), ic, j), rdz = A1(FL(rdnw), k);
  const float t0 = TEND ? 0.f : elem(tend, o);
  float hp = 0.f, mt = 0.f;
  float4 cc = make_float4(0.f, 0.f, 0.f, 0.f);
  float2 tfl = make_float2(0.f, 0.f);
  if (TEND) {
    const int ia = imin_(imax_(ic, 2), ide - 1);
    hp = hpg_term(x, 0, ia, k, j);
    cc = cor_u_terms(x, ia, k, j);
    tfl = tendf_load(FL(ru_tendf), FL(u_save), o, rk);
    mt = A2(FL(msfuy), ic, j);
  }
  const float fw0 = fx_u_sel(wx, rus, i, ide, i - lane >= 4 && i - lane + 63 <= ide - 2);
  const float fw = (i >= 2 && i <= ide) ? fw0 : 0.f;
  const float fe = lane_next(fw, lane);
  if (!own) return;
  float fn;
  if (wv < kFaceRows - 1) fn = ys[wv + 1][lane];
  else fn = (adv && j + 1 <= jde - 1) ? fy_u_sel(wn, rvn, j + 1, jde, j + 1 >= 4 && j + 1 <= jde - 3) : 0.f;
  float t = t0;
  if (adv) {
    if (j >= 2 && j <= jde - 2) t -= msf * x->rdy * (fn - fs);
    t -= msf * x->rdx * (fe - fw);
    t -= rdz * (zflux_sel(wz[1], wz[2], wz[3], wz[4], 0.5f * romt, k + 1, kde, fm1, fp1) -
                zflux_sel(wz[0], wz[1], wz[2], wz[3], 0.5f * romb, k, kde, fm0, fp0));
    if (TEND) {
      t = t - hp;
      t = cor_u_add(t, cc);
    }
  }
  if (TEND) t = t + tendf_fold(FL(ru_tendf), o, tfl, mt, rk) / mt;
  elem(tend, o) = t;
}

// ---------------- v momentum (v rows j in [2, jde-1]) -----------------------
// advect_v (module_advect_em.py:1037-1354); r covers i [1, ide-1], j [2, jde-1]
// (face-sharing blocks: x faces i in [2, ide-1], y faces j in [2, jde]).
// TEND: the v tendency of the RK step over j [1, jde] (advection, pressure
// gradient, Coriolis/curvature at j [2, jde-1], then rk_addtend_dry); the
// pressure-gradient and Coriolis terms at a row clamped into [2, jde-1].
template <bool TEND>
__global__ void __launch_bounds__(64 * kFaceRows) k_advect_v(const DevCtx xv, const float *v, float *tend, const float *ru, const float *rv,
                                                  const float *rom, R3 r, int rk) {
  const DevCtx *__restrict__ x = &xv;
  __shared__ float ys[kFaceRows][64];
  GET_FACE_BLOCK(r);
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const unsigned nb = (unsigned)x->n3, sj = (unsigned)(kde * x->ip), sk = (unsigned)x->ip;
  const BufF bq = bufv(v, nb), bv = bufv(rv, nb);
  const bool adv = !TEND || (j >= 2 && j <= jde - 1);  // whole wave
  const unsigned o = XI3(ic, k, j);
  float fs;
  {
    float w[6];
    win6b(bq, o, sj, w);
    const float f = fy_v_sel(w, bld(bv, o) + bld(bv, o - sj), j, jde, j >= 4 && j <= jde - 2);
    fs = (j >= 2 && j <= jde) ? f : 0.f;
  }
  ys[wv][lane] = fs;
  __syncthreads();
  if (!jrow) return;
  const BufF bu = bufv(ru, nb), bw = bufv(rom, nb);
  const unsigned ox = XI3(i, k, j);
  float wx[6], wz[5], wn[6];
  win6b(bq, ox, 1u, wx);
  win5z(bq, o, sk, wz);
  const float rus = bld(bu, ox) + bld(bu, ox - sj);
  const float romt = bld(bw, o + sk) + bld(bw, o + sk - sj), romb = bld(bw, o) + bld(bw, o - sj);
  float rvn = 0.f;
  if (wv == kFaceRows - 1) {
    win6b(bq, o + sj, sj, wn);
    rvn = bld(bv, o + sj) + bld(bv, o);
  }
  const int k1 = imin_(k + 1, kde);
  const float fm0 = A1(FL(fnm), k), fp0 = A1(FL(fnp), k), fm1 = A1(FL(fnm), k1), fp1 = A1(FL(fnp), k1);
  const float msy = A2(FL(msfvy), ic, j), msx = A2(FL(msfvx), ic, j), rdz = A1(FL(rdnw), k);
  const float t0 = TEND ? 0.f : elem(tend, o);
  float hp = 0.f, mvx = 0.f, mvxi = 0.f;
  float4 cc = make_float4(0.f, 0.f, 0.f, 0.f);
  float2 tfl = make_float2(0.f, 0.f);
  if (TEND) {
    const int ja = imin_(imax_(j, 2), jde - 1);
    hp = hpg_term(x, 1, ic, k, ja);
    cc = cor_v_terms(x, ic, k, ja);
    tfl = tendf_load(FL(rv_tendf), FL(v_save), o, rk);
    mvx = msx;
    mvxi = A2(FL(msfvx_inv), ic, j);
  }
  const float fw0 = fx_v_sel(wx, rus, i, ide, i - lane >= 4 && i - lane + 63 <= ide - 3);
  const float fw = (adv && i >= 2 && i <= ide - 1) ? fw0 : 0.f;
  const float fe = lane_next(fw, lane);
  if (!own) return;
  float fn;
  if (wv < kFaceRows - 1) fn = ys[wv + 1][lane];
  else fn = (j + 1 >= 2 && j + 1 <= jde) ? fy_v_sel(wn, rvn, j + 1, jde, j + 1 >= 4 && j + 1 <= jde - 2) : 0.f;
  float t = t0;
  if (adv) {
    t -= msy * x->rdy * (fn - fs);
    if (ic >= 2 && ic <= ide - 2) t -= msy * x->rdx * (fe - fw);
    t -= (msy / msx) * rdz *
         (zflux_sel(wz[1], wz[2], wz[3], wz[4], 0.5f * romt, k + 1, kde, fm1, fp1) - zflux_sel(wz[0], wz[1], wz[2], wz[3], 0.5f * romb, k, kde, fm0, fp0));
    if (TEND) {
      t = t - hp;
      t = cor_v_add(t, cc);
    }
  }
  if (TEND) t = t + tendf_fold(FL(rv_tendf), o, tfl, mvx, rk) * mvxi;
  elem(tend, o) = t;
}

// ---------------- w momentum (levels k in [2, kde]) -------------------------
// advect_w (module_advect_em.py:1357-1640); r covers i [1, ide-1], k [2, kde], j [1, jde-1]
// (face-sharing blocks: x faces i in [2, ide-1], y faces j in [2, jde-1]).
// TEND: the w tendency of the RK step over k [1, kde]: advection, buoyancy
// and pressure (pg_buoy_w, reading cqw before k_cqw_inv replaces it) and, below
// the top, Coriolis/curvature at k [2, kde]; then rk_addtend_dry.  Buoyancy
// and Coriolis terms at a level clamped into the range they apply to.
template <bool TEND>
__global__ void __launch_bounds__(64 * kFaceRows) k_advect_w(const DevCtx xv, const float *w, float *tend, const float *ru, const float *rv,
                                                  const float *rom, R3 r, int rk) {
  const DevCtx *__restrict__ x = &xv;
  __shared__ float ys[kFaceRows][64];
  GET_FACE_BLOCK_LG(r, (unsigned)x->face_lg[1]);
  const int ide = x->ide, jde = x->jde, kde = x->kde;
  const unsigned nb = (unsigned)x->n3, sj = (unsigned)(kde * x->ip), sk = (unsigned)x->ip;
  const BufF bq = bufv(w, nb), bv = bufv(rv, nb);
  const bool adv = !TEND || k >= 2;  // whole block
  const unsigned o = XI3(ic, k, j);
  float fs;
  {
    float wy[6];
    win6b(bq, o, sj, wy);
    const float f = hflux_sel(wy, rvw_sel(x, bld(bv, o), bld(bv, o - sk), bld(bv, o - 2u * sk), k), j, jde, j >= 4 && j <= jde - 3);
    fs = (adv && j >= 2 && j <= jde - 1) ? f : 0.f;
  }
  ys[wv][lane] = fs;
  __syncthreads();
  if (!jrow) return;
  const BufF bu = bufv(ru, nb), bw = bufv(rom, nb);
  const unsigned ox = XI3(i, k, j);
  float wx[6], wz[5], wn[6];
  win6b(bq, ox, 1u, wx);
  win5z(bq, o, sk, wz);
  const float uw = rvw_sel(x, bld(bu, ox), bld(bu, ox - sk), bld(bu, ox - 2u * sk), k);
  const float rp = bld(bw, o + sk), r0 = bld(bw, o), rm = bld(bw, o - sk);
  float vn = 0.f;
  if (wv == kFaceRows - 1) {
    win6b(bq, o + sj, sj, wn);
    vn = rvw_sel(x, bld(bv, o + sj), bld(bv, o + sj - sk), bld(bv, o + sj - 2u * sk), k);
  }
  const float msf = A2(FL(msftx), ic, j), rdz0 = A1(FL(rdn), imin_(k, kde - 1)), rdz1 = A1(FL(rdn), imax_(k - 1, 1));
  const float t0 = TEND ? 0.f : elem(tend, o);
  float bu_t = 0.f, mt = 0.f, cq1 = 0.f;
  float2 cc = make_float2(0.f, 0.f), tfl = make_float2(0.f, 0.f);
  if (TEND) {
    // level 1 has no buoyancy term (its result is discarded); it reads the
    // top level's inputs, which no block of this launch writes, instead of
    // level 2's cqw, which the level-2 block replaces
    bu_t = buoy_w_term(x, ic, k < 2 ? kde : k, j, &cq1);
    cc = cor_w_terms(x, ic, imin_(imax_(k, 2), kde - 1), j);
    tfl = tendf_load(FL(rw_tendf), FL(w_save), o, rk);
    mt = A2(FL(msfty), ic, j);
  }
  const float fw0 = hflux_sel(wx, uw, i, ide, i - lane >= 4 && i - lane + 63 <= ide - 3);
  const float fw = (adv && i >= 2 && i <= ide - 1) ? fw0 : 0.f;
  const float fe = lane_next(fw, lane);
  if (!own) return;
  float fn;
  if (wv < kFaceRows - 1) fn = ys[wv + 1][lane];
  else fn = (adv && j + 1 <= jde - 1) ? hflux_sel(wn, vn, j + 1, jde, j + 1 >= 4 && j + 1 <= jde - 3) : 0.f;
  float t = t0;
  if (adv) {
    if (j >= 2 && j <= jde - 2) t -= msf * x->rdy * (fn - fs);
    if (ic >= 2 && ic <= ide - 2) t -= msf * x->rdx * (fe - fw);
    const float fzb = zflux_w_sel(wz[0], wz[1], wz[2], wz[3], r0 + rm, k, kde);
    if (k <= kde - 1) t -= rdz0 * (zflux_w_sel(wz[1], wz[2], wz[3], wz[4], rp + r0, k + 1, kde) - fzb);
    else t += 2.f * rdz1 * fzb;
    if (TEND) {
      t = t + bu_t;
      if (k <= kde - 1) t = cor_w_add(t, cc);
    }
  }
  if (TEND) t = t + tendf_fold(FL(rw_tendf), o, tfl, mt, rk) / mt;
  elem(tend, o) = t;
  // pg_buoy_w's cqw := 1 / (1 + cqw) at this point (read only by this
  // thread here; level kde-1, which the top level's thread also reads, is
  // replaced after this kernel)
  if (TEND && k >= 2 && k <= kde - 2) elem(FL(cqw), o) = cq1;
}

// the phi tendency of the RK step over k [1, kde]: rhs_ph at k [2, kde] from
// zero (the cleared array), then rk_addtend_dry; r = i [1, ide-1], k [1, kde],
// j [1, jde-1]
__global__ void __launch_bounds__(256) k_tend_ph(const DevCtx xv, R3 r, int rk) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ(r);
  const unsigned o = XI3(i, k, j);
  // rk_addtend_dry's loads go out first (addtend_ph in two parts)
  const float2 tfl = tendf_load(FL(ph_tendf), FL(ph_save), o, rk);
  const float m = A2(FL(msfty), i, j);
  float t = 0.f;
  if (k >= 2) t = rhs_ph_val(x, i, k, j, t);
  t = t + tendf_fold(FL(ph_tendf), o, tfl, 1.f, rk) / m;
  elem(FL(ph_tend), o) = t;
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
