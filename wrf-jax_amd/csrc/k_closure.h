// Learned eddy viscosity for smag_km (the km of calculate_km_kh,
// module_diffusion_em_fast.py:773-1083 / smag2d_km,
// module_diffusion_em.py:1502-1646): a 55 -> 64 -> 64 -> 1 network whose
// hidden layers run on bf16 MFMA with FP32 accumulation, trained on the GPU
// by the kernels below against sub-grid stresses coarse-grained from a finer
// run of this dycore.  Plain Smagorinsky stays the fallback (no network set,
// or a non-finite output).
//
// Network (parameters theta, FP32 master copy, offsets below):
//   x   = 55 features (+ 9 zero pads): the 3x3x3 stencils of log|D| and N^2
//         around the point, normalised, and log(dx / 1 km).  Input slot
//         h*32 + s: stencil point s < 27 (di = s%3 - 1, dk = (s/3)%3 - 1,
//         dj = s/9 - 1) of log|D| (h = 0) or N^2 (h = 1); slot 27 the spacing
//         feature; the other slots zero.  Both halves share one stencil
//         order, so a lane of the MFMA layout reads the same 8 stencil
//         points for either half.
//   h1  = silu(W1 x + b1)        bf16 inputs and weights, FP32 accumulate;
//                                h1 rounded to bf16 for the next layer
//   h2  = silu(W2 h1 + b2)       the same; h2 kept in FP32
//   z   = w3 . h2 + b3           FP32
//   km  = kref * softplus(z)
// then, as for Smagorinsky, km = min(km, 10 l) / divisor (the diff_opt 2
// slope correction), kh = km / Pr.
//
// MFMA tiles: v_mfma_f32_16x16x32_bf16, C[16x16] += A[16x32] B[32x16].
// Operands are read from LDS or global arrays laid out k-contiguous for both
// A (rows m, k) and B^T (rows n, k); lane l feeds row l%16 with the 8
// consecutive k of block l/16, and receives C rows 4*(l/16) + 0..3 of
// column l%16.
#pragma once

#include "wrfdc_common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 64;    // padded input width
constexpr int kMlpFeat = 55;  // real inputs
constexpr int kMlpHid = 64;
constexpr int kOffW1 = 0, kOffB1 = 4096, kOffW2 = 4160, kOffB2 = 8256, kOffW3 = 8320, kOffB3 = 8384;
constexpr int kMlpParams = 8385;
// bf16 copies of the weights: W1 [64][64], W2 [64][64], W2^T [64][64]
constexpr int kWbfW1 = 0, kWbfW2 = 4096, kWbfW2T = 8192, kWbfSize = 12288;
// columns of a training row's targets: t1 = (tau11 - tau22)/2, t2 = tau12,
// s1 = (D11 - D22)/2, s2 = D12 (at the mass point), slope divisor, Smagorinsky km
constexpr int kTargetCols = 6;

struct ClosureNorm {
  float d_shift, d_scale, n_shift, n_scale, dx_feature, kref;
};

// SiLU and the logistic function through the hardware exp2 and reciprocal
// (a few ulp; the network is trained and evaluated with the same forms)
__device__ __forceinline__ float sigmoid_f(float v) { return __builtin_amdgcn_rcpf(1.f + __expf(-v)); }
__device__ __forceinline__ float silu_f(float v) { return v * sigmoid_f(v); }
__device__ __forceinline__ float softplus_f(float v) { return v > 20.f ? v : log1pf(expf(v)); }

__device__ __forceinline__ f32x4 mma16(const __bf16 *A, int lda, const __bf16 *Bt, int ldb, f32x4 acc) {
  const int l = threadIdx.x & 63;
  const int row = l & 15, kb = (l >> 4) * 8;
  const bf16x8 a = *reinterpret_cast<const bf16x8 *>(A + row * lda + kb);
  const bf16x8 b = *reinterpret_cast<const bf16x8 *>(Bt + row * ldb + kb);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// features of one point
// ---------------------------------------------------------------------------
// The smag2d_km range: i [2, ide-2], k [1, kde-1], j [2, jde-2]; stencil
// neighbours outside it take the nearest point of it.
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// normalised inputs of every point of the range: fd = (log(|D| + 1e-7) -
// d_shift) / d_scale and fn = (N^2 - n_shift) / n_scale (|D| as k_smag
// stored it)
__global__ void k_closure_feat(const DevCtx xv, R3 r, ClosureNorm nm, const float *__restrict__ defmag, float *__restrict__ fd,
                               float *__restrict__ fn) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  A3(fd, i, k, j) = (logf(A3(defmag, i, k, j) + 1e-7f) - nm.d_shift) / nm.d_scale;
  A3(fn, i, k, j) = (A3(FL(bn2), i, k, j) - nm.n_shift) / nm.n_scale;
}

// input slot f of the point (i, k, j) (the slot order above); stencil
// neighbours outside the range take the nearest point of it
__device__ __forceinline__ float closure_feature(const DevCtx *__restrict__ x, const float *fd, const float *fn, const ClosureNorm &nm,
                                                 int i, int k, int j, int f) {
  const int h = f >> 5, g = f & 31;
  if (g == 27) return h == 0 ? nm.dx_feature : 0.f;
  if (g > 27) return 0.f;
  const int di = g % 3 - 1, dk = (g / 3) % 3 - 1, dj = g / 9 - 1;
  const int ii = clampi(i + di, 2, x->ide - 2), kk = clampi(k + dk, 1, x->kde - 1), jj = clampi(j + dj, 2, x->jde - 2);
  return A3(h == 0 ? fd : fn, ii, kk, jj);
}

// ---------------------------------------------------------------------------
// network forward over 16 rows: A rows (bf16, k-contiguous, stride kLd) in
// LDS -> z per row into zout[16] (LDS).  H (bf16 [16][72]) is LDS scratch;
// W1, W2 rows (one per output unit) have stride ldw (64 in global memory,
// kLd in LDS).
// ---------------------------------------------------------------------------
constexpr int kLd = 72;  // LDS row stride of bf16 tiles (64 + 8: no bank conflicts on the 16-byte reads)

__device__ __forceinline__ void mlp_rows16(const __bf16 *Xs, const __bf16 *w1, const __bf16 *w2, int ldw, const float *theta,
                                           __bf16 *Hs, float *P1, float *P2, float *zout) {
  const int l = threadIdx.x & 63;
  const int col = l & 15, rb = (l >> 4) * 4;
  // layer 1
#pragma unroll
  for (int t = 0; t < 4; t++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mma16(Xs, kLd, w1 + (t * 16) * ldw, ldw, acc);
    acc = mma16(Xs + 32, kLd, w1 + (t * 16) * ldw + 32, ldw, acc);
    const int o = t * 16 + col;
    const float b = theta[kOffB1 + o];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const float pre = acc[r] + b;
      if (P1) P1[(rb + r) * 64 + o] = pre;
      Hs[(rb + r) * kLd + o] = (__bf16)silu_f(pre);
    }
  }
  __syncthreads();
  // layer 2 and the output layer
  float zpart[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 4; t++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mma16(Hs, kLd, w2 + (t * 16) * ldw, ldw, acc);
    acc = mma16(Hs + 32, kLd, w2 + (t * 16) * ldw + 32, ldw, acc);
    const int o = t * 16 + col;
    const float b = theta[kOffB2 + o], w3 = theta[kOffW3 + o];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const float pre = acc[r] + b;
      if (P2) P2[(rb + r) * 64 + o] = pre;
      zpart[r] += silu_f(pre) * w3;
    }
  }
  // sum over the 16 lanes that hold the same rows (lane bits 0..3)
#pragma unroll
  for (int r = 0; r < 4; r++) {
    float v = zpart[r];
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    zpart[r] = v;
  }
  if (col == 0) {
#pragma unroll
    for (int r = 0; r < 4; r++) zout[rb + r] = zpart[r] + theta[kOffB3];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// smag2d_km with the learned km: one wave per 64 points (64 i of one level
// and row), as 4 tiles of 16 points.  The MFMA operands come straight from
// registers: lane l feeds A row l%16 (point 16t + l%16 of tile t) with the
// 8 k of block l/16 in each half of the 64 inputs, i.e. stencil points
// s = 8*(l/16) .. +7 of log|D| (first MFMA) and of N^2 (second), so each
// lane gathers exactly its 16 operands; the stencil offsets of its 8 points
// are formed once per wave (k and j are the wave's own).  W1 and W2 (B
// operands) are staged in LDS once per block; the first hidden layer's
// output goes through the wave's own LDS tile (C layout to A layout), so
// the waves synchronise only within themselves after the staging.  k_smag has stored the slope divisor and k_closure_feat the
// normalised inputs (fd, fn: two arrays n3 apart) before.
// Block (64, 4, 1), grid as for GET_IKJ_P over i [2, ide-2], k [1, kde-1],
// j [2, jde-2].
// ---------------------------------------------------------------------------
// LDS written by some lanes of a wave, then read by others of the same
// wave: the wave's LDS operations complete and are ordered before the reads
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ bf16x8 pack8(const float v[8]) {
  bf16x8 r;
#pragma unroll
  for (int q = 0; q < 8; q++) r[q] = (__bf16)v[q];
  return r;
}
__global__ void __launch_bounds__(256) k_closure_km(const DevCtx xv, R3 r, const float *__restrict__ theta, const __bf16 *__restrict__ wbf,
                                                    ClosureNorm nm, const float *__restrict__ fd, const float *__restrict__ divisor) {
  const DevCtx *__restrict__ x = &xv;
  __shared__ __align__(16) __bf16 s_w[2][64 * kLd];
  __shared__ __align__(16) __bf16 s_h[4][16 * kLd];
  __shared__ float s_z[4][64];
  const int wv = threadIdx.y, l = threadIdx.x;
  const int col = l & 15, grp = l >> 4, rb = grp * 4;
  for (int q = wv * 64 + l; q < 2 * 64 * 8; q += 256) {  // 16-byte pieces: matrix, row, piece
    const int m = q >> 9, row = (q >> 3) & 63, c = q & 7;
    *reinterpret_cast<bf16x8 *>(&s_w[m][row * kLd + c * 8]) =
        *reinterpret_cast<const bf16x8 *>(wbf + (m == 0 ? kWbfW1 : kWbfW2) + row * 64 + c * 8);
  }
  // the tiles of GET_IKJ_P: 64-aligned in i from ib <= r.i0
  const int ib = 1 + ((r.i0 - 1) & ~63);
  const unsigned gx_ = (unsigned)(r.i1 - ib + 64) / 64u;
  const unsigned gy_ = (unsigned)(r.k1 - r.k0 + 4) / 4u;
  const int i0 = ib + (int)((blockIdx.x % gx_) * 64u);
  const int i = i0 + l;
  const int k = r.k0 + (int)(((blockIdx.x / gx_) % gy_) * 4u + threadIdx.y);
  const int j = r.j0 + (int)(blockIdx.x / (gx_ * gy_));
  // the whole wave takes part in the MFMA; points outside r compute on a
  // clamped copy and store nothing
  const bool own = i >= r.i0 && i <= r.i1 && k <= r.k1 && j <= r.j1;
  const int kc = k <= r.k1 ? k : r.k0, jc = j <= r.j1 ? j : r.j0;
  // per-column biases of the C layout
  float b1[4], b2[4], w3[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    b1[t] = theta[kOffB1 + t * 16 + col];
    b2[t] = theta[kOffB2 + t * 16 + col];
    w3[t] = theta[kOffW3 + t * 16 + col];
  }
  // this lane's 8 stencil points s = 8*grp + q: level / row part of the
  // offset (i added per tile), the i shift, and whether the slot is a
  // stencil point (s < 27), the spacing slot (27) or zero
  const unsigned nb = (unsigned)(2 * x->n3);
  const BufF bf = bufv(fd, nb);
  int koff[8], dis[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int sp = 8 * grp + q, g = sp < 27 ? sp : 13;
    const int di = g % 3 - 1, dk = (g / 3) % 3 - 1, dj = g / 9 - 1;
    const int kk = clampi(kc + dk, 1, x->kde - 1), jj = clampi(jc + dj, 2, x->jde - 2);
    koff[q] = (int)XI3(1, kk, jj) - 1;  // element offset of (i, kk, jj) is koff + i
    dis[q] = di;
  }
  const unsigned n3 = (unsigned)x->n3;
  __syncthreads();  // the staged weights
  // B operand of output tile t2, input half h: W row t2*16 + l%16, k block l/16
#define WB(m, t2, h) (*reinterpret_cast<const bf16x8 *>(&s_w[m][((t2) * 16 + col) * kLd + (h) * 32 + grp * 8]))
  for (int t = 0; t < 4; t++) {
    const int ip = imin_(imax_(i0 + 16 * t + col, r.i0), r.i1);
    float vd[8], vn[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int sp = 8 * grp + q;
      const unsigned o = (unsigned)(koff[q] + clampi(ip + dis[q], 2, x->ide - 2));
      const float a = bld(bf, o), b = bld(bf, o + n3);
      vd[q] = sp < 27 ? a : (sp == 27 ? nm.dx_feature : 0.f);
      vn[q] = sp < 27 ? b : 0.f;
    }
    const bf16x8 a0 = pack8(vd), a1 = pack8(vn);
    // layer 1 -> LDS (bf16, rows = points)
#pragma unroll
    for (int t2 = 0; t2 < 4; t2++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, WB(0, t2, 0), acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, WB(0, t2, 1), acc, 0, 0, 0);
#pragma unroll
      for (int q = 0; q < 4; q++) s_h[wv][(rb + q) * kLd + t2 * 16 + col] = (__bf16)silu_f(acc[q] + b1[t2]);
    }
    wave_lds_sync();
    const bf16x8 h0 = *reinterpret_cast<const bf16x8 *>(&s_h[wv][col * kLd + grp * 8]);
    const bf16x8 h1 = *reinterpret_cast<const bf16x8 *>(&s_h[wv][col * kLd + 32 + grp * 8]);
    // layer 2 and the output layer
    float zp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t2 = 0; t2 < 4; t2++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h0, WB(1, t2, 0), acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h1, WB(1, t2, 1), acc, 0, 0, 0);
#pragma unroll
      for (int q = 0; q < 4; q++) zp[q] += silu_f(acc[q] + b2[t2]) * w3[t2];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      float v = zp[q];
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      v += __shfl_xor(v, 4);
      v += __shfl_xor(v, 8);
      if (col == 0) s_z[wv][16 * t + rb + q] = v + theta[kOffB3];
    }
    // the next tile overwrites s_h (read above); s_z is read after the loop
    wave_lds_sync();
  }
#undef WB
  if (!own) return;
  const float smag = A3(FL(xkmh), i, k, j);
  const float mlen = sqrtf(x->dx / A2(FL(msftx), i, j) * x->dy / A2(FL(msfty), i, j));
  float km = nm.kref * softplus_f(s_z[wv][l]);
  km = fminf(km, 10.f * mlen) / A3(divisor, i, k, j);
  if (!isfinite(km)) km = smag;
  A3(FL(xkmh), i, k, j) = km;
  A3(FL(xkmv), i, k, j) = km;
  A3(FL(xkhh), i, k, j) = km / x->prandtl;
}

// ---------------------------------------------------------------------------
// training rows of the current state: features (bf16 [n][64]) and the
// target columns s1, s2, divisor, Smagorinsky km of every point of the
// smag range, row = ((j-2)*nk + (k-1))*ni + (i-2).  t1, t2 come from
// k_subgrid_stress.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_closure_rows(const DevCtx xv, R3 r, ClosureNorm nm, const float *__restrict__ fd,
                                                      const float *__restrict__ fn, const float *__restrict__ divisor, __bf16 *X, float *T) {
  const DevCtx *__restrict__ x = &xv;
  GET_IKJ_P(r);
  const int ni = r.i1 - r.i0 + 1, nk = r.k1 - r.k0 + 1;
  const size_t row = ((size_t)(j - r.j0) * nk + (size_t)(k - r.k0)) * ni + (size_t)(i - r.i0);
  for (int f = 0; f < kMlpIn; f++) X[row * kMlpIn + f] = (__bf16)closure_feature(x, fd, fn, nm, i, k, j, f);
  const float d11 = A3(FL(defor11), i, k, j), d22 = A3(FL(defor22), i, k, j);
  const float *d12 = FL(defor12);
  const float s2 = 0.25f * (A3(d12, i, k, j) + A3(d12, i, k, j + 1) + A3(d12, i + 1, k, j) + A3(d12, i + 1, k, j + 1));
  T[row * kTargetCols + 2] = 0.5f * (d11 - d22);
  T[row * kTargetCols + 3] = s2;
  T[row * kTargetCols + 4] = A3(divisor, i, k, j);
  T[row * kTargetCols + 5] = A3(FL(xkmh), i, k, j);
}

// Sub-grid stresses of a fine domain on the mass points of a grid coarser by
// `ratio` in x and y (same levels): block means of u and v at the fine mass
// points and of their products, t1 = (tau11 - tau22)/2, t2 = tau12 with
// tau_ab = <a b> - <a><b>; written to columns 0 and 1 of the training rows of
// the coarse range rc (the same row order as k_closure_rows).
__global__ void __launch_bounds__(256) k_subgrid_stress(const DevCtx xv, R3 rc, int ratio, float *T) {
  const DevCtx *__restrict__ x = &xv;  // the fine domain
  GET_IKJ_P(rc);                        // coarse point (i, k, j)
  const float *u = FL(u_2), *v = FL(v_2);
  double su = 0.0, sv = 0.0, suu = 0.0, svv = 0.0, suv = 0.0;
  for (int b = 0; b < ratio; b++)
    for (int a = 0; a < ratio; a++) {
      const int fi = (i - 1) * ratio + 1 + a, fj = (j - 1) * ratio + 1 + b;
      const float um = 0.5f * (A3(u, fi, k, fj) + A3(u, fi + 1, k, fj));
      const float vm = 0.5f * (A3(v, fi, k, fj) + A3(v, fi, k, fj + 1));
      su += um;
      sv += vm;
      suu += (double)um * um;
      svv += (double)vm * vm;
      suv += (double)um * vm;
    }
  const double n = (double)(ratio * ratio);
  const double mu = su / n, mv = sv / n;
  const double t11 = suu / n - mu * mu, t22 = svv / n - mv * mv, t12 = suv / n - mu * mv;
  const int ni = rc.i1 - rc.i0 + 1, nk = rc.k1 - rc.k0 + 1;
  const size_t row = ((size_t)(j - rc.j0) * nk + (size_t)(k - rc.k0)) * ni + (size_t)(i - rc.i0);
  T[row * kTargetCols + 0] = (float)(0.5 * (t11 - t22));
  T[row * kTargetCols + 1] = (float)t12;
}

// Block means of a fine field onto a coarser grid (factor ratio in x and y,
// same levels).  stag 0: mass points (cells), 1: x faces (u), 2: y faces (v);
// nk levels, 2-D fields with nk = 1.  dst/src carry their own pitches.
struct CoarseField {
  float *dst;
  const float *src;
  int stag, nk;
};
constexpr int kMaxCoarse = 48;
struct CoarseList {
  CoarseField f[kMaxCoarse];
  int n;
  int ide_c, jde_c, ip_c, kde_c;  // coarse extents (staggered) and pitch
  int ip_f, jde_f, kde_f;         // fine pitch and extents
  int ratio;
};
__global__ void __launch_bounds__(256) k_coarse_grain(CoarseList L) {
  const CoarseField &F = L.f[blockIdx.y];
  const size_t ncol = (size_t)L.ide_c * L.jde_c;
  const size_t total = ncol * (size_t)F.nk;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int I = 1 + (int)(t % L.ide_c), J = 1 + (int)((t / L.ide_c) % L.jde_c), k = 1 + (int)(t / ncol);
    const int R = L.ratio;
    const size_t plane_f = (size_t)L.jde_f * L.ip_f;  // unused for 2-D (nk = 1)
    auto src = [&](int fi, int fj) -> float {
      return F.src[((size_t)(fj - 1) * (F.nk == 1 ? 1 : L.kde_f) + (size_t)(k - 1)) * L.ip_f + (size_t)(fi - 1)];
    };
    (void)plane_f;
    float v = 0.f;
    bool ok = true;
    if (F.stag == 0) {  // cell I covers fine cells (I-1)R+1 .. IR
      if (I > L.ide_c - 1 || J > L.jde_c - 1) ok = false;
      else {
        double s = 0.0;
        for (int b = 0; b < R; b++)
          for (int a = 0; a < R; a++) s += src((I - 1) * R + 1 + a, (J - 1) * R + 1 + b);
        v = (float)(s / (R * R));
      }
    } else if (F.stag == 1) {  // x face I sits on fine face (I-1)R+1
      if (J > L.jde_c - 1) ok = false;
      else {
        double s = 0.0;
        for (int b = 0; b < R; b++) s += src((I - 1) * R + 1, (J - 1) * R + 1 + b);
        v = (float)(s / R);
      }
    } else {  // y face J sits on fine face (J-1)R+1
      if (I > L.ide_c - 1) ok = false;
      else {
        double s = 0.0;
        for (int a = 0; a < R; a++) s += src((I - 1) * R + 1 + a, (J - 1) * R + 1);
        v = (float)(s / R);
      }
    }
    if (ok) F.dst[((size_t)(J - 1) * (F.nk == 1 ? 1 : L.kde_c) + (size_t)(k - 1)) * L.ip_c + (size_t)(I - 1)] = v;
  }
}

// ---------------------------------------------------------------------------
// training: forward, loss and backward over rows, one wave per block; each
// wave loops over chunks of 32 rows and accumulates its gradient, written as
// one partial (kMlpParams + 1 floats: gradient, then the loss sum) per block.
// Loss of a row: ((t1 + K s1)^2 + (t2 + K s2)^2) * lossnorm with
// K = kref * softplus(z) / divisor.  BACKWARD false: forward only, z -> out.
// ---------------------------------------------------------------------------
constexpr int kChunk = 32;
constexpr int kLdT = 40;  // LDS stride of transposed [64][32] tiles

template <bool BACKWARD>
__global__ void __launch_bounds__(64) k_mlp_train(const float *__restrict__ theta, const __bf16 *__restrict__ wbf, const __bf16 *__restrict__ X,
                                                  const float *__restrict__ T, long long n, float lossnorm, float kref, int chunks_per_block,
                                                  float *__restrict__ out) {
  __shared__ __align__(16) __bf16 s_w[3][64 * kLd];  // W1, W2, W2^T
  __shared__ __align__(16) __bf16 s_x[kChunk * kLd];
  __shared__ __align__(16) __bf16 s_xt[64 * kLdT];
  __shared__ __align__(16) __bf16 s_h1[kChunk * kLd];
  __shared__ __align__(16) __bf16 s_h1t[64 * kLdT];
  __shared__ __align__(16) __bf16 s_d2[kChunk * kLd];
  __shared__ __align__(16) __bf16 s_d2t[64 * kLdT];
  __shared__ __align__(16) __bf16 s_d1t[64 * kLdT];
  __shared__ float s_p1[kChunk * 64];
  __shared__ float s_p2[kChunk * 64];
  __shared__ float s_z[kChunk];
  __shared__ float s_dz[kChunk];
  const int l = threadIdx.x;
  for (int m = 0; m < 3; m++)
    for (int e = l; e < 64 * 64; e += 64) s_w[m][(e / 64) * kLd + e % 64] = wbf[m * 4096 + e];
  __syncthreads();
  const int col = l & 15, rb = (l >> 4) * 4;
  f32x4 gw1[4][4], gw2[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      gw1[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
      gw2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, loss = 0.f;  // lane l: hidden unit l
  for (int c = 0; c < chunks_per_block; c++) {
    const long long r0 = ((long long)blockIdx.x * chunks_per_block + c) * kChunk;
    if (r0 >= n) break;
    // rows of the chunk (rows past n are zero and carry no loss)
    for (int e = l; e < kChunk * kMlpIn; e += 64) {
      const int p = e / kMlpIn, f = e % kMlpIn;
      const __bf16 v = (r0 + p < n) ? X[(r0 + p) * kMlpIn + f] : (__bf16)0.f;
      s_x[p * kLd + f] = v;
      s_xt[f * kLdT + p] = v;
    }
    __syncthreads();
    mlp_rows16(s_x, s_w[0], s_w[1], kLd, theta, s_h1, s_p1, s_p2, s_z);
    mlp_rows16(s_x + 16 * kLd, s_w[0], s_w[1], kLd, theta, s_h1 + 16 * kLd, s_p1 + 16 * 64, s_p2 + 16 * 64, s_z + 16);
    if (!BACKWARD) {
      if (l < kChunk && r0 + l < n) out[r0 + l] = s_z[l];
      __syncthreads();
      continue;
    }
    // loss and dL/dz of each row
    if (l < kChunk) {
      float dz = 0.f;
      if (r0 + l < n) {
        const float *t = T + (r0 + l) * kTargetCols;
        const float z = s_z[l], div = t[4];
        const float K = kref * softplus_f(z) / div;
        const float r1 = t[0] + K * t[2], r2 = t[1] + K * t[3];
        loss += (r1 * r1 + r2 * r2) * lossnorm;
        const float dK = 2.f * (r1 * t[2] + r2 * t[3]) * lossnorm;
        dz = dK * kref * sigmoid_f(z) / div;
      }
      s_dz[l] = dz;
    }
    __syncthreads();
    // output layer and dL/dpre2 (lane l: hidden unit l, all rows)
    {
      const float w3 = theta[kOffW3 + l];
      for (int p = 0; p < kChunk; p++) {
        const float pre = s_p2[p * 64 + l], s = sigmoid_f(pre), h = pre * s;
        const float dz = s_dz[p];
        gw3 += dz * h;
        const float d2 = dz * w3 * (s + pre * s * (1.f - s));
        gb2 += d2;
        const __bf16 d2b = (__bf16)d2;
        s_d2[p * kLd + l] = d2b;
        s_d2t[l * kLdT + p] = d2b;
        s_h1t[l * kLdT + p] = s_h1[p * kLd + l];
      }
      if (l == 0)
        for (int p = 0; p < kChunk; p++) gb3 += s_dz[p];
    }
    __syncthreads();
    // dW2[o][i] += sum_p d2[p][o] h1[p][i]
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) gw2[a][b] = mma16(s_d2t + a * 16 * kLdT, kLdT, s_h1t + b * 16 * kLdT, kLdT, gw2[a][b]);
    // dh1[p][i] = sum_o d2[p][o] W2[o][i], then dL/dpre1
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll
      for (int b = 0; b < 4; b++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mma16(s_d2 + h * 16 * kLd, kLd, s_w[2] + b * 16 * kLd, kLd, acc);
        acc = mma16(s_d2 + h * 16 * kLd + 32, kLd, s_w[2] + b * 16 * kLd + 32, kLd, acc);
        const int ihid = b * 16 + col;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int p = h * 16 + rb + r;
          const float pre = s_p1[p * 64 + ihid], s = sigmoid_f(pre);
          s_d1t[ihid * kLdT + p] = (__bf16)(acc[r] * (s + pre * s * (1.f - s)));
        }
      }
    }
    __syncthreads();
    for (int p = 0; p < kChunk; p++) gb1 += (float)s_d1t[l * kLdT + p];
    // dW1[i][f] += sum_p d1[p][i] x[p][f]
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) gw1[a][b] = mma16(s_d1t + a * 16 * kLdT, kLdT, s_xt + b * 16 * kLdT, kLdT, gw1[a][b]);
    __syncthreads();
  }
  if (!BACKWARD) return;
  float *g = out + (size_t)blockIdx.x * (kMlpParams + 1);
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = a * 16 + rb + r, cc = b * 16 + col;
        g[kOffW1 + row * 64 + cc] = gw1[a][b][r];
        g[kOffW2 + row * 64 + cc] = gw2[a][b][r];
      }
  g[kOffB1 + l] = gb1;
  g[kOffB2 + l] = gb2;
  g[kOffW3 + l] = gw3;
  // the loss of the rows is spread over lanes 0..31: sum it
  float ls = loss;
  for (int off = 32; off > 0; off >>= 1) ls += __shfl_xor(ls, off);
  if (l == 0) {
    g[kOffB3] = gb3;
    g[kMlpParams] = ls;
  }
}

// Sum of the per-block partials (fixed order), then the gradient out and,
// with adam != 0, one Adam step on theta and fresh bf16 weight copies.
// One thread per parameter (and one for the loss).
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ part, int nblocks, float *grad, float *loss, float *theta,
                                                    float *m, float *v, __bf16 *wbf, int adam, int step, float lr) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kMlpParams) return;
  float s = 0.f;
  for (int b = 0; b < nblocks; b++) s += part[(size_t)b * (kMlpParams + 1) + p];
  if (p == kMlpParams) {
    if (loss) *loss = s;
    return;
  }
  if (grad) grad[p] = s;
  if (!adam) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mm = b1 * m[p] + (1.f - b1) * s, vv = b2 * v[p] + (1.f - b2) * s * s;
  m[p] = mm;
  v[p] = vv;
  const float mh = mm / (1.f - powf(b1, (float)step)), vh = vv / (1.f - powf(b2, (float)step));
  const float th = theta[p] - lr * mh / (sqrtf(vh) + eps);
  theta[p] = th;
  if (p < kOffB1) {
    wbf[kWbfW1 + p] = (__bf16)th;
  } else if (p >= kOffW2 && p < kOffB2) {
    const int e = p - kOffW2, o = e / 64, in = e % 64;
    wbf[kWbfW2 + e] = (__bf16)th;
    wbf[kWbfW2T + in * 64 + o] = (__bf16)th;
  }
}

// bf16 copies of W1, W2 and W2^T from theta
__global__ void k_mlp_pack(const float *__restrict__ theta, __bf16 *wbf) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 4096) return;
  wbf[kWbfW1 + e] = (__bf16)theta[kOffW1 + e];
  const float w2 = theta[kOffW2 + e];
  wbf[kWbfW2 + e] = (__bf16)w2;
  wbf[kWbfW2T + (e % 64) * 64 + e / 64] = (__bf16)w2;
}
