// Shared definitions of the MI355X dycore library (device + host side).
//
// Layout in HBM: every 3-D field is [j][k][i] float32 with i contiguous and a
// row pitch `ip` (>= ide, padded to a multiple of 16 floats = 64 B); 2-D
// fields are [j][i] with the same pitch; 1-D column arrays are [k].  Indices
// in kernels are Fortran-style 1-based (i, k, j) exactly as in the reference
// dycore (src/wrf_jax/dyn_em), with ide = nx+1, jde = ny+1, kde = nz+1.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/wrfdc.h"

#define WRFDC_FIELDS(X)                                                       \
  X(u_1, '3') X(u_2, '3') X(v_1, '3') X(v_2, '3') X(w_1, '3') X(w_2, '3')     \
  X(t_1, '3') X(t_2, '3') X(ph_1, '3') X(ph_2, '3') X(p, '3') X(al, '3')      \
  X(pb, '3') X(phb, '3') X(alb, '3') X(ru, '3') X(rv, '3') X(rw, '3')         \
  X(ww, '3') X(php, '3') X(alt, '3') X(cqu, '3') X(cqv, '3') X(cqw, '3')      \
  X(c2a, '3') X(pm1, '3') X(a, '3') X(alpha, '3') X(gamma, '3')               \
  X(ru_tend, '3') X(rv_tend, '3') X(rw_tend, '3') X(t_tend, '3')              \
  X(ph_tend, '3') X(ru_tendf, '3') X(rv_tendf, '3') X(rw_tendf, '3')          \
  X(t_tendf, '3') X(ph_tendf, '3') X(u_save, '3') X(v_save, '3')              \
  X(w_save, '3') X(t_save, '3') X(ph_save, '3') X(ww_1, '3') X(t_2ave, '3')   \
  X(ru_m, '3') X(rv_m, '3') X(ww_m, '3') X(h_diabatic, '3')                  \
  X(advect_tend, '3') X(rho, '3') X(z, '3') X(rdz, '3') X(rdzw, '3')          \
  X(zx, '3') X(zy, '3') X(div, '3') X(defor11, '3') X(defor22, '3')           \
  X(defor33, '3') X(defor12, '3') X(defor13, '3') X(defor23, '3')             \
  X(xkmh, '3') X(xkmv, '3') X(xkhh, '3') X(xkhv, '3') X(bn2, '3')             \
  X(mu_1, '2') X(mu_2, '2') X(mub, '2') X(mut, '2') X(muu, '2') X(muv, '2')   \
  X(muts, '2') X(muus, '2') X(muvs, '2') X(muave, '2') X(mudf, '2')           \
  X(mu_tend, '2') X(mu_tendf, '2') X(mu_save, '2') X(ht, '2') X(msfux, '2')   \
  X(msfuy, '2') X(msfvx, '2') X(msfvx_inv, '2') X(msfvy, '2') X(msftx, '2')   \
  X(msfty, '2') X(f, '2') X(e, '2') X(sina, '2') X(cosa, '2') X(clat, '2')    \
  X(znw, '1') X(znu, '1') X(dnw, '1') X(rdnw, '1') X(dn, '1') X(rdn, '1')     \
  X(fnm, '1') X(fnp, '1') X(c1h, '1') X(c2h, '1') X(c1f, '1') X(c2f, '1')     \
  X(c3h, '1') X(c4h, '1') X(c3f, '1') X(c4f, '1') X(u_base, '1')              \
  X(v_base, '1') X(moist, 'M') X(moist_old, 'M') X(moist_tend, 'M')           \
  X(scalar, 'S') X(scalar_old, 'S') X(scalar_tend, 'S')                       \
  X(u_bxs, 'X') X(u_bxe, 'X') X(u_bys, 'Y') X(u_bye, 'Y')                     \
  X(u_btxs, 'X') X(u_btxe, 'X') X(u_btys, 'Y') X(u_btye, 'Y')                 \
  X(v_bxs, 'X') X(v_bxe, 'X') X(v_bys, 'Y') X(v_bye, 'Y')                     \
  X(v_btxs, 'X') X(v_btxe, 'X') X(v_btys, 'Y') X(v_btye, 'Y')                 \
  X(t_bxs, 'X') X(t_bxe, 'X') X(t_bys, 'Y') X(t_bye, 'Y')                     \
  X(t_btxs, 'X') X(t_btxe, 'X') X(t_btys, 'Y') X(t_btye, 'Y')                 \
  X(ph_bxs, 'X') X(ph_bxe, 'X') X(ph_bys, 'Y') X(ph_bye, 'Y')                 \
  X(ph_btxs, 'X') X(ph_btxe, 'X') X(ph_btys, 'Y') X(ph_btye, 'Y')             \
  X(qv_bxs, 'X') X(qv_bxe, 'X') X(qv_bys, 'Y') X(qv_bye, 'Y')                 \
  X(qv_btxs, 'X') X(qv_btxe, 'X') X(qv_btys, 'Y') X(qv_btye, 'Y')             \
  X(mu_bxs, 'x') X(mu_bxe, 'x') X(mu_bys, 'y') X(mu_bye, 'y')                 \
  X(mu_btxs, 'x') X(mu_btxe, 'x') X(mu_btys, 'y') X(mu_btye, 'y')

enum {
#define WRFDC_ENUM(name, kind) F_##name,
  WRFDC_FIELDS(WRFDC_ENUM)
#undef WRFDC_ENUM
  NFIELDS
};

// Library-owned scratch arrays (3-D, same layout as fields).
enum { S_T1, S_T2, S_FQX, S_FQY, S_FQZ, S_FQXL, S_FQYL, S_FQZL, S_SCALE, S_HAT2, S_VXGM, NSCRATCH };

// View of one domain (dimensions, scalars, pointers), passed to every kernel
// by value (about 1.7 KB of kernel arguments).
struct DevCtx {
  int ide, jde, kde, ip;
  int nmoist, nscalar, spec_zone, relax_zone, spec_bdy_width, top_lid;
  size_t n3, n2;
  float rdx, rdy, dx, dy, dt, g, t0, p0, ptop, r_d, cpovcv, rvovrd, epssm, smdiv, emdiv, c_s, prandtl;
  float cf1, cf2, cf3, cfn, cfn1, reradius;
  float fcx[16], gcx[16];
  // levels per level group of the face-block tile order (wrfdc_set_tile):
  // [0] advection of u, v, theta and the plain species update, [1] advection
  // of w, [2] the positive-definite species passes
  int face_lg[3];
  float *F[NFIELDS];
  float *S[NSCRATCH];
};

// Element offsets in 32-bit arithmetic, scaled to byte offsets in 32-bit
// arithmetic too (a 3-D field holds < 2^30 elements, so its byte offsets fit;
// wrfdc_create checks), added to 64-bit base pointers held in scalar
// registers: the loads and stores use the saddr + 32-bit voffset form and
// need no 64-bit address arithmetic per access.
#define XI3(i, k, j) ((unsigned)((((j) - 1) * x->kde + ((k) - 1)) * x->ip + ((i) - 1)))
#define XI2(i, j) ((unsigned)(((j) - 1) * x->ip + ((i) - 1)))
__device__ __forceinline__ float &elem(float *p, unsigned e) { return *reinterpret_cast<float *>(reinterpret_cast<char *>(p) + (e << 2)); }
__device__ __forceinline__ const float &elem(const float *p, unsigned e) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(p) + (e << 2));
}
// Buffer-resource loads of one 3-D field: a 128-bit descriptor (base and
// size in bytes, built in scalar registers from the field's pointer) and a
// 32-bit byte offset per lane.  The memory unit checks every offset against
// the size and returns 0 for an offset at or past it without an access; an
// offset below the base wraps to a large unsigned value and is out of range
// too.  Stencil windows near the domain edges are loaded unclamped: the
// slots a boundary-degraded flux does not use may fall outside the field.
struct BufF {
  __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ BufF bufv(const float *p, unsigned n) {
  // descriptor word 3: 32-bit data format, untyped access (gfx9 family)
  return BufF{__builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0, (int)(n << 2), 0x00020000)};
}
__device__ __forceinline__ float bld(const BufF &b, unsigned e) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, (int)(e << 2), 0, 0));
}
// Holds a loaded value in a register at this point (an empty asm statement
// that reads it): its load completes here, so the compiler cannot sink the
// load into a branch that uses the value later.  Placed after the last load
// of a group, it costs one wait for the whole group (and keeps later loads
// after it).
__device__ __forceinline__ void pin1(float v) { asm volatile("" : : "v"(v)); }
#define A3(fl, i, k, j) elem(fl, XI3(i, k, j))
#define A2(fl, i, j) elem(fl, XI2(i, j))
#define A1(fl, k) ((fl)[(k) - 1])
#define BX(fl, j, k, b) ((fl)[(unsigned)((((b) - 1) * x->kde + ((k) - 1)) * x->jde + ((j) - 1))])
#define BY(fl, i, k, b) ((fl)[(unsigned)((((b) - 1) * x->kde + ((k) - 1)) * x->ip + ((i) - 1))])
#define BX2(fl, j, b) ((fl)[(unsigned)(((b) - 1) * x->jde + ((j) - 1))])
#define BY2(fl, i, b) ((fl)[(unsigned)(((b) - 1) * x->ip + ((i) - 1))])
#define FL(name) (x->F[F_##name])
#define SP(name, s) (x->F[F_##name] + (size_t)(s) * x->n3)
#define SCR(id) (x->S[id])

// Inclusive Fortran index ranges of a launch.
struct R3 {
  int i0, i1, k0, k1, j0, j1;
};
struct R2 {
  int i0, i1, j0, j1;
};

// Workgroups are dispatched to the 8 XCDs round-robin by linear id, and each
// XCD has its own L2.  xcd_tile maps the linear id b of a grid of nb blocks
// to a tile number so that XCD x walks the contiguous tiles
// [x*q + min(x, r), ...) (q = nb / 8, r = nb % 8) in order: neighbouring rows
// of a stencil are then read by the same XCD shortly after one another and
// hit its L2 instead of being fetched again by another XCD.  Bijective for
// every nb.
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
  const unsigned xc = b & 7u, idx = b >> 3, q = nb >> 3, rm = nb & 7u;
  return xc * q + (xc < rm ? xc : rm) + idx;
}
// threadIdx.y of a block that is 64 threads wide: one value per wavefront.
// Read through readfirstlane so the compiler knows it is uniform across the
// wave (indices derived from it stay in scalar registers, 1-D coefficient
// loads are scalar loads and branches on them are scalar branches).
__device__ __forceinline__ unsigned wave_y() { return (unsigned)__builtin_amdgcn_readfirstlane((int)threadIdx.y); }
// Stencil kernels: thread -> (i, k, j) with blockDim (64, 4, 1) (a block is
// 64 i x 4 levels of one row) and a 1-D grid of
// ceil((i1 - ib + 1)/64) * ceil(nk/4) * 4 * ceil(nj/4) blocks.  The i tiles
// start at ib = 1 + 64 n <= i0 (lanes before i0 return): a wave's 64
// consecutive floats then sit on 256-byte boundaries of the row (the pitch
// is a multiple of 16 floats), two cache lines per load instead of three
// (the scalar diffusion passes run 15 % faster than with tiles from i0).  Tiles run i fastest,
// then 4 consecutive rows, then k, then groups of 4 rows, in XCD-contiguous
// order: the j+-1 rows of a stencil were read a few blocks earlier by the
// same XCD and are still in its L2.
#define GET_IKJ(r)                                                                           \
  const int ib_ = 1 + (((r).i0 - 1) & ~63);                                                  \
  const unsigned gx_ = (unsigned)((r).i1 - ib_ + 64) / 64u;                                  \
  const unsigned gy_ = (unsigned)((r).k1 - (r).k0 + 4) / 4u;                                 \
  const unsigned t_ = xcd_tile(blockIdx.x, gridDim.x);                                       \
  const unsigned u_ = t_ / gx_;                                                              \
  const int i = ib_ + (int)((t_ % gx_) * 64u + threadIdx.x);                                 \
  const int k = (r).k0 + (int)(((u_ / 4u) % gy_) * 4u + wave_y());                           \
  const int j = (r).j0 + (int)((u_ / 4u / gy_) * 4u + u_ % 4u);                              \
  if (i < (r).i0 || i > (r).i1 || k > (r).k1 || j > (r).j1) return
// Pointwise kernels (no neighbour in j): the same block shape and i tiles, a
// 1-D grid of ceil((i1 - ib + 1)/64) * ceil(nk/4) * nj blocks in plain order
// (i, then k, then j).
#define GET_IKJ_P(r)                                                                         \
  const int ib_ = 1 + (((r).i0 - 1) & ~63);                                                  \
  const unsigned gx_ = (unsigned)((r).i1 - ib_ + 64) / 64u;                                  \
  const unsigned gy_ = (unsigned)((r).k1 - (r).k0 + 4) / 4u;                                 \
  const int i = ib_ + (int)((blockIdx.x % gx_) * 64u + threadIdx.x);                         \
  const int k = (r).k0 + (int)(((blockIdx.x / gx_) % gy_) * 4u + wave_y());                  \
  const int j = (r).j0 + (int)(blockIdx.x / (gx_ * gy_));                                    \
  if (i < (r).i0 || i > (r).i1 || k > (r).k1 || j > (r).j1) return
// thread -> (i, j) for column kernels: blockDim (64, 1, 1), 1-D grid of
// ceil(ni/64) * nj blocks
#define GET_IJ(r)                                                                            \
  const unsigned gx_ = (unsigned)((r).i1 - (r).i0 + 64) / 64u;                               \
  const unsigned t_ = xcd_tile(blockIdx.x, gridDim.x);                                       \
  const int i = (r).i0 + (int)((t_ % gx_) * 64u + threadIdx.x);                              \
  const int j = (r).j0 + (int)(t_ / gx_);                                                    \
  if (i > (r).i1 || j > (r).j1) return

__device__ __forceinline__ float sgn1(float v) { return v >= 0.f ? 1.f : -1.f; }
// x / c correctly rounded (IEEE RN), for an integer constant c (12 or 60)
// with rc = RN(1/c).  Checked bitwise against x / c for every finite FP32 x,
// for c = 12 and c = 60.
//  - The quotient is x * rc with one FMA correction: t = x * rc * c - x
//    rounded once, then x * rc - t * rc rounded once.  Written with t (not
//    x - q * c), x = -0 gives -0 with no sign fix-up.
//  - That is exact wherever its result is not subnormal, and a subnormal
//    result (one class test) sends the lane to the rare path, which also
//    covers every x whose true quotient is subnormal.
//  - The rare path: a nonzero x below c * 2^-125 has X = |x| * 2^149, an
//    integer below 2^30, and a quotient n * 2^-149 with n = X / c rounded to
//    nearest, ties to even (n <= 2^24, so exact).  n comes from two rounded
//    estimates, each corrected by the exact remainder X - n * c (one FMA),
//    then the half-way cases; every constant is an inline operand or c / rc,
//    so the rare path takes no scalar registers of its own.
__device__ __forceinline__ float div_tiny(float xv, float c, float rc) {
  const float X = __builtin_ldexpf(__builtin_ldexpf(__builtin_ldexpf(fabsf(xv), 64), 64), 21);
  float n = __builtin_rintf(X * rc);
  float r = __builtin_fmaf(-n, c, X);
  n = n + __builtin_rintf(r * rc);
  r = __builtin_fmaf(-n, c, X);
  const float h = 0.5f * c;
  const bool odd = (n - 2.f * __builtin_rintf(0.5f * n)) != 0.f;
  if (r > h || (r == h && odd)) n = n + 1.f;
  else if (r < -h || (r == -h && odd)) n = n - 1.f;
  return copysignf(n * 0x1p-149f, xv);
}
__device__ __forceinline__ float div_const(float xv, float c, float rc) {
  const float q = xv * rc;
  const float t = __builtin_fmaf(q, c, -xv);
  float q2 = __builtin_fmaf(-t, rc, q);
  if (__builtin_expect(__builtin_isfpclass(q2, 0x0090), 0)) q2 = div_tiny(xv, c, rc);  // +-subnormal
  return q2;
}
__device__ __forceinline__ int imin_(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax_(int a, int b) { return a > b ? a : b; }
