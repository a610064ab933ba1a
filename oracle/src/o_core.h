/*
 * CPU oracle for the RK3 split-explicit dycore -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a loop-form restatement of the WRF-JAX reference dycore
 * (src/wrf_jax/dyn_em/ and src/wrf_jax/boundary/ and the stage
 * adapters in src/wrf_jax/runtime.py).  Every routine cites the reference
 * function it follows.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never does.
 *
 * Index convention: Fortran-style 1-based (i,k,j) exactly as the reference
 * loops, i in [1, ide], k in [1, kde], j in [1, jde] with ide = nx+1,
 * kde = nz+1, jde = ny+1 (staggered extents, no halo).  Storage is
 * [j][k][i] with i contiguous and an i pitch `ip` >= ide.
 *
 * `real` is float (FP32 build) or double (FP64 build, -DOREAL=double).
 */
#ifndef O_CORE_H
#define O_CORE_H

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifndef OREAL
#define OREAL float
#endif
typedef OREAL real;

/* kinds: '3' 3-D, '2' 2-D, '1' 1-D column, '4' species-major 3-D stack,
 * 'X'/'Y' lateral-boundary arrays of 3-D fields on x/y edges,
 * 'x'/'y' lateral-boundary arrays of 2-D fields on x/y edges. */
#define O_FIELDS(X)                                                          \
  X(u_1, '3') X(u_2, '3') X(v_1, '3') X(v_2, '3') X(w_1, '3') X(w_2, '3')    \
  X(t_1, '3') X(t_2, '3') X(ph_1, '3') X(ph_2, '3') X(p, '3') X(al, '3')     \
  X(pb, '3') X(phb, '3') X(alb, '3') X(ru, '3') X(rv, '3') X(rw, '3')        \
  X(ww, '3') X(php, '3') X(alt, '3') X(cqu, '3') X(cqv, '3') X(cqw, '3')     \
  X(c2a, '3') X(pm1, '3') X(a, '3') X(alpha, '3') X(gamma, '3')              \
  X(ru_tend, '3') X(rv_tend, '3') X(rw_tend, '3') X(t_tend, '3')             \
  X(ph_tend, '3') X(ru_tendf, '3') X(rv_tendf, '3') X(rw_tendf, '3')         \
  X(t_tendf, '3') X(ph_tendf, '3') X(u_save, '3') X(v_save, '3')             \
  X(w_save, '3') X(t_save, '3') X(ph_save, '3') X(ww_1, '3') X(t_2ave, '3')  \
  X(ru_m, '3') X(rv_m, '3') X(ww_m, '3') X(h_diabatic, '3')                 \
  X(advect_tend, '3') X(rho, '3') X(z, '3') X(rdz, '3') X(rdzw, '3')         \
  X(zx, '3') X(zy, '3') X(div, '3') X(defor11, '3') X(defor22, '3')          \
  X(defor33, '3') X(defor12, '3') X(defor13, '3') X(defor23, '3')            \
  X(xkmh, '3') X(xkmv, '3') X(xkhh, '3') X(xkhv, '3') X(bn2, '3')            \
  X(mu_1, '2') X(mu_2, '2') X(mub, '2') X(mut, '2') X(muu, '2') X(muv, '2')  \
  X(muts, '2') X(muus, '2') X(muvs, '2') X(muave, '2') X(mudf, '2')          \
  X(mu_tend, '2') X(mu_tendf, '2') X(mu_save, '2') X(ht, '2') X(msfux, '2')  \
  X(msfuy, '2') X(msfvx, '2') X(msfvx_inv, '2') X(msfvy, '2') X(msftx, '2')  \
  X(msfty, '2') X(f, '2') X(e, '2') X(sina, '2') X(cosa, '2') X(clat, '2')   \
  X(znw, '1') X(znu, '1') X(dnw, '1') X(rdnw, '1') X(dn, '1') X(rdn, '1')    \
  X(fnm, '1') X(fnp, '1') X(c1h, '1') X(c2h, '1') X(c1f, '1') X(c2f, '1')    \
  X(c3h, '1') X(c4h, '1') X(c3f, '1') X(c4f, '1') X(u_base, '1')             \
  X(v_base, '1') X(moist, 'M') X(moist_old, 'M') X(moist_tend, 'M')          \
  X(scalar, 'S') X(scalar_old, 'S') X(scalar_tend, 'S')                      \
  X(u_bxs, 'X') X(u_bxe, 'X') X(u_bys, 'Y') X(u_bye, 'Y')                    \
  X(u_btxs, 'X') X(u_btxe, 'X') X(u_btys, 'Y') X(u_btye, 'Y')                \
  X(v_bxs, 'X') X(v_bxe, 'X') X(v_bys, 'Y') X(v_bye, 'Y')                    \
  X(v_btxs, 'X') X(v_btxe, 'X') X(v_btys, 'Y') X(v_btye, 'Y')                \
  X(t_bxs, 'X') X(t_bxe, 'X') X(t_bys, 'Y') X(t_bye, 'Y')                    \
  X(t_btxs, 'X') X(t_btxe, 'X') X(t_btys, 'Y') X(t_btye, 'Y')                \
  X(ph_bxs, 'X') X(ph_bxe, 'X') X(ph_bys, 'Y') X(ph_bye, 'Y')                \
  X(ph_btxs, 'X') X(ph_btxe, 'X') X(ph_btys, 'Y') X(ph_btye, 'Y')            \
  X(qv_bxs, 'X') X(qv_bxe, 'X') X(qv_bys, 'Y') X(qv_bye, 'Y')                \
  X(qv_btxs, 'X') X(qv_btxe, 'X') X(qv_btys, 'Y') X(qv_btye, 'Y')            \
  X(mu_bxs, 'x') X(mu_bxe, 'x') X(mu_bys, 'y') X(mu_bye, 'y')                \
  X(mu_btxs, 'x') X(mu_btxe, 'x') X(mu_btys, 'y') X(mu_btye, 'y')

enum {
#define O_ENUM(name, kind) F_##name,
  O_FIELDS(O_ENUM)
#undef O_ENUM
  O_NFIELDS
};

/* Scalar configuration of one domain (namelist + derived constants). */
typedef struct {
  int nx, ny, nz;       /* mass points */
  int ip;               /* i pitch of every 3-D / 2-D array (>= nx+1) */
  int nmoist, nscalar;  /* species counts; moist species 0 is QV */
  int spec_bdy_width, spec_zone, relax_zone;
  int time_step_sound;  /* acoustic steps per dt (0 = WRF formula) */
  int pd_moist, pd_scalar; /* moist/scalar_adv_opt == 1 (positive definite) */
  int top_lid;
  int pad_;
  double dx, dy, dt;
  double ptop, p0, t0, g, r_d, r_v, cp, cpovcv, cvpm, rvovrd;
  double epssm, smdiv, emdiv, c_s, prandtl;
  double cf1, cf2, cf3, cfn, cfn1;
  double dtbc;          /* time since last boundary update (start of step) */
  double reradius;      /* 1 / earth radius used by curvature */
} OCfg;

typedef struct {
  OCfg c;
  int ide, jde, kde;
  size_t n3, n2;        /* elements of a 3-D / 2-D array */
  real *F[O_NFIELDS];
  /* per-step scratch owned by the oracle */
  real *fcx, *gcx;      /* relaxation weights, index 1..spec_bdy_width */
  int rk_step, nsmall, iteration;
  double dt_rk, dts_rk;
} OCtx;

#define IDE (x->ide)
#define JDE (x->jde)
#define KDE (x->kde)
#define IPITCH (x->c.ip)
#define I3(i, k, j) ((((size_t)(j) - 1) * (size_t)KDE + ((size_t)(k) - 1)) * (size_t)IPITCH + ((size_t)(i) - 1))
#define I2(i, j) (((size_t)(j) - 1) * (size_t)IPITCH + ((size_t)(i) - 1))
#define A3(fl, i, k, j) ((fl)[I3(i, k, j)])
#define A2(fl, i, j) ((fl)[I2(i, j)])
#define A1(fl, k) ((fl)[(k) - 1])
/* lateral-boundary arrays: x edges [b][k][j], y edges [b][k][i] */
#define BX(fl, j, k, b) ((fl)[(((size_t)(b) - 1) * (size_t)KDE + ((size_t)(k) - 1)) * (size_t)JDE + ((size_t)(j) - 1)])
#define BY(fl, i, k, b) ((fl)[(((size_t)(b) - 1) * (size_t)KDE + ((size_t)(k) - 1)) * (size_t)IPITCH + ((size_t)(i) - 1)])
#define BX2(fl, j, b) ((fl)[((size_t)(b) - 1) * (size_t)JDE + ((size_t)(j) - 1)])
#define BY2(fl, i, b) ((fl)[((size_t)(b) - 1) * (size_t)IPITCH + ((size_t)(i) - 1)])
#define FLD(name) (x->F[F_##name])
#define SPC(name, s) (x->F[F_##name] + (size_t)(s) * x->n3)

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline real rmax(real a, real b) { return a > b ? a : b; }
static inline real rmin(real a, real b) { return a < b ? a : b; }
static inline real sgn1(real v) { return v >= (real)0 ? (real)1 : (real)-1; }

#endif
