// MI355X (gfx950) RK3 split-explicit dycore: host orchestration and the C ABI
// declared in include/wrfdc.h.
//
// One wrfdc_ctx owns one domain's DevCtx (dimensions, scalars and the
// caller's field pointers), which every kernel receives by value so that the
// pointers it loads are known to be global, plus library scratch.  A model
// step is a fixed sequence of kernel launches on the context's stream; the
// stage order follows src/wrf_jax/core/direct_step.py:3037-3531 (RK and
// acoustic loops) and src/wrf_jax/dyn_em/solve_em.py:260-800 (final
// boundary stages).
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <array>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <utility>
#include <string>
#include <vector>

#include "wrfdc_common.h"
#include "k_advect.h"
#include "k_bc.h"
#include "k_bigstep.h"
#include "k_diff.h"
#include "k_small.h"
#include "k_acoustic.h"
#include "k_species.h"
#include "k_closure.h"
#include "k_adjoint.h"
#include "k_colsolve.h"

// ---------------------------------------------------------------------------
// small helper kernels local to the orchestration
// ---------------------------------------------------------------------------

// dst[c] = src[c] over the points of r, for every pair of the list in one
// launch (initialize_moist_old, q_old <- q, the species buffer copies).
// Block (64, 4): 256 consecutive i of 4 levels; grid (i segments x level
// groups, rows, pairs).  Each thread copies 4 consecutive i with one 16-byte
// load and store where all 4 lie in r and start on a 16-byte boundary (rows
// start 64-byte aligned), else element by element.
constexpr int kMaxCopy = 2 * kMaxSpec;
struct CopyList {
  float *dst[kMaxCopy];
  const float *src[kMaxCopy];
  int n;
};
__global__ void __launch_bounds__(256) k_copy_batch(const DevCtx xv, CopyList L, R3 r) {
  const DevCtx *__restrict__ x = &xv;
  const unsigned gi = (unsigned)(r.i1 - r.i0 + 256) / 256u;
  const int i0 = r.i0 + (int)(blockIdx.x % gi) * 256 + 4 * (int)threadIdx.x;
  const int k = r.k0 + (int)(blockIdx.x / gi) * 4 + (int)threadIdx.y;
  const int j = r.j0 + (int)blockIdx.y;
  if (k > r.k1 || i0 > r.i1) return;
  float *dst = L.dst[blockIdx.z];
  const float *src = L.src[blockIdx.z];
  const unsigned o = XI3(i0, k, j);
  if (i0 + 3 <= r.i1 && (o & 3u) == 0) {
    *reinterpret_cast<float4 *>(dst + o) = *reinterpret_cast<const float4 *>(src + o);
  } else {
    const int ie = imin_(i0 + 3, r.i1);
    for (int ii = i0; ii <= ie; ii++) elem(dst, o + (unsigned)(ii - i0)) = elem(src, o + (unsigned)(ii - i0));
  }
}

// FP64 checksums over mass points (WRFDC_NCHECKSUMS values):
// [0] sum(mu_2 + mub), [1] sum(t_2), [2] sum of all moist species,
// [3] count of non-finite prognostic values,
// [4] dry-air mass, sum of dm = (c1h*mu_t + c2h) * (-dnw) * dx*dy / (g*msftx*msfty)  (kg),
// [5] total energy, sum of dm * (kinetic + cv*T + geopotential)  (J), with the
//     winds averaged to the mass point, T = theta * ((pb+p)/p0)^(R/cp),
//     cv = R/(cp/cv - 1) and the geopotential averaged from the two w levels.
// Pass 1 writes one partial per block (vector stores, no atomics).
#define CK_BLOCK 256
#define CK_N 6
__global__ void __launch_bounds__(CK_BLOCK) k_checksum_partial(const DevCtx xv, double *part) {
  const DevCtx *__restrict__ x = &xv;
  const int ni = x->ide - 1, nk = x->kde - 1, nj = x->jde - 1;
  const size_t ncol = (size_t)ni * nj, npt = ncol * nk;
  const double cv = (double)x->r_d / ((double)x->cpovcv - 1.0);
  const double rcp = 1.0 - 1.0 / (double)x->cpovcv;  // R/cp
  const double area_g = (double)x->dx * (double)x->dy / (double)x->g;
  double s[CK_N] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < npt; t += (size_t)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(t % ni);
    const int k = 1 + (int)((t / ni) % nk);
    const int j = 1 + (int)(t / ((size_t)ni * nk));
    const size_t o = XI3(i, k, j);
    const float tv = FL(t_2)[o], uv = FL(u_2)[o], vv = FL(v_2)[o], wv = FL(w_2)[o], pv = FL(ph_2)[o];
    s[1] += (double)tv;
    float nf = 0.f;
    nf += isfinite(tv) ? 0.f : 1.f;
    nf += isfinite(uv) ? 0.f : 1.f;
    nf += isfinite(vv) ? 0.f : 1.f;
    nf += isfinite(wv) ? 0.f : 1.f;
    nf += isfinite(pv) ? 0.f : 1.f;
    for (int sp = 0; sp < x->nmoist; sp++) {
      const float q = SP(moist, sp)[o];
      s[2] += (double)q;
      nf += isfinite(q) ? 0.f : 1.f;
    }
    const double mt = (double)A2(FL(mu_2), i, j) + (double)A2(FL(mub), i, j);
    if (k == 1) {
      s[0] += mt;
      nf += isfinite(A2(FL(mu_2), i, j)) ? 0.f : 1.f;
    }
    s[3] += (double)nf;
    const double dm = ((double)A1(FL(c1h), k) * mt + (double)A1(FL(c2h), k)) * -(double)A1(FL(dnw), k) * area_g /
                      ((double)A2(FL(msftx), i, j) * (double)A2(FL(msfty), i, j));
    const double ub = 0.5 * ((double)uv + (double)A3(FL(u_2), i + 1, k, j));
    const double vb = 0.5 * ((double)vv + (double)A3(FL(v_2), i, k, j + 1));
    const double wb = 0.5 * ((double)wv + (double)A3(FL(w_2), i, k + 1, j));
    const double theta = (double)x->t0 + (double)tv;
    const double ptot = (double)A3(FL(pb), i, k, j) + (double)A3(FL(p), i, k, j);
    const double temp = theta * pow(ptot / (double)x->p0, rcp);
    const double phi = 0.5 * ((double)pv + (double)A3(FL(phb), i, k, j) + (double)A3(FL(ph_2), i, k + 1, j) +
                              (double)A3(FL(phb), i, k + 1, j));
    s[4] += dm;
    s[5] += dm * (0.5 * (ub * ub + vb * vb + wb * wb) + cv * temp + phi);
  }
  __shared__ double red[CK_N][CK_BLOCK];
  for (int c = 0; c < CK_N; c++) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int w = CK_BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < CK_N; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < CK_N) part[(size_t)blockIdx.x * CK_N + threadIdx.x] = red[threadIdx.x][0];
}
__global__ void __launch_bounds__(CK_BLOCK) k_checksum_final(const double *part, int nblocks, double *out) {
  __shared__ double red[CK_N][CK_BLOCK];
  double s[CK_N] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += CK_BLOCK)
    for (int c = 0; c < CK_N; c++) s[c] += part[(size_t)b * CK_N + c];
  for (int c = 0; c < CK_N; c++) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int w = CK_BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < CK_N; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < CK_N) out[threadIdx.x] = red[threadIdx.x][0];
}

// boundary times of the members of an ensemble, one launch before each replay
constexpr int kMaxMembers = 128;
struct DtbcList {
  float *p[kMaxMembers];
  float v[kMaxMembers];
  int n;
};
__global__ void k_put_dtbc(DtbcList L) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < L.n) *L.p[m] = L.v[m];
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static const char *kNames[] = {
#define WRFDC_NAME(name, kind) #name,
    WRFDC_FIELDS(WRFDC_NAME)
#undef WRFDC_NAME
};
static const char kKinds[] = {
#define WRFDC_KIND(name, kind) kind,
    WRFDC_FIELDS(WRFDC_KIND)
#undef WRFDC_KIND
};

static const int kCkBlocks = 512;

// Partial sums of wrfdc_checksums_noalloc: one fixed array per GPU (each
// device that loads the library holds its own copy until the process ends).
// Host code never passes this name to a kernel: it is a host-side stand-in.
// The device address comes from hipGetSymbolAddress, resolved once per
// device with that device current (noalloc_scratch below).
static __device__ __align__(256) double scratch[kCkBlocks * CK_N];

struct wrfdc_ctx {
  wrfdc_config c;
  int ide, jde, kde, ip, ns;
  size_t n3, n2;
  DevCtx h;
  hipStream_t s = nullptr;
  float *scratch = nullptr;
  float *spec_alt = nullptr;    // second buffer of every species (nmoist + nscalar 3-D arrays)
  float *spec_scale = nullptr;  // positive-definite renormalisation factors, one 3-D array per species
  float *hds = nullptr;         // horizontal scalar diffusion face fluxes, two 3-D arrays per diffused scalar
  double *ck = nullptr;  // kCkBlocks * CK_N partials + CK_N results
  // learned eddy viscosity (wrfdc_set_closure): parameters, bf16 weight
  // copies and feature normalisation; theta null = Smagorinsky
  const float *cl_theta = nullptr;
  const __bf16 *cl_wbf = nullptr;
  ClosureNorm cl_norm{};
  // per-RK-stage checkpoints of the dynamics the species transport reads
  // (wrfdc_record_stages): 5 device pointers per stage (ru_m, rv_m, ww_m,
  // mu_1, mu_2), filled in stage order until rec_n stages are recorded
  std::vector<float *> rec;
  int rec_n = 0, rec_i = 0;
  int rk = 1, nsmall = 1, iteration = 1;
  double dt_rk = 0.0, dts_rk = 0.0, dtbc = 0.0;
  float *dtbc_dev = nullptr;  // dtbc as the boundary kernels read it (one float)
  bool capturing = false;     // a step being captured into a graph leaves dtbc_dev to the replayer
  std::string err;
};

namespace {

struct HipError {
  hipError_t e;
  const char *what;
};

inline void hchk(hipError_t e, const char *what) {
  if (e != hipSuccess) throw HipError{e, what};
}

inline unsigned cdiv(int a, int b) { return (unsigned)((a + b - 1) / b); }

struct Run {
  wrfdc_ctx *X;
  // passed by value to every kernel: kernels read the field pointers from
  // kernel arguments.  The moist / scalar entries point at whichever species
  // buffer holds the current stage's values (the caller's or the library's
  // second buffer, see species_stage).
  mutable DevCtx d;
  hipStream_t s;
  int ide, jde, kde;
  float *user_moist, *user_scalar, *alt_moist, *alt_scalar;
  mutable bool old_synced = false;  // q_old == q on mass points for every species
  // within step(): the second stage's species update also does the last
  // stage's rk_update_scalar_pd of the PD species (pd_prepped records that it
  // did); a stage called on its own keeps the reference's order
  mutable bool fold_pd_prep = false, pd_prepped = false;
  explicit Run(wrfdc_ctx *x)
      : X(x), d(x->h), s(x->s), ide(x->ide), jde(x->jde), kde(x->kde), user_moist(x->h.F[F_moist]), user_scalar(x->h.F[F_scalar]),
        alt_moist(x->spec_alt), alt_scalar(x->spec_alt + x->n3 * (size_t)x->c.nmoist) {}

  float *F(int id) const { return d.F[id]; }
  float *Sp(int id, int s_) const { return d.F[id] + (size_t)s_ * X->n3; }
  float *Scr(int id) const { return X->h.S[id]; }

  R3 full3() const { return R3{1, ide, 1, kde, 1, jde}; }
  R3 mass3() const { return R3{1, ide - 1, 1, kde - 1, 1, jde - 1}; }
  R2 mass2() const { return R2{1, ide - 1, 1, jde - 1}; }
  // 1-D grids decoded by GET_IKJ (stencil kernels), GET_IKJ_P (pointwise
  // kernels) and GET_IJ
  // first i of the 64-aligned tiles of GET_IKJ / GET_IKJ_P
  static int ib(const R3 &r) { return 1 + ((r.i0 - 1) & ~63); }
  static dim3 g3(const R3 &r) { return dim3(cdiv(r.i1 - ib(r) + 1, 64) * cdiv(r.k1 - r.k0 + 1, 4) * 4u * cdiv(r.j1 - r.j0 + 1, 4)); }
  static dim3 g3p(const R3 &r) { return dim3(cdiv(r.i1 - ib(r) + 1, 64) * cdiv(r.k1 - r.k0 + 1, 4) * (unsigned)(r.j1 - r.j0 + 1)); }
  static dim3 g2(const R2 &r) { return dim3(cdiv(r.i1 - r.i0 + 1, 64) * (unsigned)(r.j1 - r.j0 + 1)); }
  static dim3 b3() { return dim3(64, 4, 1); }
  static dim3 bf() { return dim3(64, kFaceRows, 1); }  // face-sharing blocks

  void fill(float *p, size_t n, float v = 0.f) const {
    unsigned nb = (unsigned)((n + 255) / 256);
    if (nb > 8192) nb = 8192;
    k_fill<<<nb, 256, 0, s>>>(p, v, n);
  }
  void zero3(int id) const { fill(F(id), X->n3); }
  void zero2(int id) const { fill(F(id), X->n2); }

  // ---- batched physical boundary copies and fills ------------------------
  struct BcBatch {
    const Run *r;
    BcList L;
    int n = 0, kmax = 1;
    explicit BcBatch(const Run *run) : r(run) { memset(&L, 0, sizeof(L)); }
    BcBatch &add3(float *fl, char var) {
      const bool full_k = (var == 'e' || var == 'f' || var == 'w');
      return add(fl, full_k ? r->kde : r->kde - 1, var, false);
    }
    BcBatch &add2(float *fl, char var) { return add(fl, 1, var, true); }
    BcBatch &add(float *fl, int kend, char var, bool is2d) {
      if (n == kBcMax) flush();
      L.f[n] = fl;
      L.kend[n] = kend;
      L.flags[n] = ((var == 'u' || var == 'x') ? 0 : 1) | ((var == 'v' || var == 'y') ? 0 : 2) | (is2d ? 4 : 0);
      kmax = kend > kmax ? kend : kmax;
      n++;
      return *this;
    }
    void flush() {
      if (n == 0) return;
      const int along = r->ide > r->jde ? r->ide : r->jde;
      k_bc_batch<<<dim3(cdiv(along, 64), kmax, n), 64, 0, r->s>>>(r->d, L);
      n = 0;
      kmax = 1;
    }
    ~BcBatch() { flush(); }
  };
  void fill_many(std::initializer_list<std::pair<float *, size_t>> items) const {
    FillList L;
    memset(&L, 0, sizeof(L));
    int n = 0;
    size_t nmax = 0;
    for (auto &it : items) {
      if (it.second == 0) continue;
      L.p[n] = it.first;
      L.n[n] = it.second;
      nmax = it.second > nmax ? it.second : nmax;
      if (++n == kFillMax) {
        k_fill_batch<<<dim3((unsigned)std::min<size_t>((nmax + 255) / 256, 2048), n), 256, 0, s>>>(L);
        n = 0;
        nmax = 0;
      }
    }
    if (n) k_fill_batch<<<dim3((unsigned)std::min<size_t>((nmax + 255) / 256, 2048), n), 256, 0, s>>>(L);
  }

  // ---- physical boundary copies -----------------------------------------
  void bc3(float *fl, char var) const {
    const bool full_k = (var == 'e' || var == 'f' || var == 'w');
    const int kend = full_k ? kde : kde - 1;
    if (!(var == 'u' || var == 'x')) k_bc_x<<<dim3(cdiv(jde, 64), kend), 64, 0, s>>>(d, fl, kend, 0);
    if (!(var == 'v' || var == 'y')) k_bc_y<<<dim3(cdiv(ide, 64), kend), 64, 0, s>>>(d, fl, kend, 0);
  }
  void bc2(float *fl, char var) const {
    if (!(var == 'u' || var == 'x')) k_bc_x<<<dim3(cdiv(jde, 64), 1), 64, 0, s>>>(d, fl, 1, 1);
    if (!(var == 'v' || var == 'y')) k_bc_y<<<dim3(cdiv(ide, 64), 1), 64, 0, s>>>(d, fl, 1, 1);
  }

  // ---- specified / relaxation zones --------------------------------------
  struct Z {
    int ibe, jbe, kt;
  };
  Z zd(char var, bool ktop, bool is2d) const {
    Z z{ide - 1, jde - 1, ktop ? kde : kde - 1};
    if (var == 'u') z.ibe = ide;
    if (var == 'v') z.jbe = jde;
    if (is2d) z.kt = 1;
    return z;
  }
  static int zone_n(const Z &z) { return 2 * z.ibe + 2 * (z.jbe - 2); }
  BPtrs bp(int ys, int ye, int xs, int xe) const { return BPtrs{{F(ys), F(ye), F(xs), F(xe)}}; }

  void spec_overlay(float *tend, char var, bool ktop, bool is2d, BPtrs bt) const {
    Z z = zd(var, ktop, is2d);
    k_spec_overlay<<<dim3(cdiv(zone_n(z), 64), z.kt), 64, 0, s>>>(d, tend, z.ibe, z.jbe, z.kt, is2d ? 1 : 0, bt);
  }
  void spec_update(float *fl, const float *tend, double dt, char var, bool is2d) const {
    Z z = zd(var, var == 'm' || var == 'h', is2d);
    k_spec_update<<<dim3(cdiv(zone_n(z), 64), z.kt), 64, 0, s>>>(d, fl, tend, (float)dt, z.ibe, z.jbe, z.kt, is2d ? 1 : 0);
  }
  // relax_bdytend of fld (mw 0), or of its mass-weighted value (mw 1: half
  // levels, 2: w levels), formed by the relaxation kernel on the zone points
  // (RelaxBatch: several fields in one launch)
  struct RelaxBatch {
    const Run *r;
    RelaxList L;
    int n = 0, along = 1, kmax = 1;
    explicit RelaxBatch(const Run *run) : r(run) { memset(&L, 0, sizeof(L)); }
    RelaxBatch &add(const float *fld, float *tend, char var, bool ktop, bool is2d, BPtrs b, BPtrs bt, int mw = 0) {
      if (n == kRelaxMax) flush();
      const Z z = r->zd(var, ktop, is2d);
      L.fld[n] = fld;
      L.tend[n] = tend;
      L.b[n] = b;
      L.bt[n] = bt;
      L.ibe[n] = z.ibe;
      L.jbe[n] = z.jbe;
      L.kt[n] = z.kt;
      L.is2d[n] = is2d ? 1 : 0;
      L.mw[n] = mw;
      along = std::max(along, std::max(z.ibe, z.jbe));
      kmax = std::max(kmax, z.kt);
      n++;
      return *this;
    }
    void flush() {
      if (n == 0) return;
      const int rz = r->X->c.relax_zone;
      k_relax<<<dim3(cdiv(along, 64), kmax, n * 4 * (rz - 1)), 64, 0, r->s>>>(r->d, L, r->X->dtbc_dev);
      n = 0;
      along = kmax = 1;
    }
    ~RelaxBatch() { flush(); }
  };
  void relax(const float *fld, float *tend, char var, bool ktop, bool is2d, BPtrs b, BPtrs bt, int mw = 0) const {
    RelaxBatch(this).add(fld, tend, var, ktop, is2d, b, bt, mw);
  }
  // spec_bdy_final of several fields in one launch (the fields of one batch
  // must not read each other)
  struct FinalBatch {
    const Run *r;
    FinalList L;
    int n = 0, nmax = 1, kmax = 1;
    explicit FinalBatch(const Run *run) : r(run) { memset(&L, 0, sizeof(L)); }
    FinalBatch &add(float *fl, char var, bool is2d, const float *mu, const float *msf, const float *c1, const float *c2, BPtrs b,
                    BPtrs bt) {
      if (n == kFinalMax) flush();
      const Z z = r->zd(var, var == 'm' || var == 'h' || var == 'w', is2d);
      L.fl[n] = fl;
      L.mu[n] = mu;
      L.msf[n] = msf;
      L.c1[n] = c1;
      L.c2[n] = c2;
      L.b[n] = b;
      L.bt[n] = bt;
      L.ibe[n] = z.ibe;
      L.jbe[n] = z.jbe;
      L.kt[n] = z.kt;
      L.is2d[n] = is2d ? 1 : 0;
      L.msfc[n] = (var == 'u' || var == 'v' || var == 'w');
      L.muc[n] = (var != 'm');
      nmax = std::max(nmax, zone_n(z));
      kmax = std::max(kmax, z.kt);
      n++;
      return *this;
    }
    void flush() {
      if (n == 0) return;
      k_spec_final<<<dim3(cdiv(nmax, 64), kmax, n), 64, 0, r->s>>>(r->d, L, r->X->dtbc_dev);
      n = 0;
      nmax = kmax = 1;
    }
    ~FinalBatch() { flush(); }
  };
  void edge_copy(float *f0, float *f1, int axis, int dst, int src_, int kend) const {
    const int n = axis == 0 ? jde : ide;
    k_edge_copy<<<dim3(cdiv(n, 64), kend), 64, 0, s>>>(d, f0, f1, axis, dst, src_, kend);
  }

  // ---- RK-stage pieces -----------------------------------------------------
  void rk_step_prep() const {
    k_mu_uv<<<dim3(cdiv(ide, 64), jde), 64, 0, s>>>(d, F(F_mu_2), F(F_mub), F(F_muu), F(F_muv), F(F_mut));
    R3 r = full3();
    k_rk_prep3<<<g3(r), b3(), 0, s>>>(d, r);
    R2 c = mass2();
    k_ww_cp<<<g2(c), 64, sizeof(float) * 64 * (size_t)(kde - 1), s>>>(d, c);
  }
  void first_rk_step_part1() const {
    const size_t n3 = X->n3;
    fill_many({{F(F_ru_tendf), n3}, {F(F_rv_tendf), n3}, {F(F_rw_tendf), n3}, {F(F_ph_tendf), n3}, {F(F_t_tendf), n3},
               {F(F_mu_tendf), X->n2}, {F(F_moist_tend), n3 * (size_t)X->c.nmoist},
               {F(F_scalar_tend), n3 * (size_t)X->c.nscalar}});
    R3 m = mass3();
    k_rho<<<g3p(m), b3(), 0, s>>>(d, m);
  }
  void compute_diff_metrics() const {
    R3 r = full3();
    k_diff_metrics<<<g3(r), b3(), 0, s>>>(d, r);
  }
  void cal_deform_and_div() const {
    R3 m = mass3();
    k_deform_diag<<<g3(m), b3(), 0, s>>>(d, m);
    R3 r12{2, ide - 1, 1, kde - 1, 2, jde - 1};
    k_deform12<<<g3(r12), b3(), 0, s>>>(d, r12);
    edge_copy(F(F_defor12), nullptr, 0, 1, 2, kde);
    edge_copy(F(F_defor12), nullptr, 1, 1, 2, kde);
    edge_copy(F(F_defor12), nullptr, 0, ide, ide - 1, kde);
    edge_copy(F(F_defor12), nullptr, 1, jde, jde - 1, kde);
    R3 r13{2, ide - 1, 1, kde, 2, jde - 1};
    k_deform13_23<<<g3(r13), b3(), 0, s>>>(d, r13);
    edge_copy(F(F_defor13), F(F_defor23), 0, 1, 2, kde);
    edge_copy(F(F_defor13), F(F_defor23), 1, 1, 2, kde);
    edge_copy(F(F_defor13), F(F_defor23), 0, ide, ide - 1, kde);
    edge_copy(F(F_defor13), F(F_defor23), 1, jde, jde - 1, kde);
  }
  void smag2d_km() const {
    R3 r{2, ide - 2, 1, kde - 1, 2, jde - 2};
    k_smag<<<g3(r), b3(), 0, s>>>(d, r);
  }
  // N^2 of calculate_km_kh (read by no later stage of the km_opt 4 path
  // itself; the learned closure takes it as an input)
  void calculate_n2() const {
    R3 r{2, ide - 2, 1, kde - 1, 2, jde - 2};
    k_n2<<<g3p(r), b3(), 0, s>>>(d, r);
  }
  // Smagorinsky km storing |D| and the slope divisor in the scratch arrays
  // S_T1 and S_T2 (free until the diffusion's stresses), then the learned km
  // over the same points when the context has a network
  // (and the network's normalised inputs in S_FQX, S_FQY, free outside the
  // per-field positive-definite advection stage)
  void smag_for_closure(const ClosureNorm &nm) const {
    R3 r{2, ide - 2, 1, kde - 1, 2, jde - 2};
    k_smag<<<g3(r), b3(), 0, s>>>(d, r, Scr(S_T1), Scr(S_T2));
    k_closure_feat<<<g3p(r), b3(), 0, s>>>(d, r, nm, Scr(S_T1), Scr(S_FQX), Scr(S_FQY));
  }
  void closure_km() const {
    R3 r{2, ide - 2, 1, kde - 1, 2, jde - 2};
    static_assert(S_FQY == S_FQX + 1, "the network's two input arrays are adjacent scratch slots");
    k_closure_km<<<g3p(r), b3(), 0, s>>>(d, r, X->cl_theta, X->cl_wbf, X->cl_norm, Scr(S_FQX), Scr(S_T2));
  }
  // training rows of the current state (features, s1, s2, divisor, km)
  void closure_rows(const ClosureNorm &nm, __bf16 *Xr, float *T) const {
    compute_diff_metrics();
    BcBatch(this).add3(F(F_rdzw), 'w').add3(F(F_rdz), 'w').add3(F(F_z), 'w').add3(F(F_zx), 'e').add3(F(F_zy), 'f').flush();
    cal_deform_and_div();
    calculate_n2();
    smag_for_closure(nm);
    R3 r{2, ide - 2, 1, kde - 1, 2, jde - 2};
    k_closure_rows<<<g3p(r), b3(), 0, s>>>(d, r, nm, Scr(S_FQX), Scr(S_FQY), Scr(S_T2), Xr, T);
  }
  // two stress components (t1, t2 of one diffusion kernel) in one launch;
  // each entry: xk, defor, kind, i0, i1, j0, j1
  void titau(float *t1, float *t2, std::array<int, 7> a, std::array<int, 7> b) const {
    TitauPair P;
    P.tt[0] = t1;
    P.tt[1] = t2;
    const std::array<int, 7> *e[2] = {&a, &b};
    for (int n = 0; n < 2; n++) {
      P.xk[n] = F((*e[n])[0]);
      P.defor[n] = F((*e[n])[1]);
      P.kind[n] = (*e[n])[2];
      P.i0[n] = (*e[n])[3];
      P.i1[n] = (*e[n])[4];
      P.j0[n] = (*e[n])[5];
      P.j1[n] = (*e[n])[6];
    }
    R3 r = full3();
    k_titau<<<g3(r), b3(), 0, s>>>(d, P, r);
  }
  // horizontal_diffusion_s of theta and every species: faces, then tendencies
  void hdiff_scalars() const {
    HdsList L;
    memset(&L, 0, sizeof(L));
    auto add = [&](const float *v, float *t) {
      L.var[L.n] = v;
      L.tend[L.n] = t;
      L.h1[L.n] = X->hds + (size_t)(2 * L.n) * X->n3;
      L.h2[L.n] = X->hds + (size_t)(2 * L.n + 1) * X->n3;
      L.n++;
    };
    add(F(F_t_2), F(F_t_tendf));
    for (int q = 0; q < X->c.nmoist; q++) add(Sp(F_moist, q), Sp(F_moist_tend, q));
    for (int q = 0; q < X->c.nscalar; q++) add(Sp(F_scalar, q), Sp(F_scalar_tend, q));
    const R3 rf{2, ide - 1, 1, kde - 1, 2, jde - 1};
    k_hdiff_s_faces<<<g3(rf), b3(), 0, s>>>(d, L, rf);
    const R3 rt{2, ide - 2, 1, kde - 1, 2, jde - 2};
    k_hdiff_s_tend<<<g3(rt), b3(), 0, s>>>(d, L, rt);
  }
  void horizontal_diffusion_2() const {
    float *t1 = Scr(S_T1), *t2 = Scr(S_T2);
    titau(t1, t2, {F_xkmh, F_defor11, 11, 1, ide - 1, 2, jde - 1}, {F_xkmh, F_defor12, 12, 2, ide - 1, 2, jde});
    R3 ru{2, ide - 1, 1, kde - 1, 2, jde - 2};
    k_hdiff_u<<<g3(ru), b3(), 0, s>>>(d, t1, t2, ru);
    titau(t1, t2, {F_xkmh, F_defor12, 12, 2, ide, 2, jde - 1}, {F_xkmh, F_defor22, 11, 2, ide - 1, 1, jde - 1});
    R3 rv{2, ide - 2, 1, kde - 1, 2, jde - 1};
    k_hdiff_v<<<g3(rv), b3(), 0, s>>>(d, t1, t2, rv);
    titau(t1, t2, {F_xkmv, F_defor13, 13, 2, ide, 2, jde - 2}, {F_xkmv, F_defor23, 23, 2, ide - 2, 2, jde});
    R3 rw{2, ide - 2, 2, kde - 1, 2, jde - 2};
    k_hdiff_w<<<g3(rw), b3(), 0, s>>>(d, t1, t2, rw);
    hdiff_scalars();
  }
  // with_n2: also N^2 (the stage API always, the step when the learned
  // closure reads it)
  void first_rk_step_part2(bool with_n2) const {
    compute_diff_metrics();
    BcBatch(this).add3(F(F_rdzw), 'w').add3(F(F_rdz), 'w').add3(F(F_z), 'w').add3(F(F_zx), 'e').add3(F(F_zy), 'f').flush();
    cal_deform_and_div();
    if (with_n2 || X->cl_theta) calculate_n2();
    if (X->cl_theta) {
      smag_for_closure(X->cl_norm);
      closure_km();
    } else {
      smag2d_km();
    }
    BcBatch(this)
        .add3(F(F_xkmh), 't').add3(F(F_xkhh), 't').add3(F(F_xkmv), 't').add3(F(F_xkhv), 't').add3(F(F_div), 't')
        .add3(F(F_defor11), 't').add3(F(F_defor22), 't').add3(F(F_defor33), 't').add3(F(F_defor12), 'd')
        .add3(F(F_defor13), 'e').add3(F(F_defor23), 'f').add3(F(F_rho), 't').flush();
    horizontal_diffusion_2();
  }
  void advect_u() const {
    R3 r{2, ide - 1, 1, kde - 1, 1, jde - 1};
    k_advect_u<false><<<dim3(face_tiles(r, d.face_lg[0]).count()), bf(), 0, s>>>(d, F(F_u_2), F(F_ru_tend), F(F_ru), F(F_rv), F(F_ww), r, 0);
  }
  void advect_v() const {
    R3 r{1, ide - 1, 1, kde - 1, 2, jde - 1};
    k_advect_v<false><<<dim3(face_tiles(r, d.face_lg[0]).count()), bf(), 0, s>>>(d, F(F_v_2), F(F_rv_tend), F(F_ru), F(F_rv), F(F_ww), r, 0);
  }
  void advect_w() const {
    R3 r{1, ide - 1, 2, kde, 1, jde - 1};
    k_advect_w<false><<<dim3(face_tiles(r, d.face_lg[1]).count()), bf(), 0, s>>>(d, F(F_w_2), F(F_rw_tend), F(F_ru), F(F_rv), F(F_ww), r, 0);
  }
  void advect_scalar(const float *q, float *tend, const float *ru, const float *rv, const float *rom) const {
    R3 m = mass3();
    k_advect_scalar<false><<<dim3(face_tiles(m, d.face_lg[0]).count()), bf(), 0, s>>>(d, q, tend, ru, rv, rom, m, 0);
  }
  void advect_scalar_pd(const float *q, const float *q_old, float *tend, double dt) const {
    R3 r = full3();
    k_pd_fluxes<<<g3(r), b3(), 0, s>>>(d, q, q_old, F(F_ru_m), F(F_rv_m), F(F_ww_m), F(F_muts), (float)dt, r);
    k_pd_scale<<<g3(r), b3(), 0, s>>>(d, q_old, F(F_mub), F(F_mu_1), (float)dt, r);
    R3 m = mass3();
    k_pd_tend<<<g3(m), b3(), 0, s>>>(d, tend, m);
  }
  void rhs_ph() const {
    R3 r{1, ide - 1, 2, kde, 1, jde - 1};
    k_rhs_ph<<<g3(r), b3(), 0, s>>>(d, r);
  }
  void horizontal_pressure_gradient() const {
    R3 rv{1, ide - 1, 1, kde - 1, 2, jde - 1};
    k_hpg<<<g3(rv), b3(), 0, s>>>(d, 1, rv);
    R3 ru{2, ide - 1, 1, kde - 1, 1, jde - 1};
    k_hpg<<<g3(ru), b3(), 0, s>>>(d, 0, ru);
  }
  void pg_buoy_w() const {
    R2 c = mass2();
    k_pg_buoy_w<<<g2(c), 64, 0, s>>>(d, c);
  }
  void coriolis_curvature() const {
    R3 ru{2, ide - 1, 1, kde - 1, 1, jde - 1};
    k_cor_curv<<<g3(ru), b3(), 0, s>>>(d, 0, ru);
    R3 rv{1, ide - 1, 1, kde - 1, 2, jde - 1};
    k_cor_curv<<<g3(rv), b3(), 0, s>>>(d, 1, rv);
    R3 rw{1, ide - 1, 2, kde - 1, 1, jde - 1};
    k_cor_curv<<<g3(rw), b3(), 0, s>>>(d, 2, rw);
  }
  // the RK tendencies and the saved tendencies start from zero each stage.
  // The first stage clears every tendency and saved tendency.  Later stages
  // of the fused step (first = false) clear nothing: the fused tendency
  // kernels write every point of their ranges from 0 (the points outside
  // stay 0 from the first stage's clear), the zone overlay writes its
  // points, mu_tend is formed as 0 + mu_tendf by the theta kernel at rk > 1,
  // and the saved tendencies keep their values - nothing reads them before
  // small_step_prep writes them again over the same points, and the first
  // stage's addtend fold is their only earlier reader.
  void clear_tendencies(bool first) const {
    const size_t n3 = X->n3, n2 = X->n2;
    if (first)
      fill_many({{F(F_ru_tend), n3}, {F(F_rv_tend), n3}, {F(F_rw_tend), n3}, {F(F_t_tend), n3}, {F(F_ph_tend), n3},
                 {F(F_u_save), n3}, {F(F_v_save), n3}, {F(F_w_save), n3}, {F(F_ph_save), n3}, {F(F_t_save), n3},
                 {F(F_mu_tend), n2}, {F(F_mu_save), n2}});
  }
  void rk_tendency() const {
    clear_tendencies(true);
    advect_u();
    advect_v();
    advect_w();
    advect_scalar(F(F_t_2), F(F_t_tend), F(F_ru), F(F_rv), F(F_ww));
    rhs_ph();
    horizontal_pressure_gradient();
    pg_buoy_w();
    coriolis_curvature();
  }
  // rk_tendency + dry_boundary_tendency as the RK step runs them: the
  // lateral relaxation first (it only writes the physics tendencies, which the
  // tendency terms do not read), then one kernel per tendency that forms all
  // its terms in the reference's order and adds the physics tendency
  // (rk_addtend_dry), then the specified-zone overlay.
  void fused_tendency() const {
    const int rk = X->rk;
    clear_tendencies(rk == 1);
    if (rk == 1) lateral_relax();
    const R3 ru{1, ide, 1, kde - 1, 1, jde - 1};
    k_advect_u<true><<<dim3(face_tiles(ru, d.face_lg[0]).count()), bf(), 0, s>>>(d, F(F_u_2), F(F_ru_tend), F(F_ru), F(F_rv), F(F_ww), ru, rk);
    const R3 rv{1, ide - 1, 1, kde - 1, 1, jde};
    k_advect_v<true><<<dim3(face_tiles(rv, d.face_lg[0]).count()), bf(), 0, s>>>(d, F(F_v_2), F(F_rv_tend), F(F_ru), F(F_rv), F(F_ww), rv, rk);
    const R3 rw{1, ide - 1, 1, kde, 1, jde - 1};
    k_advect_w<true><<<dim3(face_tiles(rw, d.face_lg[1]).count()), bf(), 0, s>>>(d, F(F_w_2), F(F_rw_tend), F(F_ru), F(F_rv), F(F_ww), rw, rk);
    const R3 rq{1, ide - 1, kde - 1, kde - 1, 1, jde - 1};  // levels [2, kde-2]: k_advect_w<true>
    k_cqw_inv<<<g3p(rq), b3(), 0, s>>>(d, rq);
    const R3 m = mass3();
    k_advect_scalar<true><<<dim3(face_tiles(m, d.face_lg[0]).count()), bf(), 0, s>>>(d, F(F_t_2), F(F_t_tend), F(F_ru), F(F_rv), F(F_ww), m, rk);
    k_tend_ph<<<g3(rw), b3(), 0, s>>>(d, rw, rk);
    spec_overlays();
  }
  // relax_bdy_dry at the first stage (physics tendencies of u, v, ph, theta, mu)
  void lateral_relax() const {
    RelaxBatch(this)
        .add(F(F_ru), F(F_ru_tendf), 'u', false, false, bp(F_u_bys, F_u_bye, F_u_bxs, F_u_bxe), bp(F_u_btys, F_u_btye, F_u_btxs, F_u_btxe))
        .add(F(F_rv), F(F_rv_tendf), 'v', false, false, bp(F_v_bys, F_v_bye, F_v_bxs, F_v_bxe), bp(F_v_btys, F_v_btye, F_v_btxs, F_v_btxe))
        .add(F(F_ph_2), F(F_ph_tendf), 'h', true, false, bp(F_ph_bys, F_ph_bye, F_ph_bxs, F_ph_bxe),
             bp(F_ph_btys, F_ph_btye, F_ph_btxs, F_ph_btxe), 2)
        .add(F(F_t_2), F(F_t_tendf), 't', false, false, bp(F_t_bys, F_t_bye, F_t_bxs, F_t_bxe), bp(F_t_btys, F_t_btye, F_t_btxs, F_t_btxe), 1)
        .add(F(F_mu_2), F(F_mu_tend), 'm', true, true, bp(F_mu_bys, F_mu_bye, F_mu_bxs, F_mu_bxe),
             bp(F_mu_btys, F_mu_btye, F_mu_btxs, F_mu_btxe));
  }
  // spec_bdy_dry: the specified zone's tendencies
  // spec_bdytend of u, v, ph, theta and mu in one launch
  void spec_overlays() const {
    OverlayList L;
    memset(&L, 0, sizeof(L));
    int n = 0, nmax = 0, kmax = 1;
    auto add = [&](float *tend, char var, bool ktop, bool is2d, BPtrs bt) {
      const Z z = zd(var, ktop, is2d);
      L.tend[n] = tend;
      L.bt[n] = bt;
      L.ibe[n] = z.ibe;
      L.jbe[n] = z.jbe;
      L.kt[n] = z.kt;
      L.is2d[n] = is2d ? 1 : 0;
      nmax = zone_n(z) > nmax ? zone_n(z) : nmax;
      kmax = z.kt > kmax ? z.kt : kmax;
      n++;
    };
    add(F(F_ru_tend), 'u', false, false, bp(F_u_btys, F_u_btye, F_u_btxs, F_u_btxe));
    add(F(F_rv_tend), 'v', false, false, bp(F_v_btys, F_v_btye, F_v_btxs, F_v_btxe));
    add(F(F_ph_tend), 'h', true, false, bp(F_ph_btys, F_ph_btye, F_ph_btxs, F_ph_btxe));
    add(F(F_t_tend), 't', false, false, bp(F_t_btys, F_t_btye, F_t_btxs, F_t_btxe));
    add(F(F_mu_tend), 'm', true, true, bp(F_mu_btys, F_mu_btye, F_mu_btxs, F_mu_btxe));
    k_spec_overlay_batch<<<dim3(cdiv(nmax, 64), kmax, n), 64, 0, s>>>(d, L);
  }
  // rk_addtend_dry: the physics tendencies into the RK tendencies
  void rk_addtend_dry() const {
    const int rk = X->rk;
    R3 r0{1, ide, 1, kde - 1, 1, jde - 1};
    k_addtend<<<g3p(r0), b3(), 0, s>>>(d, 0, rk, r0);
    R3 r1{1, ide - 1, 1, kde - 1, 1, jde};
    k_addtend<<<g3p(r1), b3(), 0, s>>>(d, 1, rk, r1);
    R3 r2{1, ide - 1, 1, kde, 1, jde - 1};
    k_addtend<<<g3p(r2), b3(), 0, s>>>(d, 2, rk, r2);
    R3 r3 = mass3();
    k_addtend<<<g3p(r3), b3(), 0, s>>>(d, 3, rk, r3);
  }
  // relax_bdy_dry (first stage) + rk_addtend_dry + spec_bdy_dry
  void dry_boundary_tendency() const {
    if (X->rk == 1) lateral_relax();
    rk_addtend_dry();
    spec_overlays();
  }
  void small_step_prep() const {
    k_ssp_mass<<<dim3(cdiv(ide, 64), jde), 64, 0, s>>>(d, X->rk);
    R3 r = full3();
    if (X->rk == 1) k_ssp_3d<true><<<g3p(r), b3(), 0, s>>>(d, r);
    else k_ssp_3d<false><<<g3p(r), b3(), 0, s>>>(d, r);
  }
  void calc_p_rho(int step) const {
    R3 m = mass3();
    k_calc_p_rho<<<g3p(m), b3(), 0, s>>>(d, step, m);
  }
  void calc_coef_w() const {
    R2 c = mass2();
    k_calc_coef_w<<<g2(c), 64, 0, s>>>(d, (float)X->dts_rk, c);
  }
  void advance_uv() const {
    const int sz = X->c.spec_zone;
    R3 ru{1 + sz, ide - sz, 1, kde - 1, 1 + sz, jde - sz - 1};
    k_advance_uv<<<g3(ru), b3(), 0, s>>>(d, 0, (float)X->dts_rk, ru);
    R3 rv{1 + sz, ide - sz - 1, 1, kde - 1, 1 + sz, jde - sz};
    k_advance_uv<<<g3(rv), b3(), 0, s>>>(d, 1, (float)X->dts_rk, rv);
  }
  void advance_mu_t() const {
    R2 c{2, ide - 2, 2, jde - 2};
    k_advance_mu_t<<<g2(c), 64, 0, s>>>(d, (float)X->dts_rk, c);
  }
  void advance_w() const {
    R2 c{2, ide - 2, 2, jde - 2};
    k_advance_w<<<g2(c), 64, 0, s>>>(d, (float)X->dts_rk, Scr(S_HAT2), c);
  }
  void sumflux() const {
    R3 r = full3();
    k_sumflux<<<g3p(r), b3(), 0, s>>>(d, X->iteration == 1, X->iteration == X->nsmall, (float)X->nsmall, r);
  }
  void spec_bdyupdate_ph() const {
    Z z = zd('h', true, false);
    k_spec_update_ph<<<dim3(cdiv(zone_n(z), 64), kde), 64, 0, s>>>(d, (float)X->dts_rk);
  }
  void zero_grad_bdy_w() const {
    k_zero_grad_rows<<<dim3(cdiv(ide - 1, 64), kde, 2), 64, 0, s>>>(d, F(F_w_2));
    k_zero_grad_cols<<<dim3(cdiv(jde - 1, 64), kde, 2), 64, 0, s>>>(d, F(F_w_2));
  }
  // one fused acoustic substep: u/v, the column kernel, zero_grad_bdy of w;
  // all_outputs: also the intermediate t_2ave (a substep called on its own)
  void acoustic_substep(int it, bool all_outputs) const {
    X->iteration = it;
    const float dts = (float)X->dts_rk;
    R3 r{1, ide, 1, kde - 1, 1, jde};
    k_uv_substep<<<g3(r), b3(), 0, s>>>(d, dts, r);
    launch_acoustic_col(d, s, dts, it == 1, it == X->nsmall, (float)X->nsmall, all_outputs || it == X->nsmall);
    Z z = zd('w', true, false);
    k_zero_grad_w<<<dim3(cdiv(zone_n(z), 64), kde), 64, 0, s>>>(d);
  }
  // calc_mu_uv_1: the staggered column masses of muts
  void calc_mu_uv_1() const {
    k_mu_uv<<<dim3(cdiv(ide, 64), jde), 64, 0, s>>>(d, F(F_muts), nullptr, F(F_muus), F(F_muvs), nullptr);
  }
  void small_step_finish() const {
    calc_mu_uv_1();
    recouple();
  }
  // small_step_finish proper (module_small_step_em.py:1119-1332), without
  // the calc_mu_uv_1 that the RK stage runs before it
  void recouple() const {
    R3 r = full3();
    if (X->rk == 3) k_ssf<true><<<g3(r), b3(), 0, s>>>(d, (float)X->dts_rk, (float)X->nsmall, r);
    else k_ssf<false><<<g3(r), b3(), 0, s>>>(d, (float)X->dts_rk, (float)X->nsmall, r);
  }
  void calc_p_rho_phi() const {
    R3 m = mass3();
    k_calc_p_rho_phi<<<g3p(m), b3(), 0, s>>>(d, m);
  }
  void update_scalar_pd(float *q_old, float *q_tend) const {
    R3 m = mass3();
    k_update_scalar_pd<<<g3p(m), b3(), 0, s>>>(d, q_old, q_tend, (float)X->dt_rk, m);
  }
  void species_advance(float *q, float *q_old, float *q_tend, bool is_qv, bool is_moist, bool pd) const {
    float *adv = F(F_advect_tend);
    zero3(F_advect_tend);
    if (X->rk == 3 && pd) advect_scalar_pd(q, q_old, adv, X->dt_rk);
    else advect_scalar(q, adv, F(F_ru_m), F(F_rv_m), F(F_ww_m));
    R3 m = mass3();
    if (is_qv && X->rk == 1) {
      relax(q, q_tend, 'q', false, false, bp(F_qv_bys, F_qv_bye, F_qv_bxs, F_qv_bxe), bp(F_qv_btys, F_qv_btye, F_qv_btxs, F_qv_btxe), 1);
      spec_overlay(q_tend, 't', false, false, bp(F_qv_btys, F_qv_btye, F_qv_btxs, F_qv_btxe));
    }
    k_update_scalar<<<g3p(m), b3(), 0, s>>>(d, q, q_old, q_tend, X->rk, (float)X->dt_rk, m);
    if (is_moist && !is_qv) {
      k_flow_dep_rows<<<dim3(cdiv(ide - 1, 64), kde - 1, 2), 64, 0, s>>>(d, q, F(F_rv_m));
      k_flow_dep_cols<<<dim3(cdiv(jde - 1, 64), kde - 1, 2), 64, 0, s>>>(d, q, F(F_ru_m));
    }
  }

  // ---- species (all moist and scalar species of one stage) ----------------
  bool species_in_alt() const { return d.F[F_moist] != user_moist; }
  void use_species(bool alt) const {
    d.F[F_moist] = alt ? alt_moist : user_moist;
    d.F[F_scalar] = (X->c.nscalar > 0) ? (alt ? alt_scalar : user_scalar) : nullptr;
  }
  // copy the current species (every point a stage writes: i <= ide,
  // k <= kde-1, j <= jde) into the other buffer and switch to it
  void species_to(bool alt) const {
    if (species_in_alt() == alt) return;
    copy_species(R3{1, ide, 1, kde - 1, 1, jde}, alt);
    use_species(alt);
  }
  // one launch of k_copy_batch over r for the pairs of L
  void copy_batch(const CopyList &L, R3 r) const {
    if (L.n == 0) return;
    const unsigned gi = (unsigned)(r.i1 - r.i0 + 256) / 256u, gk = (unsigned)(r.k1 - r.k0 + 4) / 4u;
    k_copy_batch<<<dim3(gi * gk, (unsigned)(r.j1 - r.j0 + 1), (unsigned)L.n), dim3(64, 4), 0, s>>>(d, L, r);
  }
  void copy_species(R3 r, bool to_alt) const {
    const size_t n = X->n3;
    const float *sm = to_alt ? user_moist : alt_moist, *ss = to_alt ? user_scalar : alt_scalar;
    float *dm = to_alt ? alt_moist : user_moist, *ds = to_alt ? alt_scalar : user_scalar;
    CopyList L;
    memset(&L, 0, sizeof(L));
    for (int q = 0; q < X->c.nmoist; q++) {
      L.dst[L.n] = dm + q * n;
      L.src[L.n++] = sm + q * n;
    }
    for (int q = 0; q < X->c.nscalar; q++) {
      L.dst[L.n] = ds + q * n;
      L.src[L.n++] = ss + q * n;
    }
    copy_batch(L, r);
  }
  void fill_copy(float *dst, const float *src, size_t n) const {
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    k_copy<<<nb, 256, 0, s>>>(dst, src, n);
  }
  // q_old <- q on mass points (the reference's rk1 update sets it; for the
  // moist species the step start has already done so)
  void sync_old(bool moist, bool scalar) const {
    CopyList L;
    memset(&L, 0, sizeof(L));
    if (moist)
      for (int q = 0; q < X->c.nmoist; q++) {
        L.dst[L.n] = Sp(F_moist_old, q);
        L.src[L.n++] = Sp(F_moist, q);
      }
    if (scalar)
      for (int q = 0; q < X->c.nscalar; q++) {
        L.dst[L.n] = Sp(F_scalar_old, q);
        L.src[L.n++] = Sp(F_scalar, q);
      }
    copy_batch(L, mass3());
  }
  // Buffers: rk1 reads the stencils from q_old (== q) and writes q in place;
  // rk2 reads q and writes the second buffer, which later kernels of the step
  // then see as the species; rk3 reads that buffer and writes the caller's.
  void species_stage() const {
    const int rk = X->rk, nm = X->c.nmoist, nsc = X->c.nscalar;
    const float dt = (float)X->dt_rk;
    R3 m = mass3();
    if (rk == 1) {
      if (species_in_alt()) species_to(false);
      if (!old_synced) sync_old(true, true);
      old_synced = true;
      // relax_bdy_scalar + spec_bdy_scalar of QV into its tendency
      relax(Sp(F_moist, 0), Sp(F_moist_tend, 0), 'q', false, false, bp(F_qv_bys, F_qv_bye, F_qv_bxs, F_qv_bxe),
            bp(F_qv_btys, F_qv_btye, F_qv_btxs, F_qv_btxe), 1);
      spec_overlay(Sp(F_moist_tend, 0), 't', false, false, bp(F_qv_btys, F_qv_btye, F_qv_btxs, F_qv_btxe));
    }
    if (rk == 3 && !species_in_alt()) species_to(true);
    const bool out_alt = (rk == 2);
    SpecList plain, pd;
    memset(&plain, 0, sizeof(plain));
    memset(&pd, 0, sizeof(pd));
    const size_t n3 = X->n3;
    auto add = [&](SpecList &L, int id, int id_old, int id_tend, float *outbase, int q) {
      L.in[L.n] = (rk == 1) ? Sp(id_old, q) : Sp(id, q);
      L.out[L.n] = outbase + (size_t)q * n3;
      L.old[L.n] = Sp(id_old, q);
      L.tend[L.n] = Sp(id_tend, q);
      L.scale[L.n] = nullptr;
      L.n++;
    };
    const bool pdm = (rk == 3 && X->c.pd_moist), pds = (rk == 3 && X->c.pd_scalar);
    for (int q = 0; q < nm; q++) add(pdm ? pd : plain, F_moist, F_moist_old, F_moist_tend, out_alt ? alt_moist : user_moist, q);
    for (int q = 0; q < nsc; q++) add(pds ? pd : plain, F_scalar, F_scalar_old, F_scalar_tend, out_alt ? alt_scalar : user_scalar, q);
    if (rk == 2 && fold_pd_prep) {
      // the species of the last stage's PD list, in the list's order
      for (int q = 0; q < nm; q++)
        if (X->c.pd_moist) plain.prep_mask |= 1u << q;
      for (int q = 0; q < nsc; q++)
        if (X->c.pd_scalar) plain.prep_mask |= 1u << (nm + q);
      plain.prep_dt = (float)X->c.dt;  // the last stage's dt_rk
      pd_prepped = plain.prep_mask != 0;
    }
    if (pd.n > 0) {
      for (int q = 0; q < pd.n; q++) {
        pd.scale[q] = X->spec_scale + (size_t)q * n3;
      }
      if (!pd_prepped) k_species_pd_prep<<<g3p(m), b3(), 0, s>>>(d, pd, dt, m);
      {
        BcBatch bc(this);
        for (int q = 0; q < pd.n; q++) bc.add3(pd.old[q], 'p');
      }
      R3 r = full3();
      k_species_pd_scale<<<dim3(face_tiles(r, d.face_lg[2]).count()), bf(), sizeof(float) * 2 * 64 * kFaceRows * pd.n, s>>>(d, pd, dt, r);
      k_species_pd_update<<<dim3(face_tiles(m, d.face_lg[2]).count()), bf(), sizeof(float) * 2 * 64 * kFaceRows * pd.n, s>>>(d, pd, dt, m);
    }
    if (plain.n > 0) k_species_update<<<dim3(face_tiles(m, d.face_lg[0]).count()), bf(), sizeof(float) * 64 * kFaceRows * plain.n, s>>>(d, plain, rk, dt, m);
    // outputs become the current species
    use_species(out_alt);
    if (rk == 3) {
      // the physical-boundary ring (column ide, row jde) is last written by
      // the second stage's copies, which went to the other buffer
      copy_species(R3{ide, ide, 1, kde - 1, 1, jde}, false);
      copy_species(R3{1, ide, 1, kde - 1, jde, jde}, false);
    }
    if (nm > 1) {
      FlowList fl;
      memset(&fl, 0, sizeof(fl));
      for (int q = 1; q < nm; q++) fl.q[fl.n++] = Sp(F_moist, q);
      Z z = zd('t', false, false);
      k_flow_dep_batch<<<dim3(cdiv(zone_n(z), 64), kde - 1, fl.n), 64, 0, s>>>(d, fl);
    }
  }

  // one stage of scalar transport with recorded dynamics (tangent-linear =
  // forward) and its adjoint; rk sets dt_rk (set_rk before)
  void transport_stage(const float *const *st, const float *q_in, const float *q_old, float *q_out) const {
    float *tend = Scr(S_T1);
    R3 m = mass3();
    fill(tend, X->n3);
    advect_scalar(q_in, tend, st[0], st[1], st[2]);
    k_tl_update<<<g3p(m), b3(), 0, s>>>(d, m, StageMass{st[3], st[4]}, q_old, tend, q_out, (float)X->dt_rk);
  }
  void transport_stage_adjoint(const float *const *st, const float *qbar_out, float *qbar_old, float *qbar_in) const {
    float *tbar = Scr(S_T1);
    R3 m = mass3();
    k_ad_update<<<g3p(m), b3(), 0, s>>>(d, m, StageMass{st[3], st[4]}, qbar_out, qbar_old, tbar, (float)X->dt_rk);
    k_ad_advect<<<g3p(m), b3(), 0, s>>>(d, m, tbar, st[0], st[1], st[2], qbar_in);
  }

  // the mass fluxes and column masses of this stage's species update
  void record_stage() const {
    if (X->rec_i >= X->rec_n) return;
    float *const *dst = X->rec.data() + (size_t)5 * X->rec_i;
    const size_t b3 = sizeof(float) * X->n3, b2 = sizeof(float) * X->n2;
    hchk(hipMemcpyAsync(dst[0], F(F_ru_m), b3, hipMemcpyDeviceToDevice, s), "record ru_m");
    hchk(hipMemcpyAsync(dst[1], F(F_rv_m), b3, hipMemcpyDeviceToDevice, s), "record rv_m");
    hchk(hipMemcpyAsync(dst[2], F(F_ww_m), b3, hipMemcpyDeviceToDevice, s), "record ww_m");
    hchk(hipMemcpyAsync(dst[3], F(F_mu_1), b2, hipMemcpyDeviceToDevice, s), "record mu_1");
    hchk(hipMemcpyAsync(dst[4], F(F_mu_2), b2, hipMemcpyDeviceToDevice, s), "record mu_2");
    X->rec_i++;
  }

  // ---- physical-boundary groups and spec_bdy_final of named fields --------
  // rk_phys_bc_dry_1 (module_bc_em.py:109-216)
  BcBatch &phys_bc_dry_1(BcBatch &b) const {
    return b.add3(F(F_ru), 'u').add3(F(F_rv), 'v').add3(F(F_rw), 'w').add3(F(F_ww), 'w').add3(F(F_php), 'w')
        .add3(F(F_alt), 't').add3(F(F_p), 'p').add2(F(F_muu), 'u').add2(F(F_muv), 'v').add2(F(F_mut), 't');
  }
  // rk_phys_bc_dry_2 (module_bc_em.py:219-318)
  BcBatch &phys_bc_dry_2(BcBatch &b) const {
    return b.add3(F(F_u_2), 'u').add3(F(F_v_2), 'v').add3(F(F_w_2), 'w').add3(F(F_t_2), 'p').add3(F(F_ph_2), 'w')
        .add2(F(F_mu_2), 't');
  }
  // set_phys_bc_dry_2 (module_bc_em.py:321-445)
  BcBatch &set_phys_bc_dry_2(BcBatch &b) const {
    return b.add3(F(F_u_1), 'u').add3(F(F_u_2), 'u').add3(F(F_v_1), 'v').add3(F(F_v_2), 'v').add3(F(F_w_1), 'w')
        .add3(F(F_w_2), 'w').add3(F(F_t_1), 'p').add3(F(F_t_2), 'p').add3(F(F_ph_1), 'w').add3(F(F_ph_2), 'w')
        .add2(F(F_mu_1), 't').add2(F(F_mu_2), 't');
  }
  // spec_bdy_final of u, v, t, ph, mu or QV ('u' 'v' 't' 'h' 'm' 'q'):
  // the specified zone from the boundary data at dtbc, mass-coupled with
  // the stage's column masses and map factors (module_bc_em.py:1009-1124)
  FinalBatch &spec_final(FinalBatch &b, char which) const {
    switch (which) {
      case 'u':
        return b.add(F(F_u_2), 'u', false, F(F_muus), F(F_msfuy), F(F_c1h), F(F_c2h), bp(F_u_bys, F_u_bye, F_u_bxs, F_u_bxe),
                     bp(F_u_btys, F_u_btye, F_u_btxs, F_u_btxe));
      case 'v':
        return b.add(F(F_v_2), 'v', false, F(F_muvs), F(F_msfvx), F(F_c1h), F(F_c2h), bp(F_v_bys, F_v_bye, F_v_bxs, F_v_bxe),
                     bp(F_v_btys, F_v_btye, F_v_btxs, F_v_btxe));
      case 't':
        return b.add(F(F_t_2), 't', false, F(F_muts), F(F_msfty), F(F_c1h), F(F_c2h), bp(F_t_bys, F_t_bye, F_t_bxs, F_t_bxe),
                     bp(F_t_btys, F_t_btye, F_t_btxs, F_t_btxe));
      case 'h':
        return b.add(F(F_ph_2), 'h', false, F(F_muts), F(F_msfty), F(F_c1f), F(F_c2f), bp(F_ph_bys, F_ph_bye, F_ph_bxs, F_ph_bxe),
                     bp(F_ph_btys, F_ph_btye, F_ph_btxs, F_ph_btxe));
      case 'm':
        return b.add(F(F_mu_2), 'm', true, F(F_muts), F(F_msfty), F(F_c1h), F(F_c2h), bp(F_mu_bys, F_mu_bye, F_mu_bxs, F_mu_bxe),
                     bp(F_mu_btys, F_mu_btye, F_mu_btxs, F_mu_btxe));
      default:
        return b.add(Sp(F_moist, 0), 't', false, F(F_muts), F(F_msfty), F(F_c1h), F(F_c2h), bp(F_qv_bys, F_qv_bye, F_qv_bxs, F_qv_bxe),
                     bp(F_qv_btys, F_qv_btye, F_qv_btxs, F_qv_btxe));
    }
  }
  void set_w_surface() const { k_set_w_surface<<<dim3(cdiv(ide - 1, 64), jde - 1), 64, 0, s>>>(d); }

  void rk_stage() const {
    rk_step_prep();
    {
      BcBatch b(this);
      phys_bc_dry_1(b).add3(F(F_al), 'p').add3(F(F_ph_2), 'w');
    }
    if (X->rk == 1) {
      first_rk_step_part1();
      first_rk_step_part2(false);
    }
    fused_tendency();
    small_step_prep();
    calc_p_rho(0);
    calc_coef_w();
    BcBatch(this)
        .add3(F(F_ru_tend), 'u').add3(F(F_rv_tend), 'v').add3(F(F_ph_2), 'w').add3(F(F_al), 'p').add3(F(F_p), 'p')
        .add3(F(F_t_1), 'p').add3(F(F_t_save), 't').add2(F(F_mu_1), 't').add2(F(F_mu_2), 't').add2(F(F_mudf), 't')
        .flush();
    for (int it = 1; it <= X->nsmall; it++) acoustic_substep(it, false);
    small_step_finish();
    if (X->rk == 3)
      BcBatch(this).add3(F(F_ru_m), 'u').add3(F(F_rv_m), 'v').add3(F(F_ww_m), 'w').add2(F(F_mut), 't').add2(F(F_muts), 't').flush();
    const int nm = X->c.nmoist, nsc = X->c.nscalar;
    record_stage();
    species_stage();
    calc_p_rho_phi();
    if (X->rk < 3) {
      BcBatch bc(this);
      phys_bc_dry_2(bc);
      for (int q = 0; q < nm; q++) bc.add3(Sp(F_moist, q), 'p');
      for (int q = 0; q < nsc; q++) bc.add3(Sp(F_scalar, q), 'p');
    }
  }

  // the boundary time the kernels of this step or stage read
  void put_dtbc() const {
    if (!X->capturing) k_fill<<<1, 64, 0, s>>>(X->dtbc_dev, (float)X->dtbc, 1);
  }

  void step() const {
    put_dtbc();
    if (species_in_alt()) species_to(false);
    // initialize_moist_old (module_big_step_utilities_em.py:935-975); the
    // scalars' q_old <- q of the first-stage update is done here as well
    // (one copy launch), then the moist q_old's boundary update
    sync_old(true, true);
    {
      BcBatch bc(this);
      for (int q = 0; q < X->c.nmoist; q++) bc.add3(Sp(F_moist_old, q), 'p');
    }
    old_synced = true;
    struct FoldScope {  // cleared however the stages end
      const Run *r;
      explicit FoldScope(const Run *run) : r(run) { r->fold_pd_prep = true, r->pd_prepped = false; }
      ~FoldScope() { r->fold_pd_prep = r->pd_prepped = false; }
    } fold(this);
    for (int rk = 1; rk <= 3; rk++) {
      set_rk(rk);
      rk_stage();
    }
    {
      BcBatch b(this);
      set_phys_bc_dry_2(b).add3(Sp(F_moist, 0), 'p');
    }
    // the six fields read only coupling masses and map factors, none of them
    // another's result, so one launch serves all
    {
      FinalBatch fb(this);
      for (char w : {'u', 'v', 't', 'h', 'm', 'q'}) spec_final(fb, w);
    }
    set_w_surface();
  }

  // _rk_step_sizes (solve_em.py:243-257), rk_order = 3
  void set_rk(int rk) const {
    const double dt = X->c.dt;
    const int ns = X->ns;
    X->rk = rk;
    if (rk == 1) {
      X->dt_rk = dt / 3.0;
      X->dts_rk = X->dt_rk;
      X->nsmall = 1;
    } else if (rk == 2) {
      X->dt_rk = 0.5 * dt;
      X->dts_rk = dt / (double)ns;
      X->nsmall = ns / 2 > 1 ? ns / 2 : 1;
    } else {
      X->dt_rk = dt;
      X->dts_rk = dt / (double)ns;
      X->nsmall = ns;
    }
  }
};

// WRF's acoustic step count (solve_em.py:220-240)
int num_sound_steps(const wrfdc_config &c) {
  if (c.time_step_sound != 0) return c.time_step_sound;
  const double sp = c.dx < c.dy ? c.dx : c.dy;
  const int n = 2 * ((int)(300.0 * c.dt / sp - 0.01) + 1);
  return n > 4 ? n : 4;
}

// message of the last failure of an entry point that has no context
// (wrfdc_create, the context-free entry points), per host thread
thread_local std::string g_thread_err;

int set_err(wrfdc_ctx *x, int code, const std::string &msg) {
  if (x) x->err = msg;
  else g_thread_err = msg;
  return code;
}

int validate(const wrfdc_config &c, std::string &msg) {
  if (c.nx < 12 || c.ny < 12 || c.nz < 6) { msg = "domain too small: need nx, ny >= 12 and nz >= 6"; return WRFDC_EINVAL; }
  if (c.ip < c.nx + 1) { msg = "ip must be >= nx + 1"; return WRFDC_EINVAL; }
  if ((double)(c.ny + 1) * (c.nz + 1) * c.ip >= 1073741824.0) { msg = "a 3-D field must hold fewer than 2^30 elements"; return WRFDC_EINVAL; }
  if (c.nmoist < 1) { msg = "nmoist must be >= 1 (species 0 is QV)"; return WRFDC_EINVAL; }
  if (c.nscalar < 0) { msg = "nscalar must be >= 0"; return WRFDC_EINVAL; }
  if (c.nmoist + c.nscalar > kMaxSpec) { msg = "at most 16 advected species (nmoist + nscalar)"; return WRFDC_EINVAL; }
  if (!(c.dx > 0.0 && c.dy > 0.0 && c.dt > 0.0)) { msg = "dx, dy and dt must be positive"; return WRFDC_EINVAL; }
  if (c.spec_zone != 1) { msg = "only spec_zone = 1 is implemented"; return WRFDC_ENOTIMPL; }
  if (c.relax_zone < 2) { msg = "relax_zone must be >= 2"; return WRFDC_EINVAL; }
  if (c.spec_bdy_width != c.spec_zone + c.relax_zone || c.spec_bdy_width > 14) {
    msg = "spec_bdy_width must equal spec_zone + relax_zone (at most 14)";
    return WRFDC_EINVAL;
  }
  if (c.top_lid != 0) { msg = "top_lid is not implemented"; return WRFDC_ENOTIMPL; }
  if (c.nz > 150) { msg = "at most 150 vertical levels (LDS-resident columns)"; return WRFDC_EINVAL; }
  if (c.time_step_sound < 0) { msg = "time_step_sound must be >= 0"; return WRFDC_EINVAL; }
  return WRFDC_OK;
}

// Device address of `scratch` on the current device: resolved with
// hipGetSymbolAddress on a device's first call and kept per device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

int noalloc_scratch(double **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_err(nullptr, WRFDC_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return set_err(nullptr, WRFDC_EINVAL, "device index beyond " + std::to_string(kMaxDevices));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(nullptr, WRFDC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[dev];
  return WRFDC_OK;
}

// Dimensions, scalars and relaxation weights of a DevCtx (no pointers).
void fill_devctx(const wrfdc_config &c, DevCtx &h) {
  const wrfdc_config *cfg = &c;
  memset(&h, 0, sizeof(h));
  h.ide = cfg->nx + 1;
  h.jde = cfg->ny + 1;
  h.kde = cfg->nz + 1;
  h.ip = cfg->ip;
  h.nmoist = cfg->nmoist;
  h.nscalar = cfg->nscalar;
  h.spec_zone = cfg->spec_zone;
  h.relax_zone = cfg->relax_zone;
  h.spec_bdy_width = cfg->spec_bdy_width;
  h.top_lid = cfg->top_lid;
  h.n3 = (size_t)h.jde * h.kde * h.ip;
  h.n2 = (size_t)h.jde * h.ip;
  h.rdx = (float)(1.0 / cfg->dx);
  h.rdy = (float)(1.0 / cfg->dy);
  h.dx = (float)cfg->dx;
  h.dy = (float)cfg->dy;
  h.dt = (float)cfg->dt;
  h.g = (float)cfg->g;
  h.t0 = (float)cfg->t0;
  h.p0 = (float)cfg->p0;
  h.ptop = (float)cfg->ptop;
  h.r_d = (float)cfg->r_d;
  h.cpovcv = (float)cfg->cpovcv;
  h.rvovrd = (float)cfg->rvovrd;
  h.epssm = (float)cfg->epssm;
  h.smdiv = (float)cfg->smdiv;
  h.emdiv = (float)cfg->emdiv;
  h.c_s = (float)cfg->c_s;
  h.prandtl = (float)cfg->prandtl;
  h.cf1 = (float)cfg->cf1;
  h.cf2 = (float)cfg->cf2;
  h.cf3 = (float)cfg->cf3;
  h.cfn = (float)cfg->cfn;
  h.cfn1 = (float)cfg->cfn1;
  h.reradius = (float)cfg->reradius;
  h.face_lg[0] = (int)kFaceLG;
  h.face_lg[1] = (int)kFaceLGW;
  h.face_lg[2] = (int)kFaceLGPd;
  // lbc_fcx_gcx (boundary/lbc_fcx_gcx.py:10-56), spec_exp = 0, in FP32
  for (int b = cfg->spec_zone + 1; b <= cfg->spec_zone + cfg->relax_zone && b <= cfg->spec_bdy_width; b++) {
    const float dist = (float)(cfg->spec_zone + cfg->relax_zone - b);
    const float den = (float)(cfg->relax_zone - 1);
    h.fcx[b] = ((0.1f / (float)cfg->dt) * dist) / den;
    h.gcx[b] = (((1.0f / (float)cfg->dt) / 50.0f) * dist) / den;
  }
}

int ensure_ready(wrfdc_ctx *x) {
  for (int f = 0; f < NFIELDS; f++) {
    if (x->h.F[f]) continue;
    if (kKinds[f] == 'S' && x->c.nscalar == 0) continue;
    return set_err(x, WRFDC_EUNBOUND, std::string("field not bound: ") + kNames[f]);
  }
  return WRFDC_OK;
}

template <class Fn>
int guarded(wrfdc_ctx *x, Fn fn) {
  if (!x) return WRFDC_EINVAL;
  try {
    int rc = ensure_ready(x);
    if (rc != WRFDC_OK) return rc;
    rc = fn();
    if (rc != WRFDC_OK) return rc;
    hchk(hipGetLastError(), "kernel launch");
    return WRFDC_OK;
  } catch (const HipError &e) {
    return set_err(x, WRFDC_EHIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int wrfdc_field_count(void) { return NFIELDS; }
const char *wrfdc_field_name(int index) { return (index >= 0 && index < NFIELDS) ? kNames[index] : nullptr; }
int wrfdc_field_kind(int index) { return (index >= 0 && index < NFIELDS) ? (int)kKinds[index] : 0; }

int wrfdc_create(const wrfdc_config *cfg, wrfdc_ctx **out) {
  if (!cfg || !out) return set_err(nullptr, WRFDC_EINVAL, "cfg and out must not be null");
  *out = nullptr;
  std::string msg;
  int rc = validate(*cfg, msg);
  if (rc != WRFDC_OK) return set_err(nullptr, rc, msg);
  wrfdc_ctx *x = new wrfdc_ctx();
  x->c = *cfg;
  x->ide = cfg->nx + 1;
  x->jde = cfg->ny + 1;
  x->kde = cfg->nz + 1;
  x->ip = cfg->ip;
  x->n3 = (size_t)x->jde * x->kde * x->ip;
  x->n2 = (size_t)x->jde * x->ip;
  x->ns = num_sound_steps(*cfg);
  x->dtbc = cfg->dtbc;
  fill_devctx(*cfg, x->h);
  try {
    hchk(hipMalloc((void **)&x->scratch, sizeof(float) * x->n3 * NSCRATCH), "hipMalloc scratch");
    hchk(hipMalloc((void **)&x->ck, sizeof(double) * (kCkBlocks * CK_N + CK_N)), "hipMalloc checksum");
    hchk(hipMalloc((void **)&x->dtbc_dev, sizeof(float) * 64), "hipMalloc dtbc");
    hchk(hipMalloc((void **)&x->spec_alt, sizeof(float) * x->n3 * (size_t)(cfg->nmoist + cfg->nscalar)), "hipMalloc species");
    hchk(hipMalloc((void **)&x->spec_scale, sizeof(float) * x->n3 * (size_t)(cfg->nmoist + cfg->nscalar)), "hipMalloc species scale");
    hchk(hipMalloc((void **)&x->hds, sizeof(float) * x->n3 * 2 * (size_t)(1 + cfg->nmoist + cfg->nscalar)), "hipMalloc diffusion fluxes");
  } catch (const HipError &e) {
    set_err(nullptr, WRFDC_EHIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
    if (x->scratch) (void)hipFree(x->scratch);
    if (x->ck) (void)hipFree(x->ck);
    if (x->dtbc_dev) (void)hipFree(x->dtbc_dev);
    if (x->spec_alt) (void)hipFree(x->spec_alt);
    if (x->spec_scale) (void)hipFree(x->spec_scale);
    if (x->hds) (void)hipFree(x->hds);
    delete x;
    return WRFDC_EHIP;
  }
  for (int i = 0; i < NSCRATCH; i++) x->h.S[i] = x->scratch + (size_t)i * x->n3;
  {
    hipError_t e = acoustic_col_allow_lds(x->kde);
    if (e != hipSuccess) {
      wrfdc_destroy(x);
      return set_err(nullptr, WRFDC_EHIP, "LDS size for nz=" + std::to_string(cfg->nz) + ": " + hipGetErrorString(e));
    }
  }
  *out = x;
  return WRFDC_OK;
}

void wrfdc_destroy(wrfdc_ctx *x) {
  if (!x) return;
  if (x->s) (void)hipStreamSynchronize(x->s);
  else (void)hipDeviceSynchronize();
  if (x->scratch) (void)hipFree(x->scratch);
  if (x->ck) (void)hipFree(x->ck);
  if (x->dtbc_dev) (void)hipFree(x->dtbc_dev);
  if (x->spec_alt) (void)hipFree(x->spec_alt);
  if (x->spec_scale) (void)hipFree(x->spec_scale);
  if (x->hds) (void)hipFree(x->hds);
  delete x;
}

int wrfdc_bind(wrfdc_ctx *x, const char *name, void *device_ptr) {
  if (!x || !name) return WRFDC_EINVAL;
  for (int f = 0; f < NFIELDS; f++)
    if (strcmp(kNames[f], name) == 0) {
      x->h.F[f] = (float *)device_ptr;
      return WRFDC_OK;
    }
  return set_err(x, WRFDC_EINVAL, std::string("unknown field: ") + name);
}

int wrfdc_set_stream(wrfdc_ctx *x, void *hip_stream) {
  if (!x) return WRFDC_EINVAL;
  x->s = (hipStream_t)hip_stream;
  return WRFDC_OK;
}

int wrfdc_step(wrfdc_ctx *x, int nsteps, double dtbc0) {
  if (nsteps < 0) return set_err(x, WRFDC_EINVAL, "nsteps must be >= 0");
  return guarded(x, [&]() {
    Run run(x);
    for (int n = 0; n < nsteps; n++) {
      x->dtbc = dtbc0 + (double)(n + 1) * x->c.dt;
      run.step();
    }
    return WRFDC_OK;
  });
}

int wrfdc_set_dtbc(wrfdc_ctx *x, double dtbc) {
  if (!x) return WRFDC_EINVAL;
  if (!std::isfinite(dtbc)) return set_err(x, WRFDC_EINVAL, "dtbc must be finite");
  x->dtbc = dtbc;
  return WRFDC_OK;
}

int wrfdc_set_tile(wrfdc_ctx *x, const char *name, int value) {
  if (!x || !name) return WRFDC_EINVAL;
  if (std::string(name) != "face_levels")
    return set_err(x, WRFDC_EINVAL, std::string("unknown tile setting: ") + name + " (known: face_levels)");
  if (value < 0 || value > 32) return set_err(x, WRFDC_EINVAL, "face_levels must be 0 (the defaults) or 1..32");
  x->h.face_lg[0] = value ? value : (int)kFaceLG;
  x->h.face_lg[1] = value ? value : (int)kFaceLGW;
  x->h.face_lg[2] = value ? value : (int)kFaceLGPd;
  return WRFDC_OK;
}

// ---- the stages of solve_em by the reference's callback names -------------
// (src/wrf_jax/dyn_em/solve_em.py:470-790, src/wrf_jax/runtime.py:16235-16301)

static int field_id(const std::string &name) {
  for (int f = 0; f < NFIELDS; f++)
    if (name == kNames[f]) return f;
  return -1;
}

// boundary kind of set_physical_bc3d:<field> (runtime.py:7421-7443): 'u' and
// 'v' staggered in x / y, 'w' on full levels, anything else a mass field
static bool bc3_kind(const std::string &f, char *kind) {
  static const struct {
    const char *name;
    char kind;
  } table[] = {{"rho", 'p'},     {"al", 'p'},      {"p", 'p'},       {"ph_2", 'w'}, {"t_1", 'p'},
               {"t_save", 't'},  {"ru_tend", 'u'}, {"rv_tend", 'v'}, {"ru_m", 'u'}, {"rv_m", 'v'},
               {"ww_m", 'w'},    {"h_diabatic", 't'}};
  for (auto &e : table)
    if (f == e.name) {
      *kind = e.kind;
      return true;
    }
  return false;
}

// set_physical_bc3d:<field>[:small] and set_physical_bc2d:<field>[:small];
// false: not such a stage.  moist / moist_old / scalar / scalar_old cover
// every species of the array, moist_qv the first.  tke_1 / tke_2: the
// supported closure (km_opt 4) carries no TKE, so the library holds none
// and the stage has nothing to update.
static bool named_bc(const Run &run, wrfdc_ctx *x, const std::string &n, std::string &err) {
  const std::string p3 = "set_physical_bc3d:", p2 = "set_physical_bc2d:";
  const bool is3 = n.compare(0, p3.size(), p3) == 0, is2 = n.compare(0, p2.size(), p2) == 0;
  if (!is3 && !is2) return false;
  std::string f = n.substr(p3.size());
  const size_t colon = f.find(':');
  if (colon != std::string::npos) {
    if (f.substr(colon) != ":small") {
      err = "unknown stage: " + n;
      return true;
    }
    f = f.substr(0, colon);
  }
  Run::BcBatch b(&run);
  if (is2) {
    static const char *const ok2[] = {"mu_1", "mu_2", "mudf", "muts", "mut"};
    for (const char *o : ok2)
      if (f == o) {
        b.add2(run.F(field_id(f)), 't');
        return true;
      }
    err = "set_physical_bc2d: no boundary kind for field " + f;
    return true;
  }
  if (f == "tke_1" || f == "tke_2") return true;
  const int nm = x->c.nmoist, nsc = x->c.nscalar;
  if (f == "moist" || f == "moist_old") {
    for (int q = 0; q < nm; q++) b.add3(run.Sp(f == "moist" ? F_moist : F_moist_old, q), 'p');
  } else if (f == "scalar" || f == "scalar_old") {
    for (int q = 0; q < nsc; q++) b.add3(run.Sp(f == "scalar" ? F_scalar : F_scalar_old, q), 'p');
  } else if (f == "moist_qv") {
    b.add3(run.Sp(F_moist, 0), 'p');
  } else {
    char kind;
    if (!bc3_kind(f, &kind)) {
      err = "set_physical_bc3d: no boundary kind for field " + f;
      return true;
    }
    b.add3(run.F(field_id(f)), kind);
  }
  return true;
}

// the stages solve_em calls by name that are not plain pieces of the step;
// false: not one of them
static bool solve_em_stage(const Run &run, wrfdc_ctx *x, const std::string &n, std::string &err) {
  const int nm = x->c.nmoist, nsc = x->c.nscalar;
  const double dts = x->dts_rk;
  if (n == "initialize_moist_old") run.sync_old(true, false);
  else if (n == "rk_phys_bc_dry_1") {
    Run::BcBatch b(&run);
    run.phys_bc_dry_1(b);
  }
  else if (n == "relax_bdy_dry") run.lateral_relax();
  else if (n == "rk_addtend_dry") run.rk_addtend_dry();
  else if (n == "spec_bdy_dry") run.spec_overlays();
  else if (n == "spec_bdyupdate:u") run.spec_update(run.F(F_u_2), run.F(F_ru_tend), dts, 'u', false);
  else if (n == "spec_bdyupdate:v") run.spec_update(run.F(F_v_2), run.F(F_rv_tend), dts, 'v', false);
  else if (n == "spec_bdyupdate:t") run.spec_update(run.F(F_t_2), run.F(F_t_tend), dts, 't', false);
  else if (n == "spec_bdyupdate:mu_2") run.spec_update(run.F(F_mu_2), run.F(F_mu_tend), dts, 'm', true);
  else if (n == "spec_bdyupdate:muts") run.spec_update(run.F(F_muts), run.F(F_mu_tend), dts, 'm', true);
  else if (n == "spec_bdyupdate:w") err = "spec_bdyupdate:w is the nested-domain update of w; only specified boundaries are supported";
  else if (n == "spec_bdyupdate_ph") run.spec_bdyupdate_ph();
  else if (n == "zero_grad_bdy") run.zero_grad_bdy_w();
  else if (n == "calc_mu_uv_1") run.calc_mu_uv_1();
  else if (n == "rk_update_scalar_pd:moist_old") {
    for (int q = 0; q < nm; q++) run.update_scalar_pd(run.Sp(F_moist_old, q), run.Sp(F_moist_tend, q));
  }
  else if (n == "rk_update_scalar_pd:scalar_old") {
    for (int q = 0; q < nsc; q++) run.update_scalar_pd(run.Sp(F_scalar_old, q), run.Sp(F_scalar_tend, q));
  }
  else if (n == "moist_scalar_advance") {
    for (int q = 0; q < nm; q++)
      run.species_advance(run.Sp(F_moist, q), run.Sp(F_moist_old, q), run.Sp(F_moist_tend, q), q == 0, true, x->c.pd_moist != 0);
  }
  else if (n == "scalar_advance") {
    for (int q = 0; q < nsc; q++)
      run.species_advance(run.Sp(F_scalar, q), run.Sp(F_scalar_old, q), run.Sp(F_scalar_tend, q), false, false, x->c.pd_scalar != 0);
  }
  else if (n == "calc_p_rho_phi:final") run.calc_p_rho_phi();
  else if (n == "rk_phys_bc_dry_2") {
    Run::BcBatch b(&run);
    run.phys_bc_dry_2(b);
  }
  else if (n == "set_phys_bc_dry_2") {
    Run::BcBatch b(&run);
    run.set_phys_bc_dry_2(b);
  }
  else if (n.compare(0, 15, "spec_bdy_final:") == 0) {
    const std::string f = n.substr(15);
    char w = 0;
    if (f == "u") w = 'u';
    else if (f == "v") w = 'v';
    else if (f == "t") w = 't';
    else if (f == "ph") w = 'h';
    else if (f == "mu") w = 'm';
    else if (f == "moist" || f == "moist_qv") w = 'q';  // boundary data exist for QV only
    else {
      err = "unknown stage: " + n;
      return true;
    }
    Run::FinalBatch fb(&run);
    run.spec_final(fb, w);
  }
  else if (n == "set_w_surface") run.set_w_surface();
  else return named_bc(run, x, n, err);
  return true;
}

int wrfdc_stage(wrfdc_ctx *x, const char *name, int rk_step, int iteration) {
  if (!x || !name) return WRFDC_EINVAL;
  if (rk_step < 1 || rk_step > 3) return set_err(x, WRFDC_EINVAL, "rk_step must be 1, 2 or 3");
  return guarded(x, [&]() {
    Run run(x);
    run.set_rk(rk_step);
    if (iteration < 0 || iteration > x->nsmall) return set_err(x, WRFDC_EINVAL, "iteration outside [0, nsmall]");
    x->iteration = iteration;
    run.put_dtbc();
    std::string n(name);
    if (n == "rk_step_prep") run.rk_step_prep();
    else if (n == "first_rk_step_part1") run.first_rk_step_part1();
    else if (n == "first_rk_step_part2") run.first_rk_step_part2(true);
    else if (n == "compute_diff_metrics") run.compute_diff_metrics();
    else if (n == "cal_deform_and_div") run.cal_deform_and_div();
    else if (n == "smag2d_km") run.smag2d_km();
    else if (n == "calculate_n2") run.calculate_n2();
    else if (n == "calculate_km_kh") {
      run.calculate_n2();
      run.smag2d_km();
    }
    else if (n == "closure_km") {
      // the network's km as the first stage computes it (needs wrfdc_set_closure)
      if (!x->cl_theta) return set_err(x, WRFDC_EINVAL, "closure_km: no network set (wrfdc_set_closure)");
      run.smag_for_closure(x->cl_norm);
      run.closure_km();
    }
    else if (n == "horizontal_diffusion_2") run.horizontal_diffusion_2();
    else if (n == "rk_tendency") run.rk_tendency();
    else if (n == "advect_u") run.advect_u();
    else if (n == "advect_v") run.advect_v();
    else if (n == "advect_w") run.advect_w();
    else if (n == "advect_scalar") run.advect_scalar(run.F(F_t_2), run.F(F_t_tend), run.F(F_ru), run.F(F_rv), run.F(F_ww));
    else if (n == "advect_scalar_pd") run.advect_scalar_pd(run.Sp(F_moist, 0), run.Sp(F_moist_old, 0), run.F(F_advect_tend), x->dt_rk);
    else if (n == "rhs_ph") run.rhs_ph();
    else if (n == "horizontal_pressure_gradient") run.horizontal_pressure_gradient();
    else if (n == "pg_buoy_w") run.pg_buoy_w();
    else if (n == "dry_boundary_tendency") run.dry_boundary_tendency();
    else if (n == "small_step_prep") run.small_step_prep();
    else if (n == "calc_p_rho") run.calc_p_rho(iteration);
    else if (n == "calc_coef_w") run.calc_coef_w();
    else if (n == "advance_uv") run.advance_uv();
    else if (n == "advance_mu_t") run.advance_mu_t();
    else if (n == "advance_w") run.advance_w();
    else if (n == "sumflux") run.sumflux();
    else if (n == "acoustic_uv" || n == "acoustic_column") {
      // single kernels of the fused substep (for timing and profiling)
      const float dts = (float)x->dts_rk;
      x->iteration = iteration < 1 ? 1 : iteration;
      if (n == "acoustic_uv") {
        R3 r{1, run.ide, 1, run.kde - 1, 1, run.jde};
        k_uv_substep<<<Run::g3(r), Run::b3(), 0, x->s>>>(x->h, dts, r);
      } else {
        launch_acoustic_col(x->h, x->s, dts, x->iteration == 1, x->iteration == x->nsmall, (float)x->nsmall, 1);
      }
    }
    else if (n == "w_solve:columns" || n == "w_solve:levels") {
      // the tridiagonal w solve alone, in place on w_2 with the current
      // factors, in either layout (for measuring them against each other)
      const unsigned nb = (unsigned)((x->ide - 3 + 63) / 64) * (unsigned)(x->jde - 3);
      float *w = x->h.F[F_w_2];
      if (n == "w_solve:columns") {
        k_tridiag_lanes_columns<<<nb, 64, 0, x->s>>>(x->h, w, w);
      } else {
        if (x->kde > 64) return set_err(x, WRFDC_EINVAL, "w_solve:levels needs kde <= 64 (one level per lane)");
        const size_t lds = tridiag_levels_lds(x->kde);
        if (lds > 65536)
          hchk(hipFuncSetAttribute((const void *)k_tridiag_lanes_levels, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
               "w_solve LDS");
        k_tridiag_lanes_levels<<<nb, dim3(64, 8), lds, x->s>>>(x->h, w, w);
      }
    }
    else if (n == "acoustic_substep") {
      if (iteration < 1) return set_err(x, WRFDC_EINVAL, "acoustic_substep needs iteration >= 1");
      run.acoustic_substep(iteration, true);
    }
    else if (n == "small_step_finish") run.recouple();
    else if (n == "calc_p_rho_phi") run.calc_p_rho_phi();
    else if (n == "rk_stage") {
      run.rk_stage();
      if (run.species_in_alt()) run.species_to(false);
    }
    else if (n == "step") run.step();
    else {
      std::string err;
      if (!solve_em_stage(run, x, n, err)) err = "unknown stage: " + n;
      if (!err.empty()) return set_err(x, WRFDC_EINVAL, err);
    }
    return WRFDC_OK;
  });
}

int wrfdc_checksums(wrfdc_ctx *x, double out[WRFDC_NCHECKSUMS]) {
  static_assert(CK_N == WRFDC_NCHECKSUMS, "checksum count");
  if (!out) return set_err(x, WRFDC_EINVAL, "out is null");
  int rc = guarded(x, [&]() {
    k_checksum_partial<<<kCkBlocks, CK_BLOCK, 0, x->s>>>(x->h, x->ck);
    k_checksum_final<<<1, CK_BLOCK, 0, x->s>>>(x->ck, kCkBlocks, x->ck + kCkBlocks * CK_N);
    hchk(hipMemcpyAsync(out, x->ck + kCkBlocks * CK_N, sizeof(double) * CK_N, hipMemcpyDeviceToHost, x->s), "copy checksums");
    hchk(hipStreamSynchronize(x->s), "synchronize checksums");
    return WRFDC_OK;
  });
  if (rc != WRFDC_OK) return rc;
  if (out[3] > 0.0) return set_err(x, WRFDC_ENONFINITE, "non-finite values in the prognostic state");
  return WRFDC_OK;
}

int wrfdc_checksums_noalloc(const wrfdc_config *cfg, void *const *fields, double *out_dev, void *hip_stream) {
  if (!cfg || !fields || !out_dev) return set_err(nullptr, WRFDC_EINVAL, "cfg, fields and out_dev must not be null");
  std::string msg;
  int rc = validate(*cfg, msg);
  if (rc != WRFDC_OK) return set_err(nullptr, rc, msg);
  DevCtx h;
  fill_devctx(*cfg, h);
  for (int f = 0; f < NFIELDS; f++) h.F[f] = static_cast<float *>(fields[f]);
  static const int needed[] = {F_u_2, F_v_2, F_w_2, F_t_2, F_ph_2, F_p, F_pb, F_phb, F_mu_2, F_mub,
                               F_c1h, F_c2h, F_dnw, F_msftx, F_msfty, F_moist};
  for (int f : needed)
    if (!h.F[f]) return set_err(nullptr, WRFDC_EUNBOUND, std::string("field not given: ") + kNames[f]);
  double *part = nullptr;
  rc = noalloc_scratch(&part);
  if (rc != WRFDC_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  k_checksum_partial<<<kCkBlocks, CK_BLOCK, 0, s>>>(h, part);
  k_checksum_final<<<1, CK_BLOCK, 0, s>>>(part, kCkBlocks, out_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(nullptr, WRFDC_EHIP, std::string("checksum kernels: ") + hipGetErrorString(e));
  return WRFDC_OK;
}

// ---------------------------------------------------------------------------
// batched ensemble: one RK3 step of every member captured in one graph
// ---------------------------------------------------------------------------
struct wrfdc_ensemble {
  std::vector<wrfdc_ctx *> m;
  std::vector<int> branch;  // stream of each member
  std::vector<hipStream_t> streams;
  hipStream_t root = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t nodes = 0;
};

static void ensemble_free(wrfdc_ensemble *e) {
  if (!e) return;
  if (e->root) (void)hipStreamSynchronize(e->root);
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  for (hipEvent_t v : e->join)
    if (v) (void)hipEventDestroy(v);
  if (e->fork) (void)hipEventDestroy(e->fork);
  for (hipStream_t st : e->streams)
    if (st) (void)hipStreamDestroy(st);
  if (e->root) (void)hipStreamDestroy(e->root);
  delete e;
}

int wrfdc_ensemble_create(wrfdc_ctx *const *members, int n, const int *branch_of, int nbranches, wrfdc_ensemble **out) {
  if (!members || !out || n < 1 || n > kMaxMembers || nbranches < 1 || nbranches > 64)
    return set_err(nullptr, WRFDC_EINVAL, "ensemble: 1..128 members and 1..64 streams");
  *out = nullptr;
  for (int q = 0; q < n; q++) {
    if (!members[q]) return set_err(nullptr, WRFDC_EINVAL, "ensemble: null member");
    for (int r = 0; r < q; r++)
      if (members[r] == members[q]) return set_err(nullptr, WRFDC_EINVAL, "ensemble: a member appears twice");
    int rc = ensure_ready(members[q]);
    if (rc != WRFDC_OK) return set_err(nullptr, rc, "ensemble member " + std::to_string(q) + ": " + members[q]->err);
    if (branch_of && (branch_of[q] < 0 || branch_of[q] >= nbranches))
      return set_err(nullptr, WRFDC_EINVAL, "ensemble: stream index out of range");
  }
  wrfdc_ensemble *e = new wrfdc_ensemble();
  e->m.assign(members, members + n);
  e->branch.resize(n);
  for (int q = 0; q < n; q++) e->branch[q] = branch_of ? branch_of[q] : q % nbranches;
  std::vector<hipStream_t> saved(n);
  for (int q = 0; q < n; q++) saved[q] = members[q]->s;
  bool capturing = false;
  try {
    // every stream at the default priority
    hchk(hipStreamCreateWithFlags(&e->root, hipStreamNonBlocking), "ensemble stream");
    e->streams.resize(nbranches, nullptr);
    e->join.resize(nbranches, nullptr);
    for (int b = 0; b < nbranches; b++) {
      hchk(hipStreamCreateWithFlags(&e->streams[b], hipStreamNonBlocking), "ensemble stream");
      hchk(hipEventCreateWithFlags(&e->join[b], hipEventDisableTiming), "ensemble event");
    }
    hchk(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming), "ensemble event");
    // the members' earlier work finishes before the capture's first replay reads it
    for (int q = 0; q < n; q++)
      if (saved[q]) hchk(hipStreamSynchronize(saved[q]), "ensemble member stream");
    hchk(hipDeviceSynchronize(), "ensemble setup");
    hchk(hipStreamBeginCapture(e->root, hipStreamCaptureModeThreadLocal), "begin capture");
    capturing = true;
    hchk(hipEventRecord(e->fork, e->root), "fork");
    for (int b = 0; b < nbranches; b++) hchk(hipStreamWaitEvent(e->streams[b], e->fork, 0), "fork wait");
    for (int q = 0; q < n; q++) {
      wrfdc_ctx *x = members[q];
      x->s = e->streams[e->branch[q]];
      x->capturing = true;
      Run run(x);
      run.step();
      x->capturing = false;
      hchk(hipGetLastError(), "captured step");
    }
    for (int b = 0; b < nbranches; b++) {
      hchk(hipEventRecord(e->join[b], e->streams[b]), "join");
      hchk(hipStreamWaitEvent(e->root, e->join[b], 0), "join wait");
    }
    capturing = false;
    hchk(hipStreamEndCapture(e->root, &e->graph), "end capture");
    hchk(hipGraphGetNodes(e->graph, nullptr, &e->nodes), "graph nodes");
    hchk(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0), "graph instantiate");
  } catch (const HipError &err) {
    if (capturing) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(e->root, &g);
      if (g) (void)hipGraphDestroy(g);
    }
    for (int q = 0; q < n; q++) {
      members[q]->s = saved[q];
      members[q]->capturing = false;
    }
    ensemble_free(e);
    return set_err(nullptr, WRFDC_EHIP, std::string("ensemble: ") + err.what + ": " + hipGetErrorString(err.e));
  }
  // replays run on the root stream; the members' own streams wait for it
  for (int q = 0; q < n; q++) members[q]->s = saved[q];
  *out = e;
  return WRFDC_OK;
}

int wrfdc_ensemble_step(wrfdc_ensemble *e, int nsteps, double dtbc0) {
  if (!e || nsteps < 0) return set_err(nullptr, WRFDC_EINVAL, "ensemble step: null ensemble or nsteps < 0");
  const int n = (int)e->m.size();
  try {
    for (int s = 0; s < nsteps; s++) {
      DtbcList L;
      memset(&L, 0, sizeof(L));
      L.n = n;
      for (int q = 0; q < n; q++) {
        wrfdc_ctx *x = e->m[q];
        x->dtbc = dtbc0 + (double)(s + 1) * x->c.dt;
        L.p[q] = x->dtbc_dev;
        L.v[q] = (float)x->dtbc;
      }
      k_put_dtbc<<<(n + 63) / 64, 64, 0, e->root>>>(L);
      hchk(hipGetLastError(), "ensemble dtbc");
      hchk(hipGraphLaunch(e->exec, e->root), "graph launch");
    }
  } catch (const HipError &err) {
    return set_err(nullptr, WRFDC_EHIP, std::string("ensemble: ") + err.what + ": " + hipGetErrorString(err.e));
  }
  return WRFDC_OK;
}

int wrfdc_ensemble_sync(wrfdc_ensemble *e) {
  if (!e) return set_err(nullptr, WRFDC_EINVAL, "null ensemble");
  hipError_t err = hipStreamSynchronize(e->root);
  if (err != hipSuccess) return set_err(nullptr, WRFDC_EHIP, std::string("ensemble sync: ") + hipGetErrorString(err));
  return WRFDC_OK;
}

void *wrfdc_ensemble_stream(wrfdc_ensemble *e) { return e ? (void *)e->root : nullptr; }
long long wrfdc_ensemble_nodes(wrfdc_ensemble *e) { return e ? (long long)e->nodes : -1; }
void wrfdc_ensemble_destroy(wrfdc_ensemble *e) { ensemble_free(e); }

// ---------------------------------------------------------------------------
// tangent-linear / adjoint scalar transport
// ---------------------------------------------------------------------------
int wrfdc_record_stages(wrfdc_ctx *x, float *const *dst, int nstages) {
  if (!x) return set_err(nullptr, WRFDC_EINVAL, "null context");
  if (nstages < 0 || (nstages > 0 && !dst)) return set_err(x, WRFDC_EINVAL, "dst must hold 5 pointers per stage");
  for (int q = 0; q < 5 * nstages; q++)
    if (!dst[q]) return set_err(x, WRFDC_EINVAL, "null checkpoint pointer");
  x->rec.assign(dst, dst + (size_t)5 * nstages);
  x->rec_n = nstages;
  x->rec_i = 0;
  return WRFDC_OK;
}

int wrfdc_recorded_stages(wrfdc_ctx *x) { return x ? x->rec_i : -1; }

static int transport_args(wrfdc_ctx *x, int rk, const float *const *stage) {
  if (!x) return set_err(nullptr, WRFDC_EINVAL, "null context");
  if (rk < 1 || rk > 3) return set_err(x, WRFDC_EINVAL, "rk_step must be 1, 2 or 3");
  if (!stage) return set_err(x, WRFDC_EINVAL, "stage checkpoint is null");
  for (int q = 0; q < 5; q++)
    if (!stage[q]) return set_err(x, WRFDC_EINVAL, "stage checkpoint pointer is null");
  return WRFDC_OK;
}

int wrfdc_transport_stage(wrfdc_ctx *x, int rk, const float *const *stage, const float *q_in, const float *q_old, float *q_out) {
  int rc = transport_args(x, rk, stage);
  if (rc != WRFDC_OK) return rc;
  if (!q_in || !q_old || !q_out) return set_err(x, WRFDC_EINVAL, "q_in, q_old and q_out must not be null");
  if (q_out == q_in || q_out == q_old) return set_err(x, WRFDC_EINVAL, "q_out must not alias q_in or q_old");
  return guarded(x, [&]() {
    Run run(x);
    run.set_rk(rk);
    run.transport_stage(stage, q_in, q_old, q_out);
    return WRFDC_OK;
  });
}

int wrfdc_transport_stage_adjoint(wrfdc_ctx *x, int rk, const float *const *stage, const float *qbar_out, float *qbar_old, float *qbar_in) {
  int rc = transport_args(x, rk, stage);
  if (rc != WRFDC_OK) return rc;
  if (!qbar_out || !qbar_old || !qbar_in) return set_err(x, WRFDC_EINVAL, "adjoint arrays must not be null");
  if (qbar_out == qbar_old || qbar_out == qbar_in) return set_err(x, WRFDC_EINVAL, "qbar_out must not alias the accumulators");
  return guarded(x, [&]() {
    Run run(x);
    run.set_rk(rk);
    run.transport_stage_adjoint(stage, qbar_out, qbar_old, qbar_in);
    return WRFDC_OK;
  });
}

// ---------------------------------------------------------------------------
// learned eddy viscosity
// ---------------------------------------------------------------------------
static_assert(sizeof(wrfdc_closure_norm) == sizeof(ClosureNorm), "closure norm layout");
static ClosureNorm to_norm(const wrfdc_closure_norm &n) {
  return ClosureNorm{n.d_shift, n.d_scale, n.n_shift, n.n_scale, n.dx_feature, n.kref};
}
static bool norm_ok(const wrfdc_closure_norm &n) { return n.d_scale > 0.f && n.n_scale > 0.f && n.kref > 0.f; }

int wrfdc_set_closure(wrfdc_ctx *x, const float *theta, const void *wbf, const wrfdc_closure_norm *norm) {
  if (!x) return set_err(nullptr, WRFDC_EINVAL, "null context");
  if (!theta) {
    x->cl_theta = nullptr;
    x->cl_wbf = nullptr;
    return WRFDC_OK;
  }
  if (!wbf || !norm) return set_err(x, WRFDC_EINVAL, "the network needs its bf16 weights and feature normalisation");
  if (!norm_ok(*norm)) return set_err(x, WRFDC_EINVAL, "closure normalisation: d_scale, n_scale and kref must be > 0");
  x->cl_theta = theta;
  x->cl_wbf = static_cast<const __bf16 *>(wbf);
  x->cl_norm = to_norm(*norm);
  return WRFDC_OK;
}

long long wrfdc_closure_row_count(const wrfdc_config *cfg) {
  if (!cfg) return -1;
  return (long long)(cfg->nx - 2) * cfg->nz * (cfg->ny - 2);
}

int wrfdc_closure_rows(wrfdc_ctx *x, const wrfdc_closure_norm *norm, void *rows_x, float *rows_t) {
  if (!x) return set_err(nullptr, WRFDC_EINVAL, "null context");
  if (!norm || !rows_x || !rows_t) return set_err(x, WRFDC_EINVAL, "norm, rows_x and rows_t must not be null");
  if (!norm_ok(*norm)) return set_err(x, WRFDC_EINVAL, "closure normalisation: d_scale, n_scale and kref must be > 0");
  return guarded(x, [&]() {
    Run run(x);
    run.closure_rows(to_norm(*norm), static_cast<__bf16 *>(rows_x), rows_t);
    return WRFDC_OK;
  });
}

int wrfdc_subgrid_stress(wrfdc_ctx *fine, int nx_coarse, int ny_coarse, int ratio, float *rows_t) {
  if (!fine || !rows_t) return set_err(nullptr, WRFDC_EINVAL, "fine context and rows_t must not be null");
  if (ratio < 1 || nx_coarse * ratio != fine->c.nx || ny_coarse * ratio != fine->c.ny)
    return set_err(fine, WRFDC_EINVAL, "the fine grid must be ratio times the coarse grid in x and y");
  if (nx_coarse < 12 || ny_coarse < 12) return set_err(fine, WRFDC_EINVAL, "coarse grid below 12 x 12");
  return guarded(fine, [&]() {
    R3 rc{2, nx_coarse - 1, 1, fine->kde - 1, 2, ny_coarse - 1};
    k_subgrid_stress<<<Run::g3p(rc), Run::b3(), 0, fine->s>>>(fine->h, rc, ratio, rows_t);
    return WRFDC_OK;
  });
}

int wrfdc_coarse_grain(wrfdc_ctx *fine, wrfdc_ctx *coarse, int ratio) {
  if (!fine || !coarse) return set_err(nullptr, WRFDC_EINVAL, "null context");
  if (ratio < 1 || coarse->c.nx * ratio != fine->c.nx || coarse->c.ny * ratio != fine->c.ny || coarse->c.nz != fine->c.nz ||
      coarse->c.nmoist != fine->c.nmoist)
    return set_err(coarse, WRFDC_EINVAL, "coarse grid must be the fine one divided by ratio in x and y, same levels and species");
  int rc = ensure_ready(fine);
  if (rc != WRFDC_OK) return rc;
  return guarded(coarse, [&]() {
    CoarseList L;
    memset(&L, 0, sizeof(L));
    L.ide_c = coarse->ide;
    L.jde_c = coarse->jde;
    L.ip_c = coarse->ip;
    L.kde_c = coarse->kde;
    L.ip_f = fine->ip;
    L.jde_f = fine->jde;
    L.kde_f = fine->kde;
    L.ratio = ratio;
    auto add = [&](int f, int stag, bool two_d, int species = -1) {
      CoarseField &c = L.f[L.n++];
      const size_t off_c = species < 0 ? 0 : (size_t)species * coarse->n3, off_f = species < 0 ? 0 : (size_t)species * fine->n3;
      c.dst = coarse->h.F[f] + off_c;
      c.src = fine->h.F[f] + off_f;
      c.stag = stag;
      c.nk = two_d ? 1 : fine->kde;
    };
    add(F_u_2, 1, false);
    add(F_v_2, 2, false);
    for (int f : {F_w_2, F_t_2, F_ph_2, F_phb, F_p, F_pb, F_al, F_alb}) add(f, 0, false);
    add(F_mu_2, 0, true);
    add(F_mub, 0, true);
    add(F_ht, 0, true);
    for (int q = 0; q < fine->c.nmoist && L.n < kMaxCoarse; q++) add(F_moist, 0, false, q);
    const size_t cells = (size_t)coarse->ide * coarse->jde * coarse->kde;
    unsigned nb = (unsigned)((cells + 255) / 256);
    if (nb > 4096) nb = 4096;
    k_coarse_grain<<<dim3(nb, L.n), 256, 0, coarse->s>>>(L);
    return WRFDC_OK;
  });
}

static int mlp_err(hipError_t e, const char *what) {
  if (e == hipSuccess) return WRFDC_OK;
  return set_err(nullptr, WRFDC_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}
constexpr int kMlpChunksPerBlock = 8;  // 256 rows per block of the training kernel

long long wrfdc_mlp_work_floats(long long nrows) {
  if (nrows < 0) return -1;
  const long long rows_per_block = (long long)kMlpChunksPerBlock * kChunk;
  return ((nrows + rows_per_block - 1) / rows_per_block) * (long long)(kMlpParams + 1);
}

int wrfdc_mlp_pack(const float *theta, void *wbf, void *stream) {
  if (!theta || !wbf) return set_err(nullptr, WRFDC_EINVAL, "theta and wbf must not be null");
  k_mlp_pack<<<16, 256, 0, static_cast<hipStream_t>(stream)>>>(theta, static_cast<__bf16 *>(wbf));
  return mlp_err(hipGetLastError(), "k_mlp_pack");
}

static int mlp_pass(const float *theta, const void *wbf, const void *rows_x, const float *rows_t, long long nrows, float lossnorm, float kref,
                    float *work, float *grad, float *loss, float *m, float *v, int step, float lr, void *stream) {
  if (!theta || !wbf || !rows_x || !rows_t || !work || nrows <= 0)
    return set_err(nullptr, WRFDC_EINVAL, "theta, wbf, rows and work must not be null and nrows > 0");
  if (!(kref > 0.f)) return set_err(nullptr, WRFDC_EINVAL, "kref must be > 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long rows_per_block = (long long)kMlpChunksPerBlock * kChunk;
  const long long nb = (nrows + rows_per_block - 1) / rows_per_block;
  if (nb > (1LL << 30)) return set_err(nullptr, WRFDC_EINVAL, "too many rows");
  k_mlp_train<true><<<(unsigned)nb, 64, 0, s>>>(theta, static_cast<const __bf16 *>(wbf), static_cast<const __bf16 *>(rows_x), rows_t, nrows, lossnorm,
                                                 kref, kMlpChunksPerBlock, work);
  int rc = mlp_err(hipGetLastError(), "k_mlp_train");
  if (rc != WRFDC_OK) return rc;
  const int adam = m != nullptr;
  k_mlp_reduce<<<(kMlpParams + 1 + 255) / 256, 256, 0, s>>>(work, (int)nb, grad, loss, const_cast<float *>(theta), m, v,
                                                             const_cast<__bf16 *>(static_cast<const __bf16 *>(wbf)), adam, step, lr);
  return mlp_err(hipGetLastError(), "k_mlp_reduce");
}

int wrfdc_mlp_grad(const float *theta, const void *wbf, const void *rows_x, const float *rows_t, long long nrows, float lossnorm, float kref,
                   float *work, float *grad, float *loss, void *stream) {
  if (!grad) return set_err(nullptr, WRFDC_EINVAL, "grad must not be null");
  return mlp_pass(theta, wbf, rows_x, rows_t, nrows, lossnorm, kref, work, grad, loss, nullptr, nullptr, 0, 0.f, stream);
}

int wrfdc_mlp_train_step(float *theta, void *wbf, const void *rows_x, const float *rows_t, long long nrows, float lossnorm, float kref,
                         float *work, float *adam_m, float *adam_v, int step, float lr, float *loss, void *stream) {
  if (!adam_m || !adam_v || step < 1) return set_err(nullptr, WRFDC_EINVAL, "Adam moments must not be null and step >= 1");
  return mlp_pass(theta, wbf, rows_x, rows_t, nrows, lossnorm, kref, work, nullptr, loss, adam_m, adam_v, step, lr, stream);
}

int wrfdc_mlp_forward(const float *theta, const void *wbf, const void *rows_x, long long nrows, float *z, void *stream) {
  if (!theta || !wbf || !rows_x || !z || nrows <= 0) return set_err(nullptr, WRFDC_EINVAL, "theta, wbf, rows_x and z must not be null");
  const long long rows_per_block = (long long)kMlpChunksPerBlock * kChunk;
  const long long nb = (nrows + rows_per_block - 1) / rows_per_block;
  k_mlp_train<false><<<(unsigned)nb, 64, 0, static_cast<hipStream_t>(stream)>>>(theta, static_cast<const __bf16 *>(wbf),
                                                                              static_cast<const __bf16 *>(rows_x), nullptr, nrows, 0.f, 1.f,
                                                                              kMlpChunksPerBlock, z);
  return mlp_err(hipGetLastError(), "k_mlp_train<forward>");
}

int wrfdc_sync(wrfdc_ctx *x) {
  if (!x) return WRFDC_EINVAL;
  hipError_t e = x->s ? hipStreamSynchronize(x->s) : hipDeviceSynchronize();
  if (e != hipSuccess) return set_err(x, WRFDC_EHIP, std::string("synchronize: ") + hipGetErrorString(e));
  return WRFDC_OK;
}

const char *wrfdc_last_error(wrfdc_ctx *x) { return x ? x->err.c_str() : g_thread_err.c_str(); }

}  // extern "C"
