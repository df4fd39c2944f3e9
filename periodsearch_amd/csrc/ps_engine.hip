// ps_engine.hip -- MI355X (gfx950) engine for the trial-period x initial-pole
// Levenberg-Marquardt fit loop of period_search (Asteroids@home).
//
// Design: continuous batching over persistent fit slots.  A fixed set of
// slots lives in HBM; each slot holds the complete LM state of one fit
// (parameters, damped/accepted normal matrices, the Jacobian of the last
// model evaluation).  One LM iteration of every active slot is three
// kernels, each sized for its own work and register budget:
//
//   step_kernel   one workgroup per slot: finishes a converged fit and takes
//                 the next fit from the work queue, or damps the accepted
//                 normal equations and solves them (Gauss-Jordan with full
//                 pivoting, period_search/gauss_errc.c:9-73: in the
//                 registers of one wave up to 63 free parameters, of four
//                 waves up to 127, in LDS beyond) to form the trial
//                 parameters (period_search/mrqmin.c:53-79), and evaluates
//                 the shape at them (curv.c).
//   model kernel  one workgroup per (slot, tile of 64 points): the
//                 per-point rotation/phase geometry in FP64 and the facet
//                 loop with the brightness-derivative contraction on the
//                 matrix cores (period_search/bright.c, matrix.c, phasec.c).
//   neq_kernel    one workgroup per slot: convexity rows (conv.c), relative
//                 renormalisation and the normal equations (mrqcof.c), then
//                 accept/reject (mrqmin.c:82-107) and the driver bookkeeping
//                 (period_search/period_search_BOINC.cpp:868-892).
//
// A slot whose fit ends is refilled from the queue in the next step_kernel,
// so fits with different iteration counts keep every slot busy.  A last
// kernel reduces the poles of each trial period to one output row
// (period_search_BOINC.cpp:913-924).  iter_kernel runs neq_kernel's and
// step_kernel's work in one launch (an option: measured slower than two).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "periodsearch.h"
#include "ps_internal.h"
#include "ps_proposal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxNcoef = 128;
constexpr int kMaxMa = kMaxNcoef + 8;
constexpr int kMaxMfit = kMaxMa;
constexpr int kChunk = 32;         // points staged per normal-equation chunk
constexpr int kTilePts = 64;       // points per MFMA model tile (4 row blocks of 16)
constexpr int kFacetChunk = 32;    // facets per MFMA model chunk (8 k-steps of 4)
constexpr int kMaxFacets = 1000;
constexpr int kMaxGroups = 4;      // slot groups, one HIP stream each
constexpr double kTiny = 1e-8;     // period_search/constants.h:10
constexpr double kPi = 3.14159265358979323846;

typedef double v4d __attribute__((ext_vector_type(4)));  // one 16x16 FP64 MFMA accumulator per lane
typedef float v4f __attribute__((ext_vector_type(4)));   // one 16x16 FP32-accumulate MFMA accumulator per lane
typedef __bf16 v8bf __attribute__((ext_vector_type(8))); // bf16 operand of v_mfma_f32_16x16x32_bf16
typedef unsigned v4u __attribute__((ext_vector_type(4)));  // one 16-byte piece held in registers

// f(std::integral_constant<int, I>) for I = B..N-1: a loop whose index is a
// compile-time constant in the body, so that arrays indexed by it stay in
// registers.
template <int B, int N, class Fn>
__device__ __forceinline__ void static_for(Fn &&f) {
  if constexpr (B < N) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, N>(f);
  }
}

#define HIP_TRY(x)                                                         \
  do {                                                                     \
    hipError_t err_ = (x);                                                 \
    if (err_ != hipSuccess) {                                              \
      e->err = std::string(#x) + ": " + hipGetErrorString(err_);           \
      return PS_ERR_HIP;                                                   \
    }                                                                      \
  } while (0)

// Read-only problem tables in device memory plus the constants of the fit.
struct DevProblem {
  // points, SoA, real points only
  const double *tim, *y, *sig, *w, *ex, *ey, *ez, *sx, *sy, *sz;
  const double *swt;    // [n_points] w / sig^2 (mixed-precision normal equations)
  const float *ynorm;   // [n_points] brightness / light-curve mean - 1 (proposal features)
  const int *lc_start;  // [n_lc + 2]: real light curves then the conv "light curve"
  const int *lc_inrel;  // [n_lc + 1]
  int n_points, n_lc;
  // facets
  const double *nx, *ny, *nz, *darea;
  const double *dsph;     // [F][ncp], zero-padded columns
  // mixed-precision tables: normals, and Dsph transposed [16*nb][fpad] (zero
  // beyond ncoef and beyond F) in FP32 and in bf16
  const float *nor3f;      // [fpad][3] (nx, ny, nz), zero past the last facet: a wave's eight
                           // normals are 24 consecutive floats, fetched by wide scalar loads
  const float *dsphT32;
  const unsigned short *dsphT16, *dsphT16lo;  // bf16 of Dsph, and of Dsph - hi (split forms)
  const unsigned short *dsphT16lo2;           // bf16 of Dsph - hi - lo (three-part split form)
  // dsphT16 with each 32-facet chunk permuted for the wave-owned-points
  // kernel: chunk position 8g + j holds facet 4j + g
  const unsigned short *dsphQ16;
  int fpad;
  const double4 *nor4;    // [F] (nx, ny, nz, 0): one 32-byte scalar load per facet
  const double *dsphP;    // [fpad][16*nb] Dsph zero-padded in rows and columns (FP64 matrix-core kernel)
  const double *cpleg;    // [ngrp][F]: P_l^m of coefficient group q at facet f
  const double *cfc;      // [mmax+1][F]
  const double *cfs;      // [mmax+1][F]
  const int *grp_m;       // [ngrp]
  int nfac, ngrp, ncoef, ncp, ma, mfit;
  const int *fidx;        // [mfit] free parameter -> index into a[]
  const double *cg_first; // [ncoef]
  const double *pole_beta, *pole_lambda;  // [n_poles] deg
  double phi0, alamda_start, alamda_incr, log_cl, par0, par1, par2;
  int n_iter_max, n_iter_min;
  double iter_diff_max;
  int ndata;   // n_points + 3
  int npad;    // Jacobian leading dimension (>= n_points + 3)
  int ntiles;    // point tiles of kThreads (VALU model kernel)
  int ntiles64;  // point tiles of kTilePts (MFMA model kernel)
};

// Per-slot LM state (scalars).  Vectors and matrices live in the slot's
// block of slot memory (SlotMem).
struct SlotState {
  int item, phase, niter, err;
  int eval, sel, prob, pad1;  // prob: index of the fit's problem (asteroid) in the batch
  long long t_start;          // device wall clock when the fit took the slot
  double lam, ochisq, chisq, chitr, dev_new, dev_old, iter_diff, rchisq;
  double chck[3], chtr[3];
};
enum { PH_EMPTY = 0, PH_INIT = 1, PH_ITER = 2, PH_DRAINED = 3 };

// Offsets (in doubles) inside one slot's memory block.
struct SlotMem {
  long long J, alpha0, alpha1, beta0, beta1, da, a, atry, alast, area, g, perm, pys, psg, psw, total;
};

__host__ __device__ inline SlotMem slot_mem(int ma, int mfit, int npad) {
  SlotMem m;
  long long o = 0;
  m.area = o; o += kMaxFacets;   // curv() at atry: Area, g (written by step_kernel)
  m.g = o; o += kMaxFacets;
  m.J = o; o += (long long)(ma + 1) * npad;
  m.alpha0 = o; o += (long long)mfit * mfit;
  m.alpha1 = o; o += (long long)mfit * mfit;
  m.beta0 = o; o += kMaxMfit;
  m.beta1 = o; o += kMaxMfit;
  m.da = o; o += kMaxMfit;
  m.a = o; o += kMaxMa;
  m.atry = o; o += kMaxMa;
  m.alast = o; o += kMaxMa;
  // ordered points (point_order): sorted position -> point (int), and the
  // points' brightness, sigma and weight/sigma^2 in sorted order
  m.perm = o; o += (npad + 1) / 2;
  m.pys = o; o += npad;
  m.psg = o; o += npad;
  m.psw = o; o += npad;
  o = (o + 31) / 32 * 32;  // 256-byte aligned blocks
  m.total = o;
  return m;
}

// One batch of work: fits (mode 0) or single model evaluations (mode 1).
struct DevWork {
  int n_items, mode, nslots;  // mode 0: fits from the reference starts, 1: evaluations, 2: fits from cg_in
  int order;  // 1: each fit's points in its own order (point_order; mixed forms with the K-split normal equations)
  double us_per_tick;       // period of the device wall clock [us]
  const double *freq;       // [n_items] (mode 0)
  const int *pole;          // [n_items] (mode 0)
  const double *cg_in;      // [n_items][ma] (modes 1, 2)
  double conw_r;
  ps_fit_out *out;          // [n_items] (modes 0, 2)
  double *cg_out;           // [n_items][ma] final parameters, or null
  const int *item_prob;     // [n_items] problem of each item (mode 0, several problems), or null: problem 0
  const double *conw_rs;    // [n_problems] convexity weight per problem, or null: conw_r
  double *ev_alpha, *ev_beta, *ev_chi, *ev_chck;  // (mode 1)
  int *queue, *done;   // done[1] counts model evaluations
  SlotState *st;
  double *mem;
  long long slot_doubles;
};

// curv (period_search/curv.c:10-38): g of facet f for the shape part of a
// (Area = Darea * g).  Same operation order as the reference.
__device__ inline double curv_g(const DevProblem &P, const double *a, int f) {
#pragma clang fp contract(off)
  double s = 0;
  int n = 0;
  for (int q = 0; q < P.ngrp; q++) {
    const int m = P.grp_m[q];
    double fsum = a[n] * P.cfc[m * P.nfac + f];
    n++;
    if (m != 0) {
      fsum = fsum + a[n] * P.cfs[m * P.nfac + f];
      n++;
    }
    s = s + P.cpleg[q * P.nfac + f] * fsum;
  }
  return exp(s);
}

__device__ void eval_curv(const DevProblem &P, const double *a, double *area, double *g) {
#pragma clang fp contract(off)
  for (int f = threadIdx.x; f < P.nfac; f += blockDim.x) {
    const double gf = curv_g(P, a, f);
    g[f] = gf;
    area[f] = P.darea[f] * gf;
  }
}

// curv() at a slot's trial parameters (period_search/curv.c:10-38, the
// reference's operation order) by the threads of one workgroup: Area and g
// of every facet into the slot's memory, read by the model and normal-
// equation kernels.  The slot memory layout (SlotMem) is passed in.
struct SlotMem;
__device__ void slot_curv(const DevProblem &P, const SlotMem &M, double *mem, int tid, int nthreads);

// blmatrix (period_search/blmatrix.c:9-46) into LDS: B[9], dB/dbeta[9], dB/dlambda[9].
__device__ void eval_blmatrix(double bet, double lam, double *blm) {
  double cb = cos(bet), sb = sin(bet), cl = cos(lam), sl = sin(lam);
  double *B = blm, *D1 = blm + 9, *D2 = blm + 18;
  B[0] = cb * cl; B[1] = cb * sl; B[2] = -sb;
  B[3] = -sl;     B[4] = cl;      B[5] = 0;
  B[6] = sb * cl; B[7] = sb * sl; B[8] = cb;
  D1[0] = -sb * cl; D1[1] = -sb * sl; D1[2] = -cb;
  D1[3] = 0;        D1[4] = 0;        D1[5] = 0;
  D1[6] = cb * cl;  D1[7] = cb * sl;  D1[8] = -sb;
  D2[0] = -cb * sl; D2[1] = cb * cl;  D2[2] = 0;
  D2[3] = -cl;      D2[4] = -sl;      D2[5] = 0;
  D2[6] = -sb * sl; D2[7] = sb * cl;  D2[8] = 0;
}

// 1/x for x in (2e-8, 2]: hardware reciprocal estimate refined by two
// Newton-Raphson steps r <- r + r(1 - x r) (quadratic convergence from the
// ~2^-23 estimate to below 1 ulp); the exact IEEE division is not needed in
// the facet loop, whose sums are regrouped anyway.
__device__ inline double recip(double x) {
#ifdef PS_EXACT_DIV
  return 1.0 / x;
#else
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
#endif
}

// Per-point geometry, part 1: the phase function (period_search/phasec.c:11-24)
// and the rotated Earth/Sun directions e = tmat ee, e0 = tmat ee0 with
// tmat = Rz(f) Blmat, f = fmod(omega t + phi0, 2 pi) (period_search/matrix.c:11-58,
// bright.c:36-55).  Same operations and order as the reference.
struct PointGeom {
  double e[3], e0[3];
  double cf, sf, t;
  double scale, dphp[3];
};

__device__ void point_geometry(const DevProblem &P, int p, const double *a, const double *blm, PointGeom &G) {
#pragma clang fp contract(off)
  const int nc = P.ncoef;
  const double ee[3] = {P.ex[p], P.ey[p], P.ez[p]};
  const double e0[3] = {P.sx[p], P.sy[p], P.sz[p]};
  const double t = P.tim[p];
  double cos_alpha = ee[0] * e0[0] + ee[1] * e0[1] + ee[2] * e0[2];
  double alpha = acos(cos_alpha);
  double p1 = a[nc + 3], p2 = a[nc + 4], p3 = a[nc + 5];
  double ex = exp(-alpha / p2);
  G.scale = 1 + p1 * ex + p3 * alpha;
  G.dphp[0] = ex;
  G.dphp[1] = p1 * ex * alpha / (p2 * p2);
  G.dphp[2] = alpha;
  double f = a[nc + 2] * t + P.phi0;
  f = fmod(f, 2 * kPi);
  const double cf = cos(f), sf = sin(f);
  const double *B = blm;
  for (int i = 0; i < 3; i++) {
    // row i of tmat = fmat row i times Blmat
    double t0, t1, t2;
    if (i == 0) {
      t0 = cf * B[0] + sf * B[3];
      t1 = cf * B[1] + sf * B[4];
      t2 = cf * B[2] + sf * B[5];
    } else if (i == 1) {
      t0 = (-sf) * B[0] + cf * B[3];
      t1 = (-sf) * B[1] + cf * B[4];
      t2 = (-sf) * B[2] + cf * B[5];
    } else {
      t0 = B[6];
      t1 = B[7];
      t2 = B[8];
    }
    G.e[i] = t0 * ee[0] + t1 * ee[1] + t2 * ee[2];
    G.e0[i] = t0 * e0[0] + t1 * e0[1] + t2 * e0[2];
  }
  G.cf = cf;
  G.sf = sf;
  G.t = t;
}

// Per-point geometry, part 2: derivatives of e and e0 with respect to
// (beta, lambda, omega) contracted with the facet sums V = sum Area dsmu Nor
// and U = sum Area dsmu0 Nor: r_k = sum_i de_i,k V_i + de0_i,k U_i
// (bright.c:41-55 and :88-103 regrouped).
__device__ void rotation_derivs(const DevProblem &P, int p, const double *blm, const PointGeom &G, const double V[3],
                                const double U[3], double r[3]) {
  const double ee[3] = {P.ex[p], P.ey[p], P.ez[p]};
  const double e0[3] = {P.sx[p], P.sy[p], P.sz[p]};
  const double cf = G.cf, sf = G.sf, t = G.t;
  for (int k = 0; k < 3; k++) {
    // dtm_k rows 0..2
    double m[3][3];
    if (k < 2) {
      const double *D = blm + 9 * (k + 1);
      for (int j = 0; j < 3; j++) {
        m[0][j] = cf * D[j] + sf * D[3 + j];
        m[1][j] = (-sf) * D[j] + cf * D[3 + j];
        m[2][j] = D[6 + j];
      }
    } else {
      const double *B = blm;
      for (int j = 0; j < 3; j++) {
        m[0][j] = ((-t) * sf) * B[j] + (t * cf) * B[3 + j];
        m[1][j] = ((-t) * cf) * B[j] + ((-t) * sf) * B[3 + j];
        m[2][j] = 0;
      }
    }
    double s = 0;
    for (int i = 0; i < 3; i++) {
      const double de = m[i][0] * ee[0] + m[i][1] * ee[1] + m[i][2] * ee[2];
      const double de0 = m[i][0] * e0[0] + m[i][1] * e0[1] + m[i][2] * e0[2];
      s += de * V[i] + de0 * U[i];
    }
    r[k] = s;
  }
}

// Pass 1 for one tile of kThreads points: brightness and its derivatives
// (period_search/bright.c:14-126), rows written to the slot scratch J
// (column-major, leading dimension npad; row ma holds ymod).
// NCB = accumulator count (ncoef rounded up to 8; Dsph is zero-padded).
template <int NCB>
__device__ void model_tile(const DevProblem &P, int p0, const double *a, const double *blm, const double *area,
                           double *J, PointGeom *gst) {
  const int lane_p = p0 + (int)threadIdx.x;
  const int wave_first = p0 + ((int)threadIdx.x & ~63);
  if (wave_first >= P.n_points) return;  // whole wave idle
  const int p = lane_p < P.n_points ? lane_p : P.n_points - 1;
  {
    PointGeom G;
    point_geometry(P, p, a, blm, G);
    gst[threadIdx.x] = G;  // parked in LDS while the facet loop runs
  }
  const double cl = exp(a[P.ma - 2]);
  const double cls = a[P.ma - 1];
  double S[NCB];
#pragma unroll
  for (int c = 0; c < NCB; c++) S[c] = 0;
  double br = 0, v0 = 0, v1 = 0, v2 = 0, u0 = 0, u1 = 0, u2 = 0, scl = 0, scls = 0;
  const double e_0 = gst[threadIdx.x].e[0], e_1 = gst[threadIdx.x].e[1], e_2 = gst[threadIdx.x].e[2];
  const double s_0 = gst[threadIdx.x].e0[0], s_1 = gst[threadIdx.x].e0[1], s_2 = gst[threadIdx.x].e0[2];
  const double4 *__restrict__ n4 = P.nor4;
  const double *__restrict__ dsp = P.dsph;
#pragma unroll 1
  for (int f = 0; f < P.nfac; f++) {
    const double4 nv = n4[f];
    const double n0 = nv.x, n1 = nv.y, n2 = nv.z;
    const double mu = e_0 * n0 + e_1 * n1 + e_2 * n2;
    const double mu0 = s_0 * n0 + s_1 * n1 + s_2 * n2;
    if (mu > kTiny && mu0 > kTiny) {
      const double af = area[f];
      const double rd = recip(mu + mu0);
      const double mm = mu * mu0;
      const double s = mm * (cl + cls * rd);
      const double wgt = af * s;
      br += wgt;
      const double q0 = mu0 * rd, q1 = mu * rd;
      const double ad = af * (cls * q0 * q0 + cl * mu0);
      const double ad0 = af * (cls * q1 * q1 + cl * mu);
      v0 += ad * n0;
      v1 += ad * n1;
      v2 += ad * n2;
      u0 += ad0 * n0;
      u1 += ad0 * n1;
      u2 += ad0 * n2;
      const double am = af * mm;
      scl += am;
      scls += am * rd;
      const double *drow = dsp + (size_t)f * NCB;
#pragma unroll
      for (int c = 0; c < NCB; c++) S[c] += wgt * drow[c];
    }
  }
  if (lane_p >= P.n_points) return;
  const PointGeom G = gst[threadIdx.x];
  const int ld = P.npad;
  const double sc = G.scale;
#pragma unroll
  for (int c = 0; c < NCB; c++)
    if (c < P.ncoef) J[(size_t)c * ld + p] = sc * S[c];
  const double V[3] = {v0, v1, v2}, U[3] = {u0, u1, u2};
  double r[3];
  rotation_derivs(P, p, blm, G, V, U, r);
  for (int k = 0; k < 3; k++) J[(size_t)(P.ncoef + k) * ld + p] = sc * r[k];
  for (int i = 0; i < 3; i++) J[(size_t)(P.ncoef + 3 + i) * ld + p] = br * G.dphp[i];
  J[(size_t)(P.ncoef + 6) * ld + p] = sc * scl * cl;
  J[(size_t)(P.ncoef + 7) * ld + p] = sc * scls;
  J[(size_t)P.ma * ld + p] = br * sc;
}

// The three convexity rows (period_search/conv.c:12-29) at points
// n_points..n_points+2, and chck = sum Area*Nor.  Each sum over facets is
// split into kConvParts interleaved parts (facets f = part + kConvParts*i,
// in order), summed by separate threads with their loads issued eight
// facets at a time, and the parts are added in order; consecutive threads
// take consecutive columns, so the Dsph rows are read coalesced.  red:
// LDS scratch of kConvParts * (3 * ncoef + 3) doubles.  MIX: derivative rows
// stored as FP32 (the model value row ma stays FP64).
constexpr int kConvParts = 4;

template <bool MIX>
__device__ void conv_rows(const DevProblem &P, const double *area, const double *g, double *J, double *chck,
                          double *red) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x, ld = P.npad, n = P.n_points, F = P.nfac, nc = P.ncoef;
  float *Jf = reinterpret_cast<float *>(J);
  auto put = [&](int c, int q, double v) {
    if (MIX)
      Jf[(size_t)c * ld + q] = (float)v;
    else
      J[(size_t)c * ld + q] = v;
  };
  // tasks: (part, column) for the shape columns, then (part, k) for chck
  const int ntask = kConvParts * (nc + 3);
  for (int t = tid; t < ntask; t += blockDim.x) {
    const int part = t / (nc + 3), c = t - part * (nc + 3);
    double r0 = 0, r1 = 0, r2 = 0;
    if (c < nc) {
      for (int f0 = part; f0 < F; f0 += 8 * kConvParts) {
        double gv[8], dv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int f = min(f0 + kConvParts * u, F - 1);
          gv[u] = g[f];
          dv[u] = P.dsph[(size_t)f * P.ncp + c];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int f = f0 + kConvParts * u;
          if (f < F) {
            const double dg = gv[u] * dv[u];
            const double tt = P.darea[f] * dg;
            r0 += tt * P.nx[f];
            r1 += tt * P.ny[f];
            r2 += tt * P.nz[f];
          }
        }
      }
      red[(size_t)part * (3 * nc + 3) + 3 * c] = r0;
      red[(size_t)part * (3 * nc + 3) + 3 * c + 1] = r1;
      red[(size_t)part * (3 * nc + 3) + 3 * c + 2] = r2;
    } else {
      const int k = c - nc;
      const double *nv = k == 0 ? P.nx : (k == 1 ? P.ny : P.nz);
      double r = 0;
      for (int f0 = part; f0 < F; f0 += 8 * kConvParts) {
        double av[8], nn[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int f = min(f0 + kConvParts * u, F - 1);
          av[u] = area[f];
          nn[u] = nv[f];
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
          if (f0 + kConvParts * u < F) r += av[u] * nn[u];
      }
      red[(size_t)part * (3 * nc + 3) + 3 * nc + k] = r;
    }
  }
  __syncthreads();
  for (int t = tid; t < P.ma + 3; t += blockDim.x) {
    if (t < nc) {
      for (int k = 0; k < 3; k++) {
        double v = red[3 * t + k];
        for (int part = 1; part < kConvParts; part++) v = v + red[(size_t)part * (3 * nc + 3) + 3 * t + k];
        put(t, n + k, v);
      }
    } else if (t < P.ma) {
      put(t, n, 0);
      put(t, n + 1, 0);
      put(t, n + 2, 0);
    } else {
      const int k = t - P.ma;
      double v = red[3 * nc + k];
      for (int part = 1; part < kConvParts; part++) v = v + red[(size_t)part * (3 * nc + 3) + 3 * nc + k];
      J[(size_t)P.ma * ld + n + k] = v;
      chck[k] = v;
    }
  }
}

// Pass 2 on the matrix cores: relative renormalisation (mrqcof.c:42-104)
// and the normal equations (mrqcof.c:106-139) with alpha = WT^T Jn summed by
// v_mfma_f64_16x16x4 over blocks of 4 points.  Points are staged kNeqChunk
// at a time in LDS as renormalised rows Jn of the free parameters (columns
// padded with zeros to MP = 16*nb, row stride MP+1); the weighted row
// WT = Jn*sig2i*w is formed in registers with the reference's operations.
// The lower-triangle 16x16 tiles (jb >= kb) are spread over the 4 waves
// (tile t on wave t%4, at most NT per wave); alpha takes the j >= k entries
// and mirrors them, as the reference does.  beta and chi2 are summed by
// single threads in point order.
// MIX (mixed precision): the derivative rows are FP32, Jn is staged as FP32
// and alpha is summed by v_mfma_f32_16x16x4f32 (FP32 accumulation); the
// light-curve sums, beta, chi2 and the residuals stay FP64.
constexpr int kNeqChunk = 64;
// entries of the normal-equation kernels' free-parameter table in LDS: the
// staged kernel reads rows jf < 16*ceil(mfit/16) + 28 without a bound check
constexpr int kFidxPad = (kMaxMfit + 15) / 16 * 16 + 32;

// Doubles taken by the staged rows Jn (also the light-curve-sum and the
// convexity-row scratch; ncoef <= mfit + 1).
__host__ __device__ inline int neq_jn_doubles(int mfit, bool mix) {
  const int MS = (mfit + 15) / 16 * 16 + 1;
  int d = mix ? (kNeqChunk * MS + 1) / 2 : kNeqChunk * MS;
  if (d < 4 * (kMaxMa + 1)) d = 4 * (kMaxMa + 1);
  if (d < kConvParts * (3 * (mfit + 1) + 3)) d = kConvParts * (3 * (mfit + 1) + 3);
  return d;
}

template <int NT, bool MIX>
__device__ __forceinline__ void normal_equations_mfma(const DevProblem &P, const double *J, const int *perm, double conw_r,
                                                      double *wst, double *dave, double *drat, const int *fidx, double *beta,
                                                      double *chi_out, double *alpha) {
#pragma clang fp contract(off)
  using AccT = typename std::conditional<MIX, v4f, v4d>::type;
  using JnT = typename std::conditional<MIX, float, double>::type;
  const float *Jf = reinterpret_cast<const float *>(J);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ld = P.npad, mfit = P.mfit, ma = P.ma;
  const int nb = (mfit + 15) / 16, MP = 16 * nb, MS = MP + 1;
  const int ntile = nb * (nb + 1) / 2;
  AccT acc[NT];
  int tj[NT], tk[NT];
#pragma unroll
  for (int i = 0; i < NT; i++) {
    acc[i] = AccT{0, 0, 0, 0};
    int t = wave + 4 * i, jb = 0;
    while ((jb + 1) * (jb + 2) / 2 <= t) jb++;
    tj[i] = jb;
    tk[i] = t - jb * (jb + 1) / 2;
  }
  JnT *Jn = reinterpret_cast<JnT *>(wst);       // [kNeqChunk][MS]
  double *aux = wst + neq_jn_doubles(mfit, MIX);  // [kNeqChunk][3]: sig2i, weight, dy
  double bacc = 0, chi = 0;
  for (int lc = 0; lc <= P.n_lc; lc++) {
    const int q0 = P.lc_start[lc], q1 = P.lc_start[lc + 1], np = q1 - q0;
    const int inrel = P.lc_inrel[lc];
    if (inrel) {
      // light-curve sums (mrqcof.c:56-75): ave = sum ymod, dave = sum dyda.
      // Wave w sums columns w, w+4, ..., two columns at a time with the lanes
      // over the points (coalesced loads, four partial sums per lane and
      // column, eight loads in flight), then a butterfly over the wave: at
      // every step lanes l and l^o add the same two values, so every lane
      // ends with the same total.
      auto lc_sums = [&](auto *pa, auto *pb, bool two, int ca, int cb) __attribute__((always_inline)) {
        double sa[4] = {0, 0, 0, 0}, sb[4] = {0, 0, 0, 0};
        for (int qb = q0; qb < q1; qb += 256) {
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int q = qb + lane + 64 * u;
            const bool in = q < q1;
            const int qc = in ? q : q0;
            const double va = (double)pa[qc], vb = (double)pb[qc];
            sa[u] = sa[u] + (in ? va : 0.0);
            sb[u] = sb[u] + (in ? vb : 0.0);
          }
        }
        double ta = (sa[0] + sa[1]) + (sa[2] + sa[3]), tb = (sb[0] + sb[1]) + (sb[2] + sb[3]);
        for (int o = 32; o > 0; o >>= 1) {
          ta = ta + __shfl_xor(ta, o, 64);
          tb = tb + __shfl_xor(tb, o, 64);
        }
        if (lane == 0) {
          dave[ca] = ta;
          if (two) dave[cb] = tb;
        }
      };
      for (int ca = wave; ca <= ma; ca += 2 * kWaves) {
        const int cb0 = ca + kWaves;
        const bool two = cb0 <= ma;
        const int cb = two ? cb0 : ca;
        if (MIX && cb < ma)  // both derivative rows (FP32)
          lc_sums(Jf + (size_t)ca * ld, Jf + (size_t)cb * ld, two, ca, cb);
        else if (MIX && ca < ma)  // ca FP32, cb the model row (FP64)
          lc_sums(Jf + (size_t)ca * ld, J + (size_t)cb * ld, two, ca, cb);
        else
          lc_sums(J + (size_t)ca * ld, J + (size_t)cb * ld, two, ca, cb);
      }
      __syncthreads();
      // dave[l]/ave, once per light curve (the reference divides the product
      // ymod*dave[l] by ave for every point; the quotient first differs from
      // it by rounding only)
      for (int col = tid; col < ma; col += kThreads) drat[col] = dave[col] / dave[ma];
      __syncthreads();
    }
    // the rows of a chunk (this lane's point, kRounds batches of eight rows
    // per wave) are loaded before the chunk starts: the light curve's first
    // chunk here, each next one while the current one is being summed
    constexpr int kRounds = NT <= 1 ? 1 : (NT <= 3 ? 2 : (NT <= 6 ? 3 : 4));  // >= 16*nb / (8*kWaves)
    JnT pre[kRounds][8];
    auto load_rows = [&](int cs) __attribute__((always_inline)) {
      const int qn = cs + lane;
      const int qc = qn < q1 ? qn : q0;  // a valid point for idle lanes
#pragma unroll
      for (int r = 0; r < kRounds; r++)
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int row = fidx[wave + 8 * kWaves * r + kWaves * u];  // zero past mfit: a valid row
          if constexpr (MIX)
            pre[r][u] = Jf[(size_t)row * ld + qc];
          else
            pre[r][u] = J[(size_t)row * ld + qc];
        }
    };
    load_rows(q0);
    for (int c0 = q0; c0 < q1; c0 += kNeqChunk) {
      const int cn = min(kNeqChunk, q1 - c0);
      const int cn4 = (cn + 3) & ~3;
      // lane = point of the chunk (kNeqChunk = 64): per-point constants,
      // then the renormalised rows, one wave-uniform row jf at a time
      // (coalesced over points).  mrqcof.c:92-104 order:
      // coef = sig*np/ave; dyda = coef*(dyda - ymod*dave[l]/ave); ymod *= coef.
      {
        const int pp = lane, q = c0 + pp;
        const bool live = pp < cn;
        double ymr = 0, coef = 0;
        if (live) {
          ymr = J[(size_t)ma * ld + q];
          double sg, wg, yo;
          if (q < P.n_points) {
            const int pt = perm ? perm[q] : q;  // Jacobian column q holds point pt
            sg = P.sig[pt];
            wg = P.w[pt];
            yo = P.y[pt];
          } else {
            sg = 1 / conw_r;
            wg = 1;
            yo = 0;
          }
          if (inrel) coef = sg * np / dave[ma];
          if (wave == 0) {
            aux[pp * 3 + 0] = 1 / (sg * sg);
            aux[pp * 3 + 1] = wg;
            aux[pp * 3 + 2] = yo - (inrel ? coef * ymr : ymr);
          }
        } else if (wave == 0 && pp < cn4) {
          aux[pp * 3 + 0] = 0;
          aux[pp * 3 + 1] = 0;
          aux[pp * 3 + 2] = 0;
        }
        // eight rows per batch, from the rows loaded before the chunk
#pragma unroll
        for (int r = 0; r < kRounds; r++) {
          const int jf0 = wave + 8 * kWaves * r;
          if (jf0 >= MP) break;
          double dv[8];
          int lv[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int jf = jf0 + kWaves * u;
            lv[u] = fidx[jf];
            dv[u] = (jf < mfit && live) ? (double)pre[r][u] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int jf = jf0 + kWaves * u;
            if (jf >= MP) break;
            double d = dv[u];
            if (inrel && jf < mfit && live) {
              const int l = lv[u];
              d = (l == 0) ? 0.0 : coef * (d - ymr * drat[l]);
            }
            if (pp < cn4) Jn[pp * MS + jf] = (JnT)d;
          }
        }
      }
      if (c0 + kNeqChunk < q1) load_rows(c0 + kNeqChunk);  // in flight during this chunk's sums
      __syncthreads();
      for (int kk = 0; kk < cn4 / 4; kk++) {
        const int row = 4 * kk + (lane >> 4);
        const JnT s2 = (JnT)aux[row * 3], wg = (JnT)aux[row * 3 + 1];
#pragma unroll
        for (int i = 0; i < NT; i++) {
          // a wave's tiles past the triangle multiply zeros (no branch, the
          // accumulators stay put); they are never stored
          const bool act = wave + 4 * i < ntile;
          const int cj = act ? tj[i] : 0, ck = act ? tk[i] : 0;
          const JnT av = Jn[row * MS + 16 * cj + (lane & 15)] * s2 * wg;
          const JnT a = act ? av : (JnT)0;
          const JnT b = Jn[row * MS + 16 * ck + (lane & 15)];
          if constexpr (MIX)
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
          else
            acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
        }
      }
      // beta and chi2 in point order; four points' LDS reads are issued
      // before their (sequential) sums
      if (tid < mfit) {
        int pp = 0;
        for (; pp + 4 <= cn; pp += 4) {
          double wt[4], dy[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            wt[u] = (double)Jn[(pp + u) * MS + tid] * aux[(pp + u) * 3] * aux[(pp + u) * 3 + 1];
            dy[u] = aux[(pp + u) * 3 + 2];
          }
#pragma unroll
          for (int u = 0; u < 4; u++) bacc = bacc + dy[u] * wt[u];
        }
        for (; pp < cn; pp++) {
          const double wt = (double)Jn[pp * MS + tid] * aux[pp * 3] * aux[pp * 3 + 1];
          bacc = bacc + aux[pp * 3 + 2] * wt;
        }
      }
      if (tid == kThreads - 1) {
        int pp = 0;
        for (; pp + 4 <= cn; pp += 4) {
          double t[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const double dy = aux[(pp + u) * 3 + 2];
            t[u] = dy * dy * aux[(pp + u) * 3] * aux[(pp + u) * 3 + 1];
          }
#pragma unroll
          for (int u = 0; u < 4; u++) chi = chi + t[u];
        }
        for (; pp < cn; pp++) {
          const double dy = aux[pp * 3 + 2];
          chi = chi + dy * dy * aux[pp * 3] * aux[pp * 3 + 1];
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < NT; i++) {
    if (wave + 4 * i < ntile) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        // result rows: FP64 16x16x4 D[4r + l/16][l%16]; FP32 D[4(l/16) + r][l%16]
        const int j = 16 * tj[i] + (MIX ? 4 * (lane >> 4) + r : 4 * r + (lane >> 4));
        const int k = 16 * tk[i] + (lane & 15);
        if (j < mfit && k < mfit && j >= k) {
          alpha[j * mfit + k] = (double)acc[i][r];
          alpha[k * mfit + j] = (double)acc[i][r];
        }
      }
    }
  }
  if (tid < mfit) beta[tid] = bacc;
  if (tid == kThreads - 1) *chi_out = chi;
}

// Mixed-precision normal equations with the points split over the waves
// (mrqcof.c:42-139).  Lane (c, q) = (lane % 16, lane / 16) owns free
// parameter 16*b + c of every column block b and the points 4q..4q+3 of a
// 16-point group, whose FP32 derivative rows it reads straight from the
// Jacobian as one 16-byte load per block: no staging through LDS.  Step s of
// a group pairs the four lanes' point 4q+s, so one v_mfma_f32_16x16x4f32 per
// (lower-triangle tile, s) sums four distinct points and the four steps sum
// all sixteen.  Every wave accumulates all NB(NB+1)/2 tiles over its own
// groups (group g goes to wave g % 4 within each light curve); the four
// waves' tiles are added in wave order at the end.  The loads of a wave's
// next group are issued before its current group is computed.
// Per-point values are computed in FP64 once per point, by lane
// 16q + s for point 4q + s: the renormalisation factor coef = sig*np/ave,
// the renormalised model value coef*ymod, the residual dy and the weight
// w/sig^2; chi2 sums dy^2 w/sig^2 there.  They reach the other lanes of the
// row as FP32 by a DPP row broadcast.
// A relative light curve takes two passes: its column sums ave and dave
// (FP32 within a group of 16 points, FP64 across groups), then the rows
// renormalised in FP32 as coef*dyda - (coef*ymod)*(dave/ave), with the
// size-coefficient column zeroed.  A group that straddles two light curves
// is read once for each, with the other curve's points masked.
// beta is FP64 across groups: each lane sums its four points' Jn*w*dy/sig^2
// per column in FP32, adds that to its FP64 sum, and the lanes of a column
// and the waves are added in a fixed order at the end, so the result does
// not depend on timing.
// LDS (doubles): [4][16*NB+1] column sums, [4][16*NB] beta, [4] chi2, then
// 4 tiles x 4 waves x 256 FP32 partial tiles.
__host__ __device__ constexpr int neq_ks_doubles(int nb) { return 4 * (16 * nb + 1) + 4 * 16 * nb + 4 + 2048; }

// v broadcast from lane i of its row of 16 lanes (DPP row_newbcast)
__device__ __forceinline__ float row_bcast(float v, int i) {
  const int x = __builtin_bit_cast(int, v);
  int r;
  switch (i) {
    case 0: r = __builtin_amdgcn_update_dpp(0, x, 0x150, 0xF, 0xF, false); break;
    case 1: r = __builtin_amdgcn_update_dpp(0, x, 0x151, 0xF, 0xF, false); break;
    case 2: r = __builtin_amdgcn_update_dpp(0, x, 0x152, 0xF, 0xF, false); break;
    default: r = __builtin_amdgcn_update_dpp(0, x, 0x153, 0xF, 0xF, false); break;
  }
  return __builtin_bit_cast(float, r);
}

typedef short v4s __attribute__((ext_vector_type(4)));  // four bf16 (bit patterns) of one MFMA operand
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
// (a, b) rounded to bf16 (nearest even) and packed, a in the low half: one
// v_cvt_pk_bf16_f32.
__device__ inline unsigned bf16_pair(float a, float b) {
  const v2bf h = __builtin_convertvector(v2f{a, b}, v2bf);
  return __builtin_bit_cast(unsigned, h);
}

// ABF (PS_PREC_BF16A): alpha's products from the renormalised rows rounded
// to bf16, one v_mfma_f32_16x16x16_bf16 per tile and 16-point group (A lane
// l: row l%16, k = 4(l/16) + e, the lane's own four points) instead of four
// v_mfma_f32_16x16x4f32; beta and chi2 keep their FP32/FP64 sums.
template <int NB, bool ABF = false>
__device__ __forceinline__ void normal_equations_ks(const DevProblem &P, const double *J, const double *__restrict__ yod,
                                                    const double *__restrict__ sgd, const double *__restrict__ swd,
                                                    double conw_r, double *ks, double *dave, double *drat,
                                                    const int *fidx, double *beta, double *chi_out, double *alpha) {
#pragma clang fp contract(off)
  constexpr int NT = NB * (NB + 1) / 2, NC = 16 * NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c16 = lane & 15, kq = lane >> 4, i4 = c16 & 3;
  const int ld = P.npad, mfit = P.mfit, ma = P.ma, npts = P.n_points;
  const float *Jf = reinterpret_cast<const float *>(J);
  const double *ym = J + (size_t)ma * ld;  // FP64 model row
  double *csum = ks, *bsum = ks + 4 * (NC + 1), *csq = bsum + 4 * NC;
  float *tred = reinterpret_cast<float *>(csq + 4);
  const double sconv = 1 / conw_r, swconv = 1 / (sconv * sconv);  // convexity rows: sigma, weight/sigma^2
  // this lane's column of each block: offset of its Jacobian row (free
  // parameter 16*b + c; padding columns read row 0 and land in discarded
  // rows and columns of alpha and beta), and whether it is the size
  // coefficient (zeroed in relative light curves)
  int roff[NB];
  bool l0[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int jf = 16 * b + c16;
    const int l = jf < mfit ? fidx[jf] : 0;
    roff[b] = l * ld + 4 * kq;
    l0[b] = l == 0;
  }
  v4f acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = v4f{0, 0, 0, 0};
  double bet[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) bet[b] = 0;
  double chi = 0;
  // one group's loads: the rows of this lane's four points, and the model
  // value and data of its own point 4q + c%4 (the convexity rows and the
  // padding past them read the last real point's data, replaced at use)
  struct Grp {
    v4f raw[NB];
    double ymq, sgq, swq, yoq;
  };
  auto load = [&](int g, Grp &G, bool pts) __attribute__((always_inline)) {
#pragma unroll
    for (int b = 0; b < NB; b++) G.raw[b] = *reinterpret_cast<const v4f *>(Jf + roff[b] + 16 * g);
    const int q = 16 * g + 4 * kq + i4, qc = min(q, npts - 1);
    G.ymq = ym[q];
    if (pts) {
      G.sgq = sgd[qc];
      G.swq = swd[qc];
      G.yoq = yod[qc];
    }
  };
  for (int lc = 0; lc <= P.n_lc; lc++) {
    const int q0 = P.lc_start[lc], q1 = P.lc_start[lc + 1], np = q1 - q0;
    const bool inrel = P.lc_inrel[lc] != 0;
    const bool z0 = inrel && l0[0];  // this lane's block-0 column is the size coefficient, zeroed here
    const int g0 = q0 >> 4, g1 = (q1 + 15) >> 4;
    double rave = 0;  // np / ave
    float dr[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) dr[b] = 0;
    if (inrel) {
      // pass 1: ave = sum ymod, dave = sum dyda over the light curve
      double sy = 0, sd[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) sd[b] = 0;
      Grp cur, nxt;
      int g = g0 + wave;
      if (g < g1) load(g, nxt, false);
      for (; g < g1; g += kWaves) {
        cur = nxt;
        if (g + kWaves < g1) load(g + kWaves, nxt, false);
        const int pb = 16 * g + 4 * kq;
        {
          const int q = pb + i4;
          sy = sy + ((c16 < 4 && q >= q0 && q < q1) ? cur.ymq : 0.0);
        }
        if (16 * g >= q0 && 16 * g + 16 <= q1) {  // wave-uniform: every point of the group in the curve
#pragma unroll
          for (int b = 0; b < NB; b++)
            sd[b] = sd[b] + (double)((cur.raw[b][0] + cur.raw[b][1]) + (cur.raw[b][2] + cur.raw[b][3]));
        } else {
          bool lv[4];
#pragma unroll
          for (int i = 0; i < 4; i++) lv[i] = pb + i >= q0 && pb + i < q1;
#pragma unroll
          for (int b = 0; b < NB; b++) {
            const float s = ((lv[0] ? cur.raw[b][0] : 0.0f) + (lv[1] ? cur.raw[b][1] : 0.0f)) +
                            ((lv[2] ? cur.raw[b][2] : 0.0f) + (lv[3] ? cur.raw[b][3] : 0.0f));
            sd[b] = sd[b] + (double)s;
          }
        }
      }
      // lanes of one column, then waves, in a fixed order
#pragma unroll
      for (int b = 0; b < NB; b++) {
        sd[b] = sd[b] + __shfl_xor(sd[b], 16, 64);
        sd[b] = sd[b] + __shfl_xor(sd[b], 32, 64);
      }
      // the model-value sum: lanes c < 4 of each row hold its points
      sy = sy + __shfl_xor(sy, 1, 64);
      sy = sy + __shfl_xor(sy, 2, 64);
      sy = sy + __shfl_xor(sy, 16, 64);
      sy = sy + __shfl_xor(sy, 32, 64);
      if (kq == 0) {
#pragma unroll
        for (int b = 0; b < NB; b++) csum[wave * (NC + 1) + 16 * b + c16] = sd[b];
        if (c16 == 0) csum[wave * (NC + 1) + NC] = sy;
      }
      __syncthreads();
      if (tid <= NC) {
        double s = csum[tid];
        for (int w = 1; w < kWaves; w++) s = s + csum[w * (NC + 1) + tid];
        dave[tid] = s;
      }
      __syncthreads();
      const double ave = dave[NC];
      rave = np / ave;
      if (tid < NC) drat[tid] = dave[tid] / ave;
      __syncthreads();
#pragma unroll
      for (int b = 0; b < NB; b++) dr[b] = (float)drat[16 * b + c16];
    }
    // pass 2: renormalised rows, normal equations, beta and chi2
    Grp cur, nxt;
    int g = g0 + wave;
    if (g < g1) load(g, nxt, true);
    for (; g < g1; g += kWaves) {
      cur = nxt;
      if (g + kWaves < g1) load(g + kWaves, nxt, true);
      const int pb = 16 * g + 4 * kq;
      const bool full = 16 * g >= q0 && 16 * g + 16 <= q1;  // wave-uniform
      // this lane's own point (FP64): renormalisation factor and model
      // value, residual, weight; chi2
      float coef_f, cy_f, sw_f, swdy_f;
      {
        const int q = pb + i4;
        const bool live = q >= q0 && q < q1, real = q < npts;
        const double sg = real ? cur.sgq : sconv, sw = live ? (real ? cur.swq : swconv) : 0.0;
        const double yo = real ? cur.yoq : 0.0;
        const double coef = inrel ? sg * rave : 1.0;
        const double ymr = live ? cur.ymq : 0.0;
        const double cy = inrel ? coef * ymr : ymr;
        const double dy = live ? yo - cy : 0.0;
        if (c16 < 4) chi = chi + dy * dy * sw;
        coef_f = (float)coef;
        cy_f = inrel ? (float)cy : 0.0f;  // zero in absolute curves: the renormalisation below returns the row
        sw_f = (float)sw;
        swdy_f = (float)(sw * dy);
      }
      float bp[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) bp[b] = 0.0f;
      // ABF: the group's four points per block as bf16, each pair packed as
      // soon as both points are formed
      unsigned apk[ABF ? NB : 1][2], bpk[ABF ? NB : 1][2];
      float ae[ABF ? NB : 1], be[ABF ? NB : 1];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float cf = row_bcast(coef_f, i), cyi = row_bcast(cy_f, i), swi = row_bcast(sw_f, i),
                    swdyi = row_bcast(swdy_f, i);
        const bool live = full || (pb + i >= q0 && pb + i < q1);
        float av[NB], bv[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
          float jn = cur.raw[b][i];
          // coef*dyda - (coef*ymod)*(dave/ave), the first product fused; coef
          // 1, coef*ymod 0 and dave/ave 0 in absolute curves: jn + (-0) is jn
          // exactly
          jn = fmaf(cf, jn, -(cyi * dr[b]));
          // the size coefficient, zeroed in relative curves, can only be free
          // parameter 0 (free parameters in increasing order; the padding
          // columns past mfit are discarded)
          if (b == 0) jn = z0 ? 0.0f : jn;
          if (!full) jn = live ? jn : 0.0f;
          bp[b] = fmaf(jn, swdyi, bp[b]);
          bv[b] = jn;
          av[b] = jn * swi;
          if constexpr (ABF) {
            if (i & 1) {
              apk[b][i >> 1] = bf16_pair(ae[b], av[b]);
              bpk[b][i >> 1] = bf16_pair(be[b], bv[b]);
            } else {
              ae[b] = av[b];
              be[b] = bv[b];
            }
          }
        }
        if constexpr (!ABF)
          static_for<0, NB>([&](auto jbc) __attribute__((always_inline)) {
            constexpr int jb = decltype(jbc)::value;
            static_for<0, jb + 1>([&](auto kbc) __attribute__((always_inline)) {
              constexpr int kb = decltype(kbc)::value;
              constexpr int t = jb * (jb + 1) / 2 + kb;
              acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[jb], bv[kb], acc[t], 0, 0, 0);
            });
          });
      }
      if constexpr (ABF) {
        v4s ap[NB], bq[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
          ap[b] = __builtin_bit_cast(v4s, make_uint2(apk[b][0], apk[b][1]));
          bq[b] = __builtin_bit_cast(v4s, make_uint2(bpk[b][0], bpk[b][1]));
        }
        static_for<0, NB>([&](auto jbc) __attribute__((always_inline)) {
          constexpr int jb = decltype(jbc)::value;
          static_for<0, jb + 1>([&](auto kbc) __attribute__((always_inline)) {
            constexpr int kb = decltype(kbc)::value;
            constexpr int t = jb * (jb + 1) / 2 + kb;
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ap[jb], bq[kb], acc[t], 0, 0, 0);
          });
        });
      }
#pragma unroll
      for (int b = 0; b < NB; b++) bet[b] = bet[b] + (double)bp[b];
    }
  }
  // beta: lanes of a column, then the waves in order
#pragma unroll
  for (int b = 0; b < NB; b++) {
    bet[b] = bet[b] + __shfl_xor(bet[b], 16, 64);
    bet[b] = bet[b] + __shfl_xor(bet[b], 32, 64);
    if (kq == 0) bsum[wave * NC + 16 * b + c16] = bet[b];
  }
  // chi2: lanes c < 4 of each row hold its points' terms
  chi = chi + __shfl_xor(chi, 1, 64);
  chi = chi + __shfl_xor(chi, 2, 64);
  chi = chi + __shfl_xor(chi, 16, 64);
  chi = chi + __shfl_xor(chi, 32, 64);
  if (lane == 0) csq[wave] = chi;
  __syncthreads();
  if (tid < mfit) {
    double s = bsum[tid];
    for (int w = 1; w < kWaves; w++) s = s + bsum[w * NC + tid];
    beta[tid] = s;
  }
  if (tid == 0) *chi_out = ((csq[0] + csq[1]) + csq[2]) + csq[3];
  // alpha: four tiles at a time through LDS, each tile summed over the
  // waves in order by one wave; j >= k entries stored and mirrored.
  // FP32 D of 16x16x4: D[4(l/16) + r][l%16]
  static_for<0, (NT + 3) / 4>([&](auto rc) __attribute__((always_inline)) {
    constexpr int R = decltype(rc)::value;
    static_for<0, 4>([&](auto uc) __attribute__((always_inline)) {
      constexpr int u = decltype(uc)::value, t = 4 * R + u;
      if constexpr (t < NT) {
#pragma unroll
        for (int r = 0; r < 4; r++) tred[((u * kWaves + wave) * 4 + r) * 64 + lane] = acc[t][r];
      }
    });
    __syncthreads();
    const int t = 4 * R + wave;
    if (t < NT) {
      int jb = 0;
      while ((jb + 1) * (jb + 2) / 2 <= t) jb++;
      const int kb = t - jb * (jb + 1) / 2;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        double s = tred[((wave * kWaves + 0) * 4 + r) * 64 + lane];
        for (int w = 1; w < kWaves; w++) s = s + (double)tred[((wave * kWaves + w) * 4 + r) * 64 + lane];
        const int j = 16 * jb + 4 * kq + r, k = 16 * kb + c16;
        if (j < mfit && k < mfit && j >= k) {
          alpha[j * mfit + k] = s;
          alpha[k * mfit + j] = s;
        }
      }
    }
    __syncthreads();
  });
}

// FP64 normal equations with the points split over the waves
// (mrqcof.c:42-139): normal_equations_ks's layout and order of work in FP64,
// one v_mfma_f64_16x16x4f64 per (lower-triangle tile, step) summing four
// distinct points, no staging through LDS.  Lane (c, q) reads the FP64 rows
// of free parameters 16*b + c at points 4q..4q+3 of a 16-point group as one
// 32-byte load per block; the next group's loads are issued before the
// current group is computed.  Each per-point value is formed in FP64 with
// the reference's operations, as the staged kernel forms them:
//   coef = sig*np/ave; Jn = coef*(dyda - ymod*(dave/ave)); dy = y - coef*ymod;
//   wt = Jn*sig2i*w; alpha += wt*Jn'; beta += dy*wt; chi2 += dy*dy*sig2i*w,
// by lane 16q + s for point 4q + s, and reaches the other lanes of its row
// by DPP row broadcasts.  Only the order of the sums differs from the
// reference: per lane over its points, the lanes of a column and the waves
// in a fixed order, so the result does not depend on timing.
// perm: the fit's point order (Jacobian column q holds point perm[q]), or
// null.  LDS (doubles): [4][16*NB+1] column sums, [4][16*NB] beta, [4] chi2,
// then 4 tiles x 4 waves x 256 FP64 partial tiles.
__host__ __device__ constexpr int neq_ks64_doubles(int nb) { return 4 * (16 * nb + 1) + 4 * 16 * nb + 4 + 4096; }

// v broadcast from lane i of its row of 16 lanes (two DPP row_newbcast)
__device__ __forceinline__ double row_bcast_d(double v, int i) {
  const float lo = __builtin_bit_cast(float, __double2loint(v)), hi = __builtin_bit_cast(float, __double2hiint(v));
  return __hiloint2double(__builtin_bit_cast(int, row_bcast(hi, i)), __builtin_bit_cast(int, row_bcast(lo, i)));
}

template <int NB>
__device__ __forceinline__ void normal_equations_ks64(const DevProblem &P, const double *J, const int *perm,
                                                      double conw_r, double *ks, double *dave, double *drat,
                                                      const int *fidx, double *beta, double *chi_out, double *alpha) {
#pragma clang fp contract(off)
  constexpr int NT = NB * (NB + 1) / 2, NC = 16 * NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c16 = lane & 15, kq = lane >> 4, i4 = c16 & 3;
  const int ld = P.npad, mfit = P.mfit, ma = P.ma, npts = P.n_points;
  const double *ym = J + (size_t)ma * ld;  // model row
  double *csum = ks, *bsum = ks + 4 * (NC + 1), *csq = bsum + 4 * NC;
  double *tred = csq + 4;
  const double sconv = 1 / conw_r;  // convexity rows: sigma (weight 1, brightness 0)
  int roff[NB];
  bool l0[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int jf = 16 * b + c16;
    const int l = jf < mfit ? fidx[jf] : 0;
    roff[b] = l * ld + 4 * kq;
    l0[b] = l == 0;
  }
  v4d acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = v4d{0, 0, 0, 0};
  double bet[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) bet[b] = 0;
  double chi = 0;
  // one group's loads: the rows of this lane's four points, and the model
  // value and data of its own point 4q + c%4 (the convexity rows and the
  // padding past them read the last real point's data, replaced at use)
  struct Grp {
    v4d raw[NB];
    double ymq, sgq, wq, yoq;
  };
  auto load = [&](int g, Grp &G, bool pts) __attribute__((always_inline)) {
#pragma unroll
    for (int b = 0; b < NB; b++) G.raw[b] = *reinterpret_cast<const v4d *>(J + roff[b] + 16 * g);
    const int q = 16 * g + 4 * kq + i4, qc = min(q, npts - 1);
    G.ymq = ym[q];
    if (pts) {
      const int pt = perm ? perm[qc] : qc;
      G.sgq = P.sig[pt];
      G.wq = P.w[pt];
      G.yoq = P.y[pt];
    }
  };
  for (int lc = 0; lc <= P.n_lc; lc++) {
    const int q0 = P.lc_start[lc], q1 = P.lc_start[lc + 1], np = q1 - q0;
    const bool inrel = P.lc_inrel[lc] != 0;
    const bool z0 = inrel && l0[0];  // this lane's block-0 column is the size coefficient, zeroed here
    const int g0 = q0 >> 4, g1 = (q1 + 15) >> 4;
    double ave = 1;
    double dr[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) dr[b] = 0;
    if (inrel) {
      // pass 1: ave = sum ymod, dave = sum dyda over the light curve
      double sy = 0, sd[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) sd[b] = 0;
      Grp cur, nxt;
      int g = g0 + wave;
      if (g < g1) load(g, nxt, false);
      for (; g < g1; g += kWaves) {
        cur = nxt;
        if (g + kWaves < g1) load(g + kWaves, nxt, false);
        const int pb = 16 * g + 4 * kq;
        {
          const int q = pb + i4;
          sy = sy + ((c16 < 4 && q >= q0 && q < q1) ? cur.ymq : 0.0);
        }
        if (16 * g >= q0 && 16 * g + 16 <= q1) {  // wave-uniform: every point of the group in the curve
#pragma unroll
          for (int b = 0; b < NB; b++) sd[b] = sd[b] + ((cur.raw[b][0] + cur.raw[b][1]) + (cur.raw[b][2] + cur.raw[b][3]));
        } else {
          bool lv[4];
#pragma unroll
          for (int i = 0; i < 4; i++) lv[i] = pb + i >= q0 && pb + i < q1;
#pragma unroll
          for (int b = 0; b < NB; b++) {
            const double s = ((lv[0] ? cur.raw[b][0] : 0.0) + (lv[1] ? cur.raw[b][1] : 0.0)) +
                             ((lv[2] ? cur.raw[b][2] : 0.0) + (lv[3] ? cur.raw[b][3] : 0.0));
            sd[b] = sd[b] + s;
          }
        }
      }
      // lanes of one column, then waves, in a fixed order
#pragma unroll
      for (int b = 0; b < NB; b++) {
        sd[b] = sd[b] + __shfl_xor(sd[b], 16, 64);
        sd[b] = sd[b] + __shfl_xor(sd[b], 32, 64);
      }
      // the model-value sum: lanes c < 4 of each row hold its points
      sy = sy + __shfl_xor(sy, 1, 64);
      sy = sy + __shfl_xor(sy, 2, 64);
      sy = sy + __shfl_xor(sy, 16, 64);
      sy = sy + __shfl_xor(sy, 32, 64);
      if (kq == 0) {
#pragma unroll
        for (int b = 0; b < NB; b++) csum[wave * (NC + 1) + 16 * b + c16] = sd[b];
        if (c16 == 0) csum[wave * (NC + 1) + NC] = sy;
      }
      __syncthreads();
      if (tid <= NC) {
        double s = csum[tid];
        for (int w = 1; w < kWaves; w++) s = s + csum[w * (NC + 1) + tid];
        dave[tid] = s;
      }
      __syncthreads();
      ave = dave[NC];
      if (tid < NC) drat[tid] = dave[tid] / ave;
      __syncthreads();
#pragma unroll
      for (int b = 0; b < NB; b++) dr[b] = drat[16 * b + c16];
    }
    // pass 2: renormalised rows, normal equations, beta and chi2
    Grp cur, nxt;
    int g = g0 + wave;
    if (g < g1) load(g, nxt, true);
    for (; g < g1; g += kWaves) {
      cur = nxt;
      if (g + kWaves < g1) load(g + kWaves, nxt, true);
      const int pb = 16 * g + 4 * kq;
      const bool full = 16 * g >= q0 && 16 * g + 16 <= q1;  // wave-uniform
      // this lane's own point: renormalisation factor, model value,
      // residual, 1/sigma^2 and weight (zero weight outside the curve); chi2
      double coef, ymr, s2, wg, dy;
      {
        const int q = pb + i4;
        const bool live = q >= q0 && q < q1, real = q < npts;
        const double sg = real ? cur.sgq : sconv;
        wg = live ? (real ? cur.wq : 1.0) : 0.0;
        const double yo = real ? cur.yoq : 0.0;
        coef = inrel ? sg * np / ave : 1.0;
        ymr = live ? cur.ymq : 0.0;
        dy = live ? yo - (inrel ? coef * ymr : ymr) : 0.0;
        s2 = 1 / (sg * sg);
        if (c16 < 4) chi = chi + dy * dy * s2 * wg;
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const double cf = row_bcast_d(coef, i), ymi = row_bcast_d(inrel ? ymr : 0.0, i);
        const double s2i = row_bcast_d(s2, i), wgi = row_bcast_d(wg, i), dyi = row_bcast_d(dy, i);
        const bool live = full || (pb + i >= q0 && pb + i < q1);
        double av[NB], bv[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
          double jn = cur.raw[b][i];
          // coef 1, ymod 0 and dave/ave 0 in absolute curves: 1*(jn - 0*0) is jn exactly
          jn = cf * (jn - ymi * dr[b]);
          // the size coefficient, zeroed in relative curves, can only be free
          // parameter 0 (free parameters in increasing order; the padding
          // columns past mfit are discarded)
          if (b == 0) jn = z0 ? 0.0 : jn;
          if (!full) jn = live ? jn : 0.0;
          const double wt = jn * s2i * wgi;
          bet[b] = bet[b] + dyi * wt;
          bv[b] = jn;
          av[b] = wt;
        }
        static_for<0, NB>([&](auto jbc) __attribute__((always_inline)) {
          constexpr int jb = decltype(jbc)::value;
          static_for<0, jb + 1>([&](auto kbc) __attribute__((always_inline)) {
            constexpr int kb = decltype(kbc)::value;
            constexpr int t = jb * (jb + 1) / 2 + kb;
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[jb], bv[kb], acc[t], 0, 0, 0);
          });
        });
      }
    }
  }
  // beta: lanes of a column, then the waves in order
#pragma unroll
  for (int b = 0; b < NB; b++) {
    bet[b] = bet[b] + __shfl_xor(bet[b], 16, 64);
    bet[b] = bet[b] + __shfl_xor(bet[b], 32, 64);
    if (kq == 0) bsum[wave * NC + 16 * b + c16] = bet[b];
  }
  // chi2: lanes c < 4 of each row hold its points' terms
  chi = chi + __shfl_xor(chi, 1, 64);
  chi = chi + __shfl_xor(chi, 2, 64);
  chi = chi + __shfl_xor(chi, 16, 64);
  chi = chi + __shfl_xor(chi, 32, 64);
  if (lane == 0) csq[wave] = chi;
  __syncthreads();
  if (tid < mfit) {
    double s = bsum[tid];
    for (int w = 1; w < kWaves; w++) s = s + bsum[w * NC + tid];
    beta[tid] = s;
  }
  if (tid == 0) *chi_out = ((csq[0] + csq[1]) + csq[2]) + csq[3];
  // alpha: four tiles at a time through LDS, each tile summed over the
  // waves in order by one wave; j >= k entries stored and mirrored.
  // FP64 D of 16x16x4: D[4r + l/16][l%16]
  static_for<0, (NT + 3) / 4>([&](auto rc) __attribute__((always_inline)) {
    constexpr int R = decltype(rc)::value;
    static_for<0, 4>([&](auto uc) __attribute__((always_inline)) {
      constexpr int u = decltype(uc)::value, t = 4 * R + u;
      if constexpr (t < NT) {
#pragma unroll
        for (int r = 0; r < 4; r++) tred[((u * kWaves + wave) * 4 + r) * 64 + lane] = acc[t][r];
      }
    });
    __syncthreads();
    const int t = 4 * R + wave;
    if (t < NT) {
      int jb = 0;
      while ((jb + 1) * (jb + 2) / 2 <= t) jb++;
      const int kb = t - jb * (jb + 1) / 2;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        double s = tred[((wave * kWaves + 0) * 4 + r) * 64 + lane];
        for (int w = 1; w < kWaves; w++) s = s + tred[((wave * kWaves + w) * 4 + r) * 64 + lane];
        const int j = 16 * jb + 4 * r + kq, k = 16 * kb + c16;
        if (j < mfit && k < mfit && j >= k) {
          alpha[j * mfit + k] = s;
          alpha[k * mfit + j] = s;
        }
      }
    }
    __syncthreads();
  });
}

// Scratch at the start of the normal-equation kernel's LDS: the convexity
// rows' partial sums, then the staged rows (ks = 0) or the K-split
// reductions (ks = column blocks).
__host__ __device__ inline int neq_scratch_doubles(int mfit, bool mix, int ks) {
  const int d = neq_jn_doubles(mfit, mix);
  const int k = ks > 0 ? (mix ? neq_ks_doubles(ks) : neq_ks64_doubles(ks)) : 0;
  return k > d ? k : d;
}

// Fit initialisation (period_search_BOINC.cpp:819-844): shape from the
// ellipsoid, pole m, omega from the trial frequency, phase and scattering
// parameters.
__device__ void init_params(const DevProblem &P, double freq, int m, double *a) {
#pragma clang fp contract(off)
  const int nc = P.ncoef;
  for (int l = threadIdx.x; l < P.ma; l += blockDim.x) {
    double v;
    if (l < nc)
      v = P.cg_first[l];
    else if (l == nc)
      v = (kPi / 180) * (90 - P.pole_beta[m]);
    else if (l == nc + 1)
      v = (kPi / 180) * P.pole_lambda[m];
    else if (l == nc + 2)
      v = 24 * 2 * kPi / (1 / freq);
    else if (l == nc + 3)
      v = P.par0;
    else if (l == nc + 4)
      v = P.par1;
    else if (l == nc + 5)
      v = P.par2;
    else if (l == nc + 6)
      v = P.log_cl;
    else
      v = 1;
    a[l] = v;
  }
}

// ---------------------------------------------------------------------
// Gauss-Jordan elimination with full pivoting (period_search/gauss_errc.c:9-73)
// by one 256-thread workgroup on the augmented matrix [A | b] (n x (n+1),
// row-major, b as column n) held twice in LDS: each pivot step reads buffer
// i&1 and writes buffer (i+1)&1, so a step needs one barrier.  Thread
// (wave w, lane l) owns columns l, l+64, ... and rows w, w+4, ...  While it
// eliminates, every thread also scans its new entries for the next pivot
// (|a| over unpivoted rows and columns; ties go to the last entry in
// row-major order, fabs(a) >= big); the per-wave winners meet in LDS.
// Only the solution b is needed, so pivoted columns are not updated (the
// reference's in-place inverse); every value b depends on is computed with
// the reference's operations in the reference's order:
//   a[icol][l] *= pivinv; b[icol] *= pivinv;
//   a[ll][l] -= a[icol][l]*dum; b[ll] -= b[icol]*dum  (dum = a[ll][icol]).
// Rows irow and icol are swapped by reading the old buffer through the
// permutation.  Returns 0, 1 (singular) or 2 (zero pivot), uniform across the
// workgroup; on 0 the solution is in x (LDS or global, n entries).
// LDS: 2 * n * (n+1) + 16 doubles from buf.
// INPLACE (for n whose two copies exceed the LDS; up to 128 free parameters,
// lmax = mmax = 10): one copy, n * (n+1) + 16 doubles, updated in place.  In a
// pivot step the only entries that one thread writes and others read are
// those of rows irow and icol (the pivot row and its swap partner; column
// icol is pivoted and never written again), so every thread first reads
// those two rows at its columns into registers, and a barrier separates
// these reads from the writes.  The arithmetic and its order are the same.
__device__ inline void pivot_max(double &bv, int &be, double ov, int oe) {
  const bool take = (ov > bv) | ((ov == bv) & (oe > be));  // no short-circuit: selects, not branches
  bv = take ? ov : bv;
  be = take ? oe : be;
}

constexpr size_t kSolveLdsDoubles = (160 * 1024 - 1024) / sizeof(double);  // dynamic LDS of the step kernel
constexpr int kSolveCols = (kMaxMfit + 1 + 63) / 64;                      // columns of [A|b] per lane
__host__ __device__ inline bool solve_in_place(int n) {
  return 2 * (size_t)n * (n + 1) + 4 * kWaves > kSolveLdsDoubles;
}

template <bool INPLACE>
__device__ __forceinline__ int gauss_solve(double *buf, int n, double *x) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ld = n + 1, nn = n * ld;
  double *red = buf + (INPLACE ? 1 : 2) * nn;  // [2][kWaves] x (value, index)
  unsigned long long done[2] = {0ull, 0ull};  // pivoted columns (ipiv != 0) = pivoted rows
  auto is_done = [&](int k) { return (done[k >> 6] >> (k & 63)) & 1ull; };
  // pivot search over the initial matrix
  {
    double bv = -1.0;
    int be = -1;
    for (int j = wave; j < n; j += kWaves)
      for (int c = lane; c < n; c += 64) {
        const double v = fabs(buf[j * ld + c]);
        if (v >= bv) {
          bv = v;
          be = j * n + c;
        }
      }
    for (int o = 32; o > 0; o >>= 1) pivot_max(bv, be, __shfl_xor(bv, o, 64), __shfl_xor(be, o, 64));
    if (lane == 0) {
      red[2 * wave] = bv;
      red[2 * wave + 1] = (double)be;
    }
  }
  __syncthreads();
  int prev_irow = -1, prev_icol = -1;
  for (int i = 0; i < n; i++) {
    const double *Ao = INPLACE ? buf : buf + (i & 1) * nn;
    double *An = INPLACE ? buf : buf + ((i + 1) & 1) * nn;
    const double *rd = red + (i & 1) * 2 * kWaves;
    double bv = rd[0];
    int be = (int)rd[1];
    for (int w = 1; w < kWaves; w++) pivot_max(bv, be, rd[2 * w], (int)rd[2 * w + 1]);
    int irow, icol;
    if (be < 0) {  // no candidate (NaN matrix): the reference keeps the previous pivot
      if (prev_icol < 0) return 1;
      irow = prev_irow;
      icol = prev_icol;
    } else {
      irow = be / n;
      icol = be - irow * n;
    }
    if (is_done(icol)) return 1;  // column pivoted twice: the next scan reports singular
    prev_irow = irow;
    prev_icol = icol;
    const double piv = Ao[irow * ld + icol];  // a[icol][icol] after the swap
    if (piv == 0.0) return 2;
    const double pivinv = 1.0 / piv;
    done[icol >> 6] |= 1ull << (icol & 63);
    double nbv = -1.0;
    int nbe = -1;
    if constexpr (INPLACE) {
      // rows irow (pivot row) and icol (its swap partner) at this lane's
      // columns, read before any thread writes them
      double prow[kSolveCols], crow[kSolveCols];
#pragma unroll
      for (int k = 0; k < kSolveCols; k++) {
        const int c = min(lane + 64 * k, n);
        prow[k] = Ao[irow * ld + c];
        crow[k] = Ao[icol * ld + c];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kSolveCols; k++) {
        const int c = lane + 64 * k;
        if (c > n || (c < n && is_done(c))) continue;  // pivoted (incl. icol): never read again
        const double pr = prow[k] * pivinv;            // new a[icol][c] (b for c == n)
        for (int j0 = wave; j0 < n; j0 += 8 * kWaves) {
          double av_[8], ac_[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int j = min(j0 + kWaves * u, n - 1);
            const int sj = (j == irow) ? icol : j;  // row j after the swap (row icol takes pr below)
            av_[u] = (j == irow) ? crow[k] : Ao[j * ld + c];
            ac_[u] = Ao[sj * ld + icol];            // column icol: never written
          }
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int j = j0 + kWaves * u;
            if (j < n) {
              const double v = (j == icol) ? pr : av_[u] - pr * ac_[u];
              An[j * ld + c] = v;
              if (c < n && !is_done(j)) {
                const double a = fabs(v);
                if (a >= nbv) {
                  nbv = a;
                  nbe = j * n + c;
                }
              }
            }
          }
        }
      }
    } else
    for (int c = lane; c <= n; c += 64) {
      if (c < n && is_done(c)) continue;  // pivoted (incl. icol): never read again
      const double pr = Ao[irow * ld + c] * pivinv;  // new a[icol][c] (b for c == n)
      // rows j = wave + 4u, eight at a time with their LDS loads issued together
      for (int j0 = wave; j0 < n; j0 += 8 * kWaves) {
        double av_[8], ac_[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int j = min(j0 + kWaves * u, n - 1);
          const int sj = (j == irow) ? icol : ((j == icol) ? irow : j);  // row j after the swap
          av_[u] = Ao[sj * ld + c];
          ac_[u] = Ao[sj * ld + icol];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int j = j0 + kWaves * u;
          if (j < n) {
            const double v = (j == icol) ? pr : av_[u] - pr * ac_[u];
            An[j * ld + c] = v;
            if (c < n && !is_done(j)) {
              const double a = fabs(v);
              if (a >= nbv) {
                nbv = a;
                nbe = j * n + c;
              }
            }
          }
        }
      }
    }
    for (int o = 32; o > 0; o >>= 1) pivot_max(nbv, nbe, __shfl_xor(nbv, o, 64), __shfl_xor(nbe, o, 64));
    if (lane == 0) {
      double *wr = red + ((i + 1) & 1) * 2 * kWaves;
      wr[2 * wave] = nbv;
      wr[2 * wave + 1] = (double)nbe;
    }
    __syncthreads();
  }
  const double *Af = INPLACE ? buf : buf + (n & 1) * nn;
  for (int j = tid; j < n; j += kThreads) x[j] = Af[j * ld + n];
  return 0;
}

// Broadcast lane `lane` (wave-uniform) of a double through scalar registers.
__device__ __forceinline__ double lane_value(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// One DPP exchange of a pivot candidate: the partner lane's (value, index).
template <int CTRL>
__device__ __forceinline__ void dpp_pivot(double &bv, int &be) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(bv), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(bv), CTRL, 0xF, 0xF, false);
  const int oe = __builtin_amdgcn_update_dpp(0, be, CTRL, 0xF, 0xF, false);
  pivot_max(bv, be, __hiloint2double(hi, lo), oe);
}

// Pivot candidate (value, row-major index) reduced over the wave: larger
// value; equal values: larger index (the reference's ">=" scan keeps the
// last).  Inside each row of 16 lanes four DPP exchanges pair every lane
// with one partner (xor 1, xor 2, 7 - i, 15 - i), both lanes of a pair
// combine the same two candidates, so the row ends uniform; the four rows
// then meet through scalar registers.  The result is wave-uniform.
__device__ __forceinline__ void wave_pivot_max(double &bv, int &be) {
  dpp_pivot<0xB1>(bv, be);   // quad_perm [1,0,3,2]
  dpp_pivot<0x4E>(bv, be);   // quad_perm [2,3,0,1]
  dpp_pivot<0x141>(bv, be);  // row_half_mirror
  dpp_pivot<0x140>(bv, be);  // row_mirror
  double v = lane_value(bv, 0);
  int e = __builtin_amdgcn_readlane(be, 0);
#pragma unroll
  for (int r = 16; r < 64; r += 16) pivot_max(v, e, lane_value(bv, r), __builtin_amdgcn_readlane(be, r));
  bv = v;
  be = e;
}

// gauss_errc (period_search/gauss_errc.c:9-73) by one wave with the
// augmented matrix in registers: lane c holds column c of [A|b] (c <= n <=
// R - 1, R <= 64), a[p] its entry in physical row p.  Rows never move: the
// reference's swap of rows irow and icol becomes a swap in a logical-row
// permutation (perm, lane L: physical row of logical row L; inv, lane p:
// logical row of physical row p).  Pivot candidates and ties are judged by
// the logical row-major index j*n + c, as the reference scans, and every
// value b depends on is computed with the reference's operations:
//   pr = a[icol][c] * pivinv; a[j][c] = a[j][c] - pr * a[j][icol].
// Pivoted columns (lanes) keep being updated but are never read again.
// No LDS and no barriers; returns 0, 1 (singular) or 2 (zero pivot), the
// solution in x[0..n) (written by lane 0) on 0.
template <int R>
__device__ __forceinline__ int gauss_solve_wave(double (&a)[R], int n, double *x) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  int perm = lane, inv = lane;
  unsigned long long done = 0ull;
  const bool colc = lane < n;
  double bv = -1.0;
  int be = -1;
#pragma unroll
  for (int p = 0; p < R; p++) {  // physical = logical order here: indices increase with p
    const double v = fabs(a[p]);
    const bool take = colc & (p < n) & (v >= bv);
    bv = take ? v : bv;
    be = take ? p * n + lane : be;
  }
  wave_pivot_max(bv, be);
  int prev_irow = -1, prev_icol = -1;
  for (int i = 0; i < n; i++) {
    int irow, icol;
    if (be < 0) {  // no candidate (NaN matrix): the reference keeps the previous pivot
      if (prev_icol < 0) return 1;
      irow = prev_irow;
      icol = prev_icol;
    } else {
      irow = be / n;
      icol = be - irow * n;
    }
    if ((done >> icol) & 1ull) return 1;  // column pivoted twice: the next scan reports singular
    prev_irow = irow;
    prev_icol = icol;
    const int phr = __builtin_amdgcn_readlane(perm, irow), phc = __builtin_amdgcn_readlane(perm, icol);
    if (irow != icol) {
      perm = (lane == irow) ? phc : ((lane == icol) ? phr : perm);
      inv = (lane == phr) ? icol : ((lane == phc) ? irow : inv);
    }
    const int pp = phr;  // physical row of the pivot row (logical row icol after the swap)
    double prow = 0.0;
#pragma unroll
    for (int p = 0; p < R; p++) prow = (p == pp) ? a[p] : prow;
    const double piv = lane_value(prow, icol);
    if (piv == 0.0) return 2;
    const double pivinv = 1.0 / piv;
    done |= 1ull << icol;
    const double pr = prow * pivinv;  // new a[icol][c] (b for c == n)
    const bool candc = colc && !((done >> lane) & 1ull);
    bv = -1.0;
    be = -1;
    // rows p >= n (register padding) are updated too, harmlessly, but are
    // never candidates; no branches in the row loop
#pragma unroll
    for (int p = 0; p < R; p++) {
      const double dum = lane_value(a[p], icol);
      const double v = (p == pp) ? pr : a[p] - pr * dum;
      a[p] = v;
      const int j = __builtin_amdgcn_readlane(inv, p);
      const bool cand = candc & (p < n) & !((done >> j) & 1ull);
      const double av = fabs(v);
      const int idx = j * n + lane;
      const bool take = cand & ((av > bv) | ((av == bv) & (idx > be)));
      bv = take ? av : bv;
      be = take ? idx : be;
    }
    wave_pivot_max(bv, be);
  }
#pragma unroll
  for (int p = 0; p < R; p++) {
    if (p < n) {
      const double bp = lane_value(a[p], n);
      const int j = __builtin_amdgcn_readlane(inv, p);
      if (lane == 0) x[j] = bp;
    }
  }
  return 0;
}

// gauss_solve_wave with the pivots on the diagonal (mixed forms; as
// gauss_solve_block_diag): the pivot is the not-yet-pivoted row j of largest
// |a_jj|, ties to the larger j, and no row moves.  Lane j holds a_jj in
// register row j, so after each elimination the candidates are one value
// per lane and one wave reduction finds the next pivot.
template <int R>
__device__ __forceinline__ int gauss_solve_wave_diag(double (&a)[R], int n, double *x) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  unsigned long long done = 0ull;
  auto pick = [&]() __attribute__((always_inline)) {
    double d = 0.0;
#pragma unroll
    for (int p = 0; p < R; p++) d = (p == lane) ? a[p] : d;
    double bv = -1.0;
    int be = -1;
    if (lane < n && !((done >> lane) & 1ull)) {
      bv = fabs(d);
      be = lane;
    }
    wave_pivot_max(bv, be);
    return be;
  };
  int icol = pick();
  for (int i = 0; i < n; i++) {
    if (icol < 0) return 1;
    double prow = 0.0;
#pragma unroll
    for (int p = 0; p < R; p++) prow = (p == icol) ? a[p] : prow;
    const double piv = lane_value(prow, icol);
    if (piv == 0.0) return 2;
    const double pivinv = 1.0 / piv;
    done |= 1ull << icol;
    const double pr = prow * pivinv;  // new a[icol][c] (b for c == n)
#pragma unroll
    for (int p = 0; p < R; p++) {
      const double dum = lane_value(a[p], icol);
      a[p] = (p == icol) ? pr : a[p] - pr * dum;
    }
    icol = pick();
  }
#pragma unroll
  for (int p = 0; p < R; p++) {
    if (p < n) {
      const double bp = lane_value(a[p], n);
      if (lane == 0) x[p] = bp;
    }
  }
  return 0;
}

// gauss_errc (period_search/gauss_errc.c:9-73) for 64 <= n <= 127 by four
// waves with the augmented matrix [A|b] in registers: lane l holds columns l
// (a0) and l + 64 (a1); wave w holds the physical rows p = w + 4 r, r <
// RB.  As in gauss_solve_wave rows never move (a logical-row permutation,
// replicated in every wave: perm/inv in two registers per lane, logical or
// physical index l and l + 64), candidates and ties are judged by the
// logical row-major index j*n + c, and every value b depends on is computed
// with the reference's operations:
//   pr = a[icol][c] * pivinv; a[j][c] = a[j][c] - pr * a[j][icol].
// Per pivot step the wave holding the pivot row writes it to LDS (prow,
// n + 1 doubles), a barrier, every wave eliminates its rows and reduces its
// next-pivot candidate (red, two doubles per wave), a barrier.  Returns 0,
// 1 (singular) or 2 (zero pivot), uniform over the workgroup; on 0 the
// solution is in x[0..n) (written by the waves owning the rows).
constexpr int kBlockSolveMax = 127;
template <int RB>
__device__ __forceinline__ int gauss_solve_block(double (&a0)[RB], double (&a1)[RB], int n, double *x,
                                                 double *prow, double *red) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = lane, c1 = lane + 64;
  int perm0 = c0, perm1 = c1, inv0 = c0, inv1 = c1;  // logical -> physical and physical -> logical
  unsigned long long done0 = 0ull, done1 = 0ull;        // pivoted columns (= pivoted logical rows)
  auto is_done = [&](int k) { return ((k < 64 ? done0 : done1) >> (k & 63)) & 1ull; };
  auto perm_of = [&](int L) {  // wave-uniform L
    return L < 64 ? __builtin_amdgcn_readlane(perm0, L) : __builtin_amdgcn_readlane(perm1, L - 64);
  };
  auto inv_of = [&](int p) { return p < 64 ? __builtin_amdgcn_readlane(inv0, p) : __builtin_amdgcn_readlane(inv1, p - 64); };
  auto meet = [&](double bv, int be, double &ov, int &oe) {
    // the waves' candidates meet in LDS; every thread ends with the same pair
    wave_pivot_max(bv, be);
    if (lane == 0) {
      red[2 * wave] = bv;
      red[2 * wave + 1] = (double)be;
    }
    __syncthreads();
    ov = red[0];
    oe = (int)red[1];
#pragma unroll
    for (int w = 1; w < kWaves; w++) pivot_max(ov, oe, red[2 * w], (int)red[2 * w + 1]);
  };
  double bv = -1.0;
  int be = -1;
#pragma unroll
  for (int r = 0; r < RB; r++) {  // physical = logical order here
    const int p = wave + 4 * r;
    const double v0 = fabs(a0[r]), v1 = fabs(a1[r]);
    const bool t0 = (p < n) & (c0 < n) & (v0 >= bv);
    bv = t0 ? v0 : bv;
    be = t0 ? p * n + c0 : be;
    const bool t1 = (p < n) & (c1 < n) & (v1 > bv || (v1 == bv && p * n + c1 > be));
    bv = t1 ? v1 : bv;
    be = t1 ? p * n + c1 : be;
  }
  meet(bv, be, bv, be);
  int prev_irow = -1, prev_icol = -1;
  for (int i = 0; i < n; i++) {
    int irow, icol;
    if (be < 0) {  // no candidate (NaN matrix): the reference keeps the previous pivot
      if (prev_icol < 0) return 1;
      irow = prev_irow;
      icol = prev_icol;
    } else {
      irow = be / n;
      icol = be - irow * n;
    }
    if (is_done(icol)) return 1;  // column pivoted twice: the next scan reports singular
    prev_irow = irow;
    prev_icol = icol;
    const int phr = perm_of(irow), phc = perm_of(icol);
    if (irow != icol) {
      perm0 = (c0 == irow) ? phc : ((c0 == icol) ? phr : perm0);
      perm1 = (c1 == irow) ? phc : ((c1 == icol) ? phr : perm1);
      inv0 = (c0 == phr) ? icol : ((c0 == phc) ? irow : inv0);
      inv1 = (c1 == phr) ? icol : ((c1 == phc) ? irow : inv1);
    }
    const int pp = phr;  // physical row of the pivot row (logical row icol after the swap)
    if (wave == (pp & 3)) {
      const int rp = pp >> 2;
      double v0 = 0.0, v1 = 0.0;
#pragma unroll
      for (int r = 0; r < RB; r++) {
        v0 = (r == rp) ? a0[r] : v0;
        v1 = (r == rp) ? a1[r] : v1;
      }
      if (c0 <= n) prow[c0] = v0;
      if (c1 <= n) prow[c1] = v1;
    }
    __syncthreads();
    const double piv = prow[icol];
    if (piv == 0.0) return 2;
    const double pivinv = 1.0 / piv;
    if (icol < 64)
      done0 |= 1ull << icol;
    else
      done1 |= 1ull << (icol - 64);
    const double pr0 = prow[min(c0, n)] * pivinv, pr1 = prow[min(c1, n)] * pivinv;  // new a[icol][c]
    const bool cand0 = (c0 < n) && !is_done(c0), cand1 = (c1 < n) && !is_done(c1);
    const int hi = icol >= 64, il = icol & 63;
    bv = -1.0;
    be = -1;
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int p = wave + 4 * r;
      const double dum = lane_value(hi ? a1[r] : a0[r], il);
      const double v0 = (p == pp) ? pr0 : a0[r] - pr0 * dum;
      const double v1 = (p == pp) ? pr1 : a1[r] - pr1 * dum;
      a0[r] = v0;
      a1[r] = v1;
      const int j = inv_of(min(p, 127));
      const bool rowc = (p < n) && !is_done(j);
      const double f0 = fabs(v0), f1 = fabs(v1);
      const int i0 = j * n + c0, i1 = j * n + c1;
      const bool t0 = rowc & cand0 & ((f0 > bv) | ((f0 == bv) & (i0 > be)));
      bv = t0 ? f0 : bv;
      be = t0 ? i0 : be;
      const bool t1 = rowc & cand1 & ((f1 > bv) | ((f1 == bv) & (i1 > be)));
      bv = t1 ? f1 : bv;
      be = t1 ? i1 : be;
    }
    meet(bv, be, bv, be);
  }
  const int bh = n >= 64, bl = n & 63;  // column n (b)
#pragma unroll
  for (int r = 0; r < RB; r++) {
    const int p = wave + 4 * r;
    if (p < n) {
      const double bp = lane_value(bh ? a1[r] : a0[r], bl);
      const int j = inv_of(p);
      if (lane == 0) x[j] = bp;
    }
  }
  return 0;
}

// Gauss-Jordan with the pivots on the diagonal (the mixed forms' solve, in
// FP64): each pivot step does gauss_solve_block's arithmetic, with the pivot
// the not-yet-pivoted row j of largest |a_jj|, ties to the larger j (the
// reference's ">=" row-major scan restricted to the diagonal).  The damped
// normal matrices of a fit are symmetric and positive definite in exact
// arithmetic, so |a_jk| <= sqrt(a_jj a_kk): complete pivoting
// (gauss_errc.c:21-36) finds its maximum on the diagonal as well, up to
// rounding-level ties.  Diagonal pivots need no row swaps, and each step
// reduces n diagonal entries instead of the whole remaining matrix.
// Layout as gauss_solve_block: wave w holds physical = logical rows
// w + 4r, lane l columns l and l + 64.  LDS: prow [kMaxMfit + 1], dg
// [kMaxMfit + 1] (the current diagonal).  Returns 2 on a zero pivot, 1 when
// no candidate is left (NaN diagonal), as gauss_errc reports singular
// matrices.
template <int RB>
__device__ __forceinline__ int gauss_solve_block_diag(double (&a0)[RB], double (&a1)[RB], int n, double *x,
                                                      double *prow, double *dg) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = lane, c1 = lane + 64;
  unsigned long long done0 = 0ull, done1 = 0ull;  // pivoted rows (= columns)
  // the diagonal entries of this wave's rows, in the lanes of their columns
  auto put_diag = [&]() __attribute__((always_inline)) {
    double d0 = 0.0, d1 = 0.0;
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int p = wave + 4 * r;
      if (r < 16)
        d0 = (lane == p) ? a0[r] : d0;
      else
        d1 = (lane == p - 64) ? a1[r] : d1;
    }
    if ((lane & 3) == wave) {
      if (c0 < n) dg[c0] = d0;
      if (c1 < n) dg[c1] = d1;
    }
  };
  // the next pivot row: argmax of |a_jj| over the rows not pivoted yet
  auto pick = [&]() __attribute__((always_inline)) {
    double bv = -1.0;
    int be = -1;
    if (c0 < n && !((done0 >> c0) & 1ull)) {
      bv = fabs(dg[c0]);
      be = c0;
    }
    if (c1 < n && !((done1 >> lane) & 1ull)) pivot_max(bv, be, fabs(dg[c1]), c1);
    wave_pivot_max(bv, be);
    return be;
  };
  put_diag();
  __syncthreads();
  int icol = pick();
  for (int i = 0; i < n; i++) {
    if (icol < 0) return 1;
    if (wave == (icol & 3)) {
      const int rp = icol >> 2;
      double v0 = 0.0, v1 = 0.0;
#pragma unroll
      for (int r = 0; r < RB; r++) {
        v0 = (r == rp) ? a0[r] : v0;
        v1 = (r == rp) ? a1[r] : v1;
      }
      if (c0 <= n) prow[c0] = v0;
      if (c1 <= n) prow[c1] = v1;
    }
    __syncthreads();
    const double piv = prow[icol];
    if (piv == 0.0) return 2;
    const double pivinv = 1.0 / piv;
    if (icol < 64)
      done0 |= 1ull << icol;
    else
      done1 |= 1ull << (icol - 64);
    const double pr0 = prow[min(c0, n)] * pivinv, pr1 = prow[min(c1, n)] * pivinv;  // new a[icol][c]
    const int hi = icol >= 64, il = icol & 63;
#pragma unroll
    for (int r = 0; r < RB; r++) {
      const int p = wave + 4 * r;
      const double dum = lane_value(hi ? a1[r] : a0[r], il);
      a0[r] = (p == icol) ? pr0 : a0[r] - pr0 * dum;
      a1[r] = (p == icol) ? pr1 : a1[r] - pr1 * dum;
    }
    put_diag();
    __syncthreads();
    icol = pick();
  }
  const int bh = n >= 64, bl = n & 63;  // column n (b)
#pragma unroll
  for (int r = 0; r < RB; r++) {
    const int p = wave + 4 * r;
    if (p < n) {
      const double bp = lane_value(bh ? a1[r] : a0[r], bl);
      if (lane == 0) x[p] = bp;
    }
  }
  return 0;
}

__device__ void slot_curv(const DevProblem &P, const SlotMem &M, double *mem, int tid, int nthreads) {
#pragma clang fp contract(off)
  const double *atry = mem + M.atry;
  for (int f = tid; f < P.nfac; f += nthreads) {
    const double gf = curv_g(P, atry, f);
    mem[M.g + f] = gf;
    mem[M.area + f] = P.darea[f] * gf;
  }
}

// Problem of a slot.  One problem (MULTI false): the kernel argument P0, as
// constant as any other argument.  Several (MULTI true): the slot's entry of
// the device array of problems.
template <bool MULTI>
__device__ __forceinline__ const DevProblem &pick_problem(const DevProblem &P0, const DevProblem *probs, int idx) {
  if constexpr (MULTI) {
    // The problem array is read-only for the whole kernel: reading it through
    // the constant address space lets its (wave-uniform) fields come in by
    // scalar loads even after the kernel's own global stores.
    typedef const __attribute__((address_space(4))) DevProblem ConstProblem;
    ConstProblem *cp = (ConstProblem *)probs + __builtin_amdgcn_readfirstlane(idx);
    return *(const DevProblem *)cp;
  } else {
    return P0;
  }
}

// step_kernel: one workgroup per slot.  Finishes a fit whose loop predicate
// is false (per-fit summary) and takes the next fit from the queue, or damps
// the accepted normal equations and solves them, forming the trial
// parameters and their shape (curv()).
// LDS: the solver's two augmented matrices and pivot exchange; the final
// curv of a finished fit (area/g [F]) reuses the same space.
// WAVE (mfit <= 63): one wave per slot, the solve in registers
// (gauss_solve_wave); BLOCK (64 <= mfit <= 127): 256 threads, the solve in
// registers of four waves (gauss_solve_block); otherwise 256 threads and the
// LDS solver.
constexpr int kWaveSolveMax = 63;
// Point order of a fit: within each light curve, the points by
// the body-frame longitude of the Earth-Sun bisector at the fit's starting
// pole and period, in kOrderBins bins, ties in point order (a stable
// counting sort, so the order is deterministic).  Points that see nearly
// the same facets then share a model tile, and the model kernels' wave-wide
// skip of facets that no point of the wave sees takes effect more often.
// Only the order in which the normal equations sum the points changes.
// Also writes the points' brightness, sigma and weight/sigma^2 in that
// order for the normal equations.  lds: 16 doubles + kOrderBins + 1 ints.
constexpr int kOrderBins = 64;
__device__ void point_order(const DevProblem &P, const SlotMem &M, double *mem, int tid, int nthreads, double *lds) {
  const double *a = mem + M.a;
  int *perm = reinterpret_cast<int *>(mem + M.perm);
  int *bin = reinterpret_cast<int *>(mem + M.psw);  // bins first; the weights overwrite them at the end
  double *blm = lds;
  int *cnt = reinterpret_cast<int *>(lds + 16);
  const int nc = P.ncoef, n = P.n_points;
  if (tid == 0) eval_blmatrix(a[nc], a[nc + 1], blm);
  __syncthreads();
  const double om = a[nc + 2];
  for (int p = tid; p < n; p += nthreads) {
    const double bx = P.ex[p] + P.sx[p], by = P.ey[p] + P.sy[p], bz = P.ez[p] + P.sz[p];
    const double ux = blm[0] * bx + blm[1] * by + blm[2] * bz, uy = blm[3] * bx + blm[4] * by + blm[5] * bz;
    double lon = atan2(uy, ux) - fmod(om * P.tim[p] + P.phi0, 2 * kPi);  // body-frame longitude
    lon -= 2 * kPi * floor(lon / (2 * kPi));
    bin[p] = min(max((int)(lon * (kOrderBins / (2 * kPi))), 0), kOrderBins - 1);
  }
  __syncthreads();
  for (int lc = 0; lc < P.n_lc; lc++) {
    const int q0 = P.lc_start[lc], q1 = P.lc_start[lc + 1];
    if (tid < kOrderBins) {
      int c = 0;
      for (int q = q0; q < q1; q++) c += bin[q] == tid ? 1 : 0;
      cnt[tid + 1] = c;
    }
    __syncthreads();
    if (tid == 0) {
      cnt[0] = 0;
      for (int b = 0; b < kOrderBins; b++) cnt[b + 1] += cnt[b];
    }
    __syncthreads();
    if (tid < kOrderBins) {
      int k = q0 + cnt[tid];
      for (int q = q0; q < q1; q++)
        if (bin[q] == tid) perm[k++] = q;
    }
    __syncthreads();
  }
  double *pys = mem + M.pys, *psg = mem + M.psg, *psw = mem + M.psw;
  for (int q = tid; q < n; q += nthreads) {
    const int p = perm[q];
    pys[q] = P.y[p];
    psg[q] = P.sig[p];
  }
  __syncthreads();
  for (int q = tid; q < n; q += nthreads) psw[q] = P.swt[perm[q]];
  __syncthreads();
}

template <bool MULTI, int RW, int BR, bool DIAG = false>  // RW: register rows of the one-wave solver (>= mfit); BR:
                                                         // rows per wave of the four-wave solver (>= mfit / 4); both
                                                         // 0: LDS solver; DIAG: the four-wave solver with diagonal
                                                         // pivots (mixed forms)
__device__ __forceinline__ void step_body(const DevProblem &P0, const DevProblem *__restrict__ probs, const DevWork &W) {
  constexpr bool WAVE = RW > 0;
  constexpr bool BLOCK = BR > 0;
  constexpr int kStepThreads = WAVE ? 64 : kThreads;
  extern __shared__ double sm[];
  const int tid = threadIdx.x, slot = blockIdx.x;
  SlotState &S = W.st[slot];
  double *mem = W.mem + (size_t)slot * W.slot_doubles;
  __shared__ int s_phase, s_go, s_item, s_prob;

  if (tid == 0) {
    int ph = S.phase;
    int go = 0;
    if (ph == PH_ITER) {
      // loop predicate (period_search_BOINC.cpp:868)
      const DevProblem &Pc = pick_problem<MULTI>(P0, probs, S.prob);
      go = (((S.niter < Pc.n_iter_max) && (S.iter_diff > Pc.iter_diff_max)) || (S.niter < Pc.n_iter_min)) ? 1 : 0;
    }
    s_phase = ph;
    s_go = go;
    s_item = S.item;
    s_prob = S.prob;
  }
  __syncthreads();
  int ph = s_phase;
  if (ph == PH_DRAINED || ph == PH_INIT) {
    if (tid == 0) S.eval = (ph == PH_INIT) ? 1 : 0;
    return;
  }
  if (ph == PH_ITER && !s_go) {
    // per-fit summary (period_search_BOINC.cpp:900-911); totarea from the
    // shape of the last curv() call.
    const DevProblem &P = pick_problem<MULTI>(P0, probs, s_prob);
    const int ma = P.ma, nc = P.ncoef;
    const SlotMem M = slot_mem(ma, P.mfit, P.npad);
    double *area = sm, *g = area + P.nfac;
    eval_curv(P, mem + M.alast, area, g);
    if (W.cg_out)
      for (int l = tid; l < ma; l += kStepThreads) W.cg_out[(size_t)s_item * ma + l] = mem[M.a + l];
    __syncthreads();
    if (tid == 0) {
#pragma clang fp contract(off)
      double tot = 0;
      for (int f = 0; f < P.nfac; f++) tot = tot + area[f];
      const double *a = mem + M.a;
      const double c0 = S.chck[0], c1 = S.chck[1], c2 = S.chck[2];
      ps_fit_out o;
      o.dev = S.dev_new;
      o.chisq = S.chisq;
      o.rchisq = S.rchisq;
      o.prd = 2 * kPi / a[nc + 2];
      o.la = (180 / kPi) * a[nc + 1];
      o.be = 90 - (180 / kPi) * a[nc];
      o.dark = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
      o.totarea = tot;
      o.alamda = S.lam;
      o.niter = S.niter;
      o.err = S.err;
      o.latency_us = (double)(wall_clock64() - S.t_start) * W.us_per_tick;
      W.out[s_item] = o;
      atomicAdd(W.done, 1);
    }
    ph = PH_EMPTY;
  }
  if (ph == PH_EMPTY) {
    if (tid == 0) {
      int item = atomicAdd(W.queue, 1);
      s_item = item;
      s_phase = item < W.n_items ? PH_INIT : PH_DRAINED;
      s_prob = (item < W.n_items && W.item_prob) ? W.item_prob[item] : 0;
    }
    __syncthreads();
    const int item = s_item;
    if (s_phase == PH_DRAINED) {
      if (tid == 0) {
        S.phase = PH_DRAINED;
        S.eval = 0;
      }
      return;
    }
    const DevProblem &P = pick_problem<MULTI>(P0, probs, s_prob);
    const int ma = P.ma;
    const SlotMem M = slot_mem(ma, P.mfit, P.npad);
    double *a = mem + M.a, *atry = mem + M.atry;
    if (W.mode != 0) {  // given starting parameters (evaluation or fit)
      for (int l = tid; l < ma; l += kStepThreads) a[l] = W.cg_in[(size_t)item * ma + l];
    } else {
      init_params(P, W.freq[item], W.pole[item], a);
    }
    for (int l = tid; l < ma; l += kStepThreads) atry[l] = a[l];
    if (tid == 0) {
      S.item = item;
      S.prob = s_prob;
      S.t_start = wall_clock64();
      S.phase = PH_INIT;
      S.eval = 1;
      S.err = 0;
    }
    __syncthreads();
    if (W.order) point_order(P, M, mem, tid, kStepThreads, sm);
    slot_curv(P, M, mem, tid, kStepThreads);
    return;
  }
  // damped normal equations and solve (mrqmin.c:53-79)
  const DevProblem &P = pick_problem<MULTI>(P0, probs, s_prob);
  const int mfit = P.mfit, ma = P.ma;
  const SlotMem M = slot_mem(ma, mfit, P.npad);
  double *A = sm;
  const double *aacc = mem + (S.sel ? M.alpha1 : M.alpha0);
  const double *bacc = mem + (S.sel ? M.beta1 : M.beta0);
  double *da = mem + M.da;
  int err;
  if constexpr (WAVE) {
#pragma clang fp contract(off)
    // lane c loads column c of the damped [alpha | beta] (c <= mfit)
    const double lam = S.lam;
    const int lane = tid;
    // lanes <= mfit read alpha column min(lane, mfit-1); lane mfit takes beta
    const double *colp = (lane == mfit) ? bacc : aacc + min(lane, mfit - 1);
    const int rstride = (lane == mfit) ? 1 : mfit;
    double am[WAVE ? RW : 1];
#pragma unroll
    for (int p = 0; p < RW; p++) {
      const int j = min(p, mfit - 1);
      const double av = colp[j * rstride];
      am[p] = av * ((p == lane) ? (1 + lam) : 1.0);  // times 1.0 is exact; beta's lane never matches
    }
    if constexpr (DIAG)
      err = gauss_solve_wave_diag(am, mfit, da);
    else
      err = gauss_solve_wave(am, mfit, da);
    __syncthreads();
  } else if constexpr (BLOCK) {
#pragma clang fp contract(off)
    // lane l loads columns l and l + 64 of the damped [alpha | beta] for the
    // wave's rows p = wave + 4 r (column mfit is beta)
    const double lam = S.lam;
    const int lane = tid & 63, wave = tid >> 6;
    double a0[BLOCK ? BR : 1], a1[BLOCK ? BR : 1];
#pragma unroll
    for (int r = 0; r < BR; r++) {
      const int p = wave + 4 * r, j = min(p, mfit - 1);
      const int ca = lane, cb = lane + 64;
      const double va = (ca == mfit) ? bacc[j] : aacc[j * mfit + min(ca, mfit - 1)];
      const double vb = (cb == mfit) ? bacc[j] : aacc[j * mfit + min(cb, mfit - 1)];
      a0[r] = va * ((p == ca) ? (1 + lam) : 1.0);  // times 1.0 is exact; beta's column never matches
      a1[r] = vb * ((p == cb) ? (1 + lam) : 1.0);
    }
    if constexpr (DIAG)
      err = gauss_solve_block_diag<BLOCK ? BR : 1>(a0, a1, mfit, da, sm, sm + 128);
    else
      err = gauss_solve_block<BLOCK ? BR : 1>(a0, a1, mfit, da, sm, sm + 128);
    __syncthreads();
  } else {
  {
#pragma clang fp contract(off)
    const double lam = S.lam;
    const int ld = mfit + 1, wave = tid >> 6, lane = tid & 63;
    // four rows at a time, loads without branches (clamped indices) so that
    // they are in flight together
    for (int j0 = wave; j0 < mfit; j0 += 4 * kWaves) {
      for (int k = lane; k <= mfit; k += 64) {
        const int kc = min(k, mfit - 1);
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int j = min(j0 + kWaves * q, mfit - 1);
          const double av = aacc[j * mfit + kc], bv = bacc[j];
          const double damp = av * ((j == k) ? (1 + lam) : 1.0);  // times 1.0 is exact
          v[q] = (k == mfit) ? bv : damp;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int j = j0 + kWaves * q;
          if (j < mfit) A[j * ld + k] = v[q];
        }
      }
    }
  }
  __syncthreads();
  err = solve_in_place(mfit) ? gauss_solve<true>(A, mfit, da) : gauss_solve<false>(A, mfit, da);
  __syncthreads();
  }
  double *a = mem + M.a, *atry = mem + M.atry;
  if (err == 0) {
#pragma clang fp contract(off)
    for (int l = tid; l < ma; l += kStepThreads) atry[l] = a[l];
    __syncthreads();
    for (int jf = tid; jf < mfit; jf += kStepThreads) {
      const int l = P.fidx[jf];
      atry[l] = a[l] + da[jf];
    }
    __syncthreads();
    slot_curv(P, M, mem, tid, kStepThreads);  // the shape at the trial parameters
  }
  if (tid == 0) {
    S.eval = err == 0 ? 1 : 0;
    S.err = err;
  }
}

template <bool MULTI, int RW, int BR = 0, bool DIAG = false>
__global__ void __launch_bounds__(RW ? 64 : kThreads, RW ? 1 : (BR ? 2 : 4)) step_kernel(DevProblem P0,
                                                                        const DevProblem *__restrict__ probs,
                                                                        DevWork W) {
  step_body<MULTI, RW, BR, DIAG>(P0, probs, W);
}

// model_kernel: one workgroup per (slot, tile); blockIdx.x = tile * nslots + slot.
#ifndef PS_MODEL_MIN_WAVES
#define PS_MODEL_MIN_WAVES 1
#endif
template <int NCB, bool MULTI>
__global__ void __launch_bounds__(kThreads, PS_MODEL_MIN_WAVES) model_kernel(DevProblem P0,
                                                                            const DevProblem *__restrict__ probs,
                                                                            DevWork W) {
  extern __shared__ double sm[];
  const int slot = blockIdx.x % W.nslots, tile = blockIdx.x / W.nslots;
  if (!W.st[slot].eval) return;
  const DevProblem &P = pick_problem<MULTI>(P0, probs, W.st[slot].prob);
  if (tile >= P.ntiles) return;
  const SlotMem M = slot_mem(P.ma, P.mfit, P.npad);
  double *mem = W.mem + (size_t)slot * W.slot_doubles;
  double *x = sm, *blm = x + kMaxMa, *area = blm + 32, *g = area + P.nfac;
  PointGeom *gst = reinterpret_cast<PointGeom *>(g + P.nfac);
  for (int l = threadIdx.x; l < P.ma; l += kThreads) x[l] = mem[M.atry + l];
  for (int f = threadIdx.x; f < P.nfac; f += kThreads) area[f] = mem[M.area + f];
  if (threadIdx.x == 0) eval_blmatrix(x[P.ncoef], x[P.ncoef + 1], blm);
  __syncthreads();
  (void)g;
  model_tile<NCB>(P, tile * kThreads, x, blm, area, mem + M.J, gst);
}

// model_mfma_kernel: one workgroup per (slot, tile of kTilePts points);
// blockIdx.x = tile * nslots + slot.  The same model as model_tile
// (period_search/bright.c:14-126) with the shape-derivative contraction on
// the matrix cores:
//   facet stage (VALU, FP64): thread (point pl, sub-slot s) evaluates the
//     facets f = f0 + s + 4i of each chunk: visibility, scattering law,
//     the weight W[p][f] = Area_f s_pf (0 if not visible) into an LDS tile,
//     and its partial sums of br, V, U, cl- and cls-derivatives;
//   contraction (MFMA, FP64): wave w owns points 16w..16w+15 and computes
//     S[p][c] = sum_f W[p][f] Dsph[f][c] with v_mfma_f64_16x16x4 over the
//     chunk's Dsph rows staged in LDS (A = W, B = Dsph, NB blocks of 16
//     columns).  D layout: lane holds D[4r + lane/16][lane%16], r = 0..3.
// LDS: x[kMaxMa] blm[32] area[F] geom[kTilePts] W[kFacetChunk][kTilePts]
//      Dsph chunk[kFacetChunk][16 NB]; the last two regions are reused for
//      the partial-sum reduction at the end.
template <int NB, bool MULTI>
__global__ void __launch_bounds__(kThreads, NB <= 4 ? 3 : 2) model_mfma_kernel(DevProblem P0, const DevProblem *__restrict__ probs,
                                                              DevWork W) {
  extern __shared__ double sm[];
  const int slot = blockIdx.x % W.nslots, tile = blockIdx.x / W.nslots;
  if (!W.st[slot].eval) return;
  const DevProblem &P = pick_problem<MULTI>(P0, probs, W.st[slot].prob);
  if (tile >= P.ntiles64) return;
  constexpr int NC = 16 * NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const SlotMem M = slot_mem(P.ma, P.mfit, P.npad);
  double *mem = W.mem + (size_t)slot * W.slot_doubles;
  const int F = P.nfac;
  double *x = sm, *blm = x + kMaxMa, *area = blm + 32;
  PointGeom *gst = reinterpret_cast<PointGeom *>(area + F);
  double *wt = reinterpret_cast<double *>(gst + kTilePts);
  double *dch = wt + kFacetChunk * kTilePts;
  double *red = wt;  // [4][kTilePts][9] after the last chunk

  for (int l = tid; l < P.ma; l += kThreads) x[l] = mem[M.atry + l];
  for (int f = tid; f < F; f += kThreads) area[f] = mem[M.area + f];
  if (tid == 0) eval_blmatrix(mem[M.atry + P.ncoef], mem[M.atry + P.ncoef + 1], blm);
  __syncthreads();
  const int p0 = tile * kTilePts;
  if (tid < kTilePts) {
    const int p = min(p0 + tid, P.n_points - 1);
    PointGeom G;
    point_geometry(P, p, x, blm, G);
    gst[tid] = G;
  }
  __syncthreads();
  const int pl = tid % kTilePts, sub = tid / kTilePts;
  const double e_0 = gst[pl].e[0], e_1 = gst[pl].e[1], e_2 = gst[pl].e[2];
  const double s_0 = gst[pl].e0[0], s_1 = gst[pl].e0[1], s_2 = gst[pl].e0[2];
  const double cl = exp(x[P.ma - 2]);
  const double cls = x[P.ma - 1];
  double br = 0, v0 = 0, v1 = 0, v2 = 0, u0 = 0, u1 = 0, u2 = 0, scl = 0, scls = 0;
  v4d acc[NB];
#pragma unroll
  for (int cb = 0; cb < NB; cb++) acc[cb] = v4d{0, 0, 0, 0};
  const double4 *__restrict__ n4 = P.nor4;
  // Dsph chunks ([fpad][NC], zero-padded: a chunk is 32 contiguous rows) pass
  // through registers one chunk ahead, so their loads are in flight during
  // the previous chunk's facet stage and contraction
  const v4u *__restrict__ dsq = reinterpret_cast<const v4u *>(P.dsphP);
  constexpr int kPieces = kFacetChunk * NC / 2;  // 16-byte pieces per chunk
  constexpr int kStageIt = (kPieces + kThreads - 1) / kThreads;
  v4u pre[kStageIt];
  // (threads past the last piece repeat it: identical loads and stores, no
  // branches, so the pieces stay in registers)
  auto load_chunk = [&](int fc) __attribute__((always_inline)) {
    static_for<0, kStageIt>([&](auto c) __attribute__((always_inline)) {
      constexpr int it = decltype(c)::value;
      const int idx = min(tid + it * kThreads, kPieces - 1);
      pre[it] = dsq[(size_t)fc * (NC / 2) + idx];
    });
  };
  load_chunk(0);
  for (int f0 = 0; f0 < F; f0 += kFacetChunk) {
    static_for<0, kStageIt>([&](auto c) __attribute__((always_inline)) {
      constexpr int it = decltype(c)::value;
      const int idx = min(tid + it * kThreads, kPieces - 1);
      reinterpret_cast<v4u *>(dch)[idx] = pre[it];
    });
    if (f0 + kFacetChunk < F) load_chunk(f0 + kFacetChunk);
    // facet stage; a wave's facets are wave-uniform: normals by scalar
    // loads and areas fetched four facets at a time, before their use
    const int fb = __builtin_amdgcn_readfirstlane(f0 + sub);
#pragma unroll
    for (int g = 0; g < kFacetChunk / 16; g++) {
      double4 nvs[4];
      double afs[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int f = min(fb + 4 * (4 * g + j), F - 1);
        nvs[j] = n4[f];
        afs[j] = area[f];
      }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int i = 4 * g + j;
      const int fl = sub + 4 * i, f = fb + 4 * i;
      double wv = 0.0;
      if (f < F) {
        const double4 nv = nvs[j];
        const double mu = e_0 * nv.x + e_1 * nv.y + e_2 * nv.z;
        const double mu0 = s_0 * nv.x + s_1 * nv.y + s_2 * nv.z;
        if (mu > kTiny && mu0 > kTiny) {
          const double af = afs[j];
          const double rd = recip(mu + mu0);
          const double mm = mu * mu0;
          wv = af * (mm * (cl + cls * rd));
          br += wv;
          const double q0 = mu0 * rd, q1 = mu * rd;
          const double ad = af * (cls * q0 * q0 + cl * mu0);
          const double ad0 = af * (cls * q1 * q1 + cl * mu);
          v0 += ad * nv.x;
          v1 += ad * nv.y;
          v2 += ad * nv.z;
          u0 += ad0 * nv.x;
          u1 += ad0 * nv.y;
          u2 += ad0 * nv.z;
          const double am = af * mm;
          scl += am;
          scls += am * rd;
        }
      }
      wt[fl * kTilePts + pl] = wv;
    }
    }
    __syncthreads();
    // contraction on the matrix cores: wave = row block of 16 points
#pragma unroll
    for (int kk = 0; kk < kFacetChunk / 4; kk++) {
      const int kr = 4 * kk + (lane >> 4);
      const double av = wt[kr * kTilePts + 16 * wave + (lane & 15)];
#pragma unroll
      for (int cb = 0; cb < NB; cb++) {
        const double bv = dch[kr * NC + 16 * cb + (lane & 15)];
        acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[cb], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // reduce the per-point partial sums of the 4 sub-slots
  {
    double *rp = red + ((size_t)sub * kTilePts + pl) * 9;
    rp[0] = br;
    rp[1] = v0;
    rp[2] = v1;
    rp[3] = v2;
    rp[4] = u0;
    rp[5] = u1;
    rp[6] = u2;
    rp[7] = scl;
    rp[8] = scls;
  }
  __syncthreads();
  const int ld = P.npad;
  double *J = mem + M.J;
  if (tid < kTilePts && p0 + tid < P.n_points) {
    const int p = p0 + tid;
    double t[9];
    for (int k = 0; k < 9; k++) t[k] = 0;
    for (int q = 0; q < 4; q++)
      for (int k = 0; k < 9; k++) t[k] += red[((size_t)q * kTilePts + tid) * 9 + k];
    const PointGeom G = gst[tid];
    const double sc = G.scale;
    const double V[3] = {t[1], t[2], t[3]}, U[3] = {t[4], t[5], t[6]};
    double r[3];
    rotation_derivs(P, p, blm, G, V, U, r);
    for (int k = 0; k < 3; k++) J[(size_t)(P.ncoef + k) * ld + p] = sc * r[k];
    for (int i = 0; i < 3; i++) J[(size_t)(P.ncoef + 3 + i) * ld + p] = t[0] * G.dphp[i];
    J[(size_t)(P.ncoef + 6) * ld + p] = sc * t[7] * cl;
    J[(size_t)(P.ncoef + 7) * ld + p] = sc * t[8];
    J[(size_t)P.ma * ld + p] = t[0] * sc;
  }
  // shape derivatives from the accumulators
#pragma unroll
  for (int cb = 0; cb < NB; cb++) {
    const int c = 16 * cb + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int pp = 16 * wave + 4 * r + (lane >> 4);
      const int p = p0 + pp;
      if (p < P.n_points && c < P.ncoef) J[(size_t)c * ld + p] = gst[pp].scale * acc[cb][r];
    }
  }
}

// Mixed-precision model (PS_PREC_FP32, PS_PREC_BF16, PS_PREC_BF16X3, PS_PREC_BF16X6; PS_PREC_BF16A as PS_PREC_BF16).
// Per-point geometry (phase reduction, rotation matrices, phase function)
// stays FP64 (point_geometry); the facet stage runs in FP32 and the
// shape-derivative contraction S = W Dsph on the matrix cores with FP32
// accumulation, with W and Dsph in FP32 (MODE 0: v_mfma_f32_16x16x4f32,
// 8 k-steps per 32-facet chunk), in bf16 (MODE 1: v_mfma_f32_16x16x32_bf16,
// one k-step per chunk) or split into bf16 high and low parts (MODE 2:
// x = hi + lo, S = Whi Dhi + Whi Dlo + Wlo Dhi, three bf16 products per
// chunk, about 16 significant bits).  Derivative
// rows are stored as FP32; the model value ymod (row ma) stays FP64, so the
// residuals, chi2 and the LM decisions are FP64.
//   facet stage: thread (point pl, sub-slot s) takes facets f0+8s..f0+8s+7
//     of each chunk and writes their weights as one 16/32-byte row piece;
//   contraction: wave w owns points 16w..16w+15; A = W rows, B = Dsph^T
//     rows staged from the pre-converted table.  Operand layout (A lane l:
//     row l%16, k-group l/16; B lane l: column l%16, k-group l/16) and
//     result layout D[4*(l/16)+r][l%16].
// LDS layout in bytes (mixed_layout): x, blm, area (FP32), geometry, W tile
// [kTilePts][WS], Dsph^T chunk [16 NB][WS]; the tiles are reused for the
// partial-sum reduction.
struct MixedLds {
  size_t x, blm, area, geom, wt, dch, red, total;
};
__host__ __device__ inline size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }
__host__ __device__ inline MixedLds mixed_layout(int nfac, int nb, int mode) {
  MixedLds L;
  size_t o = 0;
  const bool bf = mode != 0;
  const size_t es = bf ? 2 : 4, ws = kFacetChunk + (bf ? 8 : 4), parts = mode == 3 ? 3 : (mode == 2 ? 2 : 1);
  L.x = o;
  o += kMaxMa * sizeof(double);
  L.blm = o;
  o += 32 * sizeof(double);
  L.area = o;  // zero-padded to whole facet chunks
  o = align16(o + (size_t)((nfac + kFacetChunk - 1) / kFacetChunk * kFacetChunk) * sizeof(float));
  L.geom = o;
  o = align16(o + kTilePts * sizeof(PointGeom));
  L.wt = o;
  o = align16(o + parts * kTilePts * ws * es);
  L.dch = o;
  o = align16(o + parts * 16 * nb * ws * es);
  L.red = L.wt;
  const size_t red = (size_t)kWaves * kTilePts * 8 * sizeof(float);
  if (o - L.wt < red) o = L.wt + red;
  L.total = o;
  return L;
}


template <int NB, int MODE, bool MULTI>
__global__ void __launch_bounds__(kThreads, ((MODE == 1 || MODE == 2) && NB <= 4) ? 5 : 4) model_mixed_kernel(DevProblem P0, const DevProblem *__restrict__ probs,
                                                               DevWork W) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smb[];
  const int slot = blockIdx.x % W.nslots, tile = blockIdx.x / W.nslots;
  if (!W.st[slot].eval) return;
  const DevProblem &P = pick_problem<MULTI>(P0, probs, W.st[slot].prob);
  if (tile >= P.ntiles64) return;
  constexpr bool BF = MODE != 0, SPLIT = MODE >= 2, SPLIT3 = MODE == 3;
  constexpr int NC = 16 * NB;
  constexpr int WS = kFacetChunk + (BF ? 8 : 4);   // padded row, elements (80 / 144 bytes)
  using T = typename std::conditional<BF, unsigned short, float>::type;
  constexpr int PER16 = 16 / (int)sizeof(T);      // elements per 16-byte piece
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const SlotMem M = slot_mem(P.ma, P.mfit, P.npad);
  double *mem = W.mem + (size_t)slot * W.slot_doubles;
  const int F = P.nfac;
  const MixedLds L = mixed_layout(F, NB, MODE);
  double *x = reinterpret_cast<double *>(smb + L.x), *blm = reinterpret_cast<double *>(smb + L.blm);
  float *areaf = reinterpret_cast<float *>(smb + L.area);
  PointGeom *gst = reinterpret_cast<PointGeom *>(smb + L.geom);
  T *wt = reinterpret_cast<T *>(smb + L.wt);     // [kTilePts][WS] (then the low parts)
  T *dch = reinterpret_cast<T *>(smb + L.dch);   // [NC][WS] (then the low parts)
  T *wtl = wt + kTilePts * WS, *dchl = dch + NC * WS;
  T *wtl2 = wtl + kTilePts * WS, *dchl2 = dchl + NC * WS;  // third parts (SPLIT3)
  float *red = reinterpret_cast<float *>(smb + L.red);

  for (int l = tid; l < P.ma; l += kThreads) x[l] = mem[M.atry + l];
  for (int f = tid; f < P.fpad; f += kThreads) areaf[f] = f < F ? (float)mem[M.area + f] : 0.0f;
  if (tid == 0) eval_blmatrix(mem[M.atry + P.ncoef], mem[M.atry + P.ncoef + 1], blm);
  __syncthreads();
  // tile positions p0..p0+63: points in the fit's order (point_order) or in
  // input order; the Jacobian's columns follow the positions
  const int p0 = tile * kTilePts;
  const int *__restrict__ perm = reinterpret_cast<const int *>(mem + M.perm);
  if (tid < kTilePts) {
    const int q = min(p0 + tid, P.n_points - 1);
    const int p = W.order ? perm[q] : q;
    PointGeom G;
    point_geometry(P, p, x, blm, G);
    gst[tid] = G;
  }
  __syncthreads();
  const int pl = tid % kTilePts, sub = tid / kTilePts;
  const float e_0 = (float)gst[pl].e[0], e_1 = (float)gst[pl].e[1], e_2 = (float)gst[pl].e[2];
  const float s_0 = (float)gst[pl].e0[0], s_1 = (float)gst[pl].e0[1], s_2 = (float)gst[pl].e0[2];
  const double cld = exp(x[P.ma - 2]);
  const float cl = (float)cld, cls = (float)x[P.ma - 1], tiny = (float)kTiny;
  float v0 = 0, v1 = 0, v2 = 0, u0 = 0, u1 = 0, u2 = 0, scl = 0, scls = 0;
  v4f acc[NB];
#pragma unroll
  for (int cb = 0; cb < NB; cb++) acc[cb] = v4f{0, 0, 0, 0};
  const float *__restrict__ n3 = P.nor3f;
  const T *__restrict__ dsT = BF ? reinterpret_cast<const T *>(P.dsphT16) : reinterpret_cast<const T *>(P.dsphT32);
  const T *__restrict__ dsTl = reinterpret_cast<const T *>(P.dsphT16lo);
  const T *__restrict__ dsTl2 = reinterpret_cast<const T *>(P.dsphT16lo2);
  const int fpad = P.fpad;
  // Dsph^T chunks pass through registers one chunk ahead: the loads of chunk
  // k+1 are in flight during chunk k's facet stage and contraction (fpad is
  // a multiple of the chunk, so a chunk never reads past the table)
  constexpr int kPieces = kFacetChunk / PER16;  // 16-byte pieces per row
  constexpr int kStageIt = (NC * kPieces + kThreads - 1) / kThreads;
  v4u pre[kStageIt], prel[kStageIt], prel2[SPLIT3 ? kStageIt : 1];
  // (threads past the last piece repeat it: identical loads and stores, no
  // branches, so the pieces stay in registers)
  auto load_chunk = [&](int fc) __attribute__((always_inline)) {
    static_for<0, kStageIt>([&](auto ic) __attribute__((always_inline)) {
      constexpr int it = decltype(ic)::value;
      const int idx = min(tid + it * kThreads, NC * kPieces - 1);
      const int c = idx / kPieces, pc = idx - c * kPieces;
      pre[it] = *reinterpret_cast<const v4u *>(dsT + (size_t)c * fpad + fc + pc * PER16);
      if (SPLIT) prel[it] = *reinterpret_cast<const v4u *>(dsTl + (size_t)c * fpad + fc + pc * PER16);
      if constexpr (SPLIT3) prel2[it] = *reinterpret_cast<const v4u *>(dsTl2 + (size_t)c * fpad + fc + pc * PER16);
    });
  };
  load_chunk(0);
  for (int f0 = 0; f0 < F; f0 += kFacetChunk) {
    static_for<0, kStageIt>([&](auto ic) __attribute__((always_inline)) {
      constexpr int it = decltype(ic)::value;
      const int idx = min(tid + it * kThreads, NC * kPieces - 1);
      const int c = idx / kPieces, pc = idx - c * kPieces;
      *reinterpret_cast<v4u *>(dch + c * WS + pc * PER16) = pre[it];
      if (SPLIT) *reinterpret_cast<v4u *>(dchl + c * WS + pc * PER16) = prel[it];
      if constexpr (SPLIT3) *reinterpret_cast<v4u *>(dchl2 + c * WS + pc * PER16) = prel2[it];
    });
    if (f0 + kFacetChunk < F) load_chunk(f0 + kFacetChunk);
    // facet stage (FP32).  The eight facets of a wave are wave-uniform: their
    // normals (scalar loads) and areas are fetched before the first use.
    // fb + 8 <= fpad (fpad is a multiple of the chunk), so the 24 floats of
    // the normals are inside the zero-padded table and need no clamp.
    const int fb = __builtin_amdgcn_readfirstlane(f0 + 8 * sub);
    float nv3[24];
#pragma unroll
    for (int k = 0; k < 24; k++) nv3[k] = n3[3 * fb + k];
    // areas of the eight facets (zero past the last facet, whose zero normals
    // are never visible)
    const float4 a03 = *reinterpret_cast<const float4 *>(areaf + fb);
    const float4 a47 = *reinterpret_cast<const float4 *>(areaf + fb + 4);
    const float afs[8] = {a03.x, a03.y, a03.z, a03.w, a47.x, a47.y, a47.z, a47.w};
    float wv[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      float w = 0.0f;
      {
        const float3 nv = make_float3(nv3[3 * i], nv3[3 * i + 1], nv3[3 * i + 2]);
        const float mu = e_0 * nv.x + e_1 * nv.y + e_2 * nv.z;
        const float mu0 = s_0 * nv.x + s_1 * nv.y + s_2 * nv.z;
        if (mu > tiny && mu0 > tiny) {
          // w = Area mu mu0 (cl + cls/(mu+mu0)) and its derivatives
          // dw/dmu = Area mu0 (cl + cls mu0/(mu+mu0)^2), dw/dmu0 likewise;
          // the brightness sum is cl*scl + cls*scls, formed after the loop
          const float af = afs[i];
          const float rd = __builtin_amdgcn_rcpf(mu + mu0);
          const float amu = af * mu, amu0 = af * mu0;
          const float am = amu * mu0;
          w = am * (cl + cls * rd);
          scl += am;
          scls += am * rd;
          const float k = cls * (rd * rd);
          const float ad = amu0 * (cl + k * mu0);
          const float ad0 = amu * (cl + k * mu);
          v0 += ad * nv.x;
          v1 += ad * nv.y;
          v2 += ad * nv.z;
          u0 += ad0 * nv.x;
          u1 += ad0 * nv.y;
          u2 += ad0 * nv.z;
        }
      }
      wv[i] = w;
    }
    if constexpr (BF) {
      // v_cvt_pk_bf16_f32: two weights per instruction, round to nearest even
      unsigned h[4];
#pragma unroll
      for (int q = 0; q < 4; q++) h[q] = bf16_pair(wv[2 * q], wv[2 * q + 1]);
      *reinterpret_cast<uint4 *>(wt + pl * WS + 8 * sub) = make_uint4(h[0], h[1], h[2], h[3]);
      if constexpr (SPLIT) {
        unsigned lo[4], lo2[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const float ra = wv[2 * q] - __uint_as_float(h[q] << 16);
          const float rb = wv[2 * q + 1] - __uint_as_float(h[q] & 0xffff0000u);
          lo[q] = bf16_pair(ra, rb);
          if constexpr (SPLIT3)  // what the low part leaves, exact in FP32
            lo2[q] = bf16_pair(ra - __uint_as_float(lo[q] << 16), rb - __uint_as_float(lo[q] & 0xffff0000u));
        }
        *reinterpret_cast<uint4 *>(wtl + pl * WS + 8 * sub) = make_uint4(lo[0], lo[1], lo[2], lo[3]);
        if constexpr (SPLIT3)
          *reinterpret_cast<uint4 *>(wtl2 + pl * WS + 8 * sub) = make_uint4(lo2[0], lo2[1], lo2[2], lo2[3]);
      }
    } else {
      *reinterpret_cast<float4 *>(wt + pl * WS + 8 * sub) = make_float4(wv[0], wv[1], wv[2], wv[3]);
      *reinterpret_cast<float4 *>(wt + pl * WS + 8 * sub + 4) = make_float4(wv[4], wv[5], wv[6], wv[7]);
    }
    __syncthreads();
    // contraction on the matrix cores
    const int ar = 16 * wave + (lane & 15), kg = lane >> 4;
    if constexpr (BF) {
      const v8bf a = *reinterpret_cast<const v8bf *>(wt + ar * WS + 8 * kg);
      v8bf al, al2;
      if constexpr (SPLIT) al = *reinterpret_cast<const v8bf *>(wtl + ar * WS + 8 * kg);
      if constexpr (SPLIT3) al2 = *reinterpret_cast<const v8bf *>(wtl2 + ar * WS + 8 * kg);
#pragma unroll
      for (int cb = 0; cb < NB; cb++) {
        const int br_ = (16 * cb + (lane & 15)) * WS + 8 * kg;
        const v8bf b = *reinterpret_cast<const v8bf *>(dch + br_);
        if constexpr (SPLIT3) {  // the third-order products first (smallest)
          const v8bf bl = *reinterpret_cast<const v8bf *>(dchl + br_);
          const v8bf bl2 = *reinterpret_cast<const v8bf *>(dchl2 + br_);
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al2, b, acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bl, acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl2, acc[cb], 0, 0, 0);
        }
        if constexpr (SPLIT) {
          const v8bf bl = *reinterpret_cast<const v8bf *>(dchl + br_);
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, b, acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl, acc[cb], 0, 0, 0);
        }
        acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[cb], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < kFacetChunk / 4; ks++) {
        const float a = wt[ar * WS + 4 * ks + kg];
#pragma unroll
        for (int cb = 0; cb < NB; cb++) {
          const float b = dch[(16 * cb + (lane & 15)) * WS + 4 * ks + kg];
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[cb], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  {
    float *rp = red + ((size_t)sub * kTilePts + pl) * 8;
    rp[0] = v0;
    rp[1] = v1;
    rp[2] = v2;
    rp[3] = u0;
    rp[4] = u1;
    rp[5] = u2;
    rp[6] = scl;
    rp[7] = scls;
  }
  __syncthreads();
  const int ld = P.npad;
  double *J = mem + M.J;                       // row ma (ymod) in FP64
  float *Jf = reinterpret_cast<float *>(J);    // rows 0..ma-1 in FP32
  if (tid < kTilePts && p0 + tid < P.n_points) {
    const int p = p0 + tid, pt = W.order ? perm[p] : p;  // position, point
    double t[8];
    for (int k = 0; k < 8; k++) t[k] = 0;
    for (int q = 0; q < kWaves; q++)
      for (int k = 0; k < 8; k++) t[k] += (double)red[((size_t)q * kTilePts + tid) * 8 + k];
    const double br = cld * t[6] + x[P.ma - 1] * t[7];  // sum of Area mu mu0 (cl + cls/(mu+mu0))
    const PointGeom G = gst[tid];
    const double sc = G.scale;
    const double V[3] = {t[0], t[1], t[2]}, U[3] = {t[3], t[4], t[5]};
    double r[3];
    rotation_derivs(P, pt, blm, G, V, U, r);
    for (int k = 0; k < 3; k++) Jf[(size_t)(P.ncoef + k) * ld + p] = (float)(sc * r[k]);
    for (int i = 0; i < 3; i++) Jf[(size_t)(P.ncoef + 3 + i) * ld + p] = (float)(br * G.dphp[i]);
    Jf[(size_t)(P.ncoef + 6) * ld + p] = (float)(sc * t[6] * cld);
    Jf[(size_t)(P.ncoef + 7) * ld + p] = (float)(sc * t[7]);
    J[(size_t)P.ma * ld + p] = br * sc;
  }
#pragma unroll
  for (int cb = 0; cb < NB; cb++) {
    const int c = 16 * cb + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int pp = 16 * wave + 4 * (lane >> 4) + r;
      const int p = p0 + pp;
      if (p < P.n_points && c < P.ncoef) Jf[(size_t)c * ld + p] = (float)(gst[pp].scale * (double)acc[cb][r]);
    }
  }
}

// Mixed-precision model with wave-owned points, single-part forms
// (PS_PREC_FP32: MODE 0, PS_PREC_BF16 and PS_PREC_BF16A: MODE 1; the split forms keep
// model_mixed_kernel, whose staging needs less LDS).  The model of
// period_search/bright.c:14-126 as model_mixed_kernel.  One workgroup per
// (slot, tile of kTilePts positions), blockIdx.x = tile * nslots + slot;
// wave w owns the tile's positions 16w..16w+15, lane l the point l%16 and
// the facet group g = l/16.
//   mu and mu0 on the matrix cores: per half-chunk of 16 facets one
//     v_mfma_f32_16x16x4f32 with A = the facets' normals (row m: facet
//     4(m%4) + m/4 of the half-chunk; columns x, y, z and the area, which
//     meets a zero row of B) and B = the directions e (or e0) of the wave's
//     16 points: lane l receives mu for its point and the facets 4r + g,
//     r = 0..3 (D lane l: rows 4(l/16)+r, column l%16).
//   facet stage (FP32 VALU): for the lane's eight facets 16h + 4r + g the
//     visibility test, and for a visible pair the weight
//     W = Area mu mu0 (cl + cls/(mu+mu0)) and its derivative terms.  One
//     step covers 16 points x 4 adjacent facets, and a step that no point of
//     the wave sees is jumped over as a whole.
//   contraction: the lane's eight weights are its A operand as they stand:
//     FP32 k-step 4h + r takes facet 16h + 4r + g (the natural order of the
//     Dsph^T table); bf16 k-slot 8g + j takes facet 4j + g (the order of the
//     chunk-permuted bf16 tables).
//   Dsph^T chunks are double-buffered in LDS (one barrier per chunk); the
//   normals and areas are one LDS table, the lane's eight entries read
//   together at the top of each chunk.
//   On gfx950 the matrix and the vector instructions of a SIMD share its
//   issue cycles (VALU and MFMA busy add up to 97% here), so the mu products
//   pay off by their cycle count: 32 per 256 values against 48 for the three
//   FMAs on the vector units.
struct MixedWLds {
  size_t x, blm, ftab, geom, dch, total;
};
__host__ __device__ inline MixedWLds mixedw_layout(int fpad, int nb, int mode) {
  MixedWLds L;
  size_t o = 0;
  const bool bf = mode != 0;
  const size_t es = bf ? 2 : 4, ws = kFacetChunk + (bf ? 8 : 4);
  L.x = o;
  o += kMaxMa * sizeof(double);
  L.blm = o;
  o += 32 * sizeof(double);
  L.ftab = o;  // [fpad] (nx, ny, nz, Area) FP32, zero past the last facet
  o = align16(o + (size_t)fpad * 16);
  L.geom = o;
  o = align16(o + kTilePts * sizeof(PointGeom));
  L.dch = o;  // [2][16 nb][ws]
  o = align16(o + 2 * 16 * (size_t)nb * ws * es);
  L.total = o;
  return L;
}

template <int NB, int MODE, bool MULTI>
__global__ void __launch_bounds__(kThreads, (MODE == 1 && NB <= 4) ? 5 : 4) model_mixedw_kernel(DevProblem P0,
                                                                             const DevProblem *__restrict__ probs,
                                                                             DevWork W) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smb[];
  const int slot = blockIdx.x % W.nslots, tile = blockIdx.x / W.nslots;
  if (!W.st[slot].eval) return;
  const DevProblem &P = pick_problem<MULTI>(P0, probs, W.st[slot].prob);
  if (tile >= P.ntiles64) return;
  static_assert(MODE == 0 || MODE == 1, "single-part forms");
  constexpr bool BF = MODE != 0;
  constexpr int NC = 16 * NB;
  constexpr int WS = kFacetChunk + (BF ? 8 : 4);  // padded row, elements (80 / 144 bytes)
  constexpr int BUF = NC * WS;                    // elements per buffer
  using T = typename std::conditional<BF, unsigned short, float>::type;
  constexpr int PER16 = 16 / (int)sizeof(T);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, pi = lane & 15, kg = lane >> 4;
  const SlotMem M = slot_mem(P.ma, P.mfit, P.npad);
  double *mem = W.mem + (size_t)slot * W.slot_doubles;
  const int F = P.nfac, fpad = P.fpad, nch = fpad / kFacetChunk;
  const MixedWLds L = mixedw_layout(fpad, NB, MODE);
  double *x = reinterpret_cast<double *>(smb + L.x), *blm = reinterpret_cast<double *>(smb + L.blm);
  float4 *ftab = reinterpret_cast<float4 *>(smb + L.ftab);
  const float *ftf = reinterpret_cast<const float *>(smb + L.ftab);
  PointGeom *gst = reinterpret_cast<PointGeom *>(smb + L.geom);
  T *dch = reinterpret_cast<T *>(smb + L.dch);
  const T *__restrict__ dsT = BF ? reinterpret_cast<const T *>(P.dsphQ16) : reinterpret_cast<const T *>(P.dsphT32);
  constexpr int kPieces = kFacetChunk / PER16;  // 16-byte pieces per row
  constexpr int kStageIt = (NC * kPieces + kThreads - 1) / kThreads;
  v4u pre[kStageIt];
  // (threads past the last piece repeat it: identical loads and stores)
  auto load_chunk = [&](int fc) __attribute__((always_inline)) {
    static_for<0, kStageIt>([&](auto ic) __attribute__((always_inline)) {
      constexpr int it = decltype(ic)::value;
      const int idx = min(tid + it * kThreads, NC * kPieces - 1);
      const int c = idx / kPieces, pc = idx - c * kPieces;
      pre[it] = *reinterpret_cast<const v4u *>(dsT + (size_t)c * fpad + fc + pc * PER16);
    });
  };
  auto store_chunk = [&](int b) __attribute__((always_inline)) {
    T *d = dch + b * BUF;
    static_for<0, kStageIt>([&](auto ic) __attribute__((always_inline)) {
      constexpr int it = decltype(ic)::value;
      const int idx = min(tid + it * kThreads, NC * kPieces - 1);
      const int c = idx / kPieces, pc = idx - c * kPieces;
      *reinterpret_cast<v4u *>(d + c * WS + pc * PER16) = pre[it];
    });
  };

  for (int l = tid; l < P.ma; l += kThreads) x[l] = mem[M.atry + l];
  for (int f = tid; f < fpad; f += kThreads)  // nor3f is zero-padded to fpad
    ftab[f] = make_float4(P.nor3f[3 * f], P.nor3f[3 * f + 1], P.nor3f[3 * f + 2], f < F ? (float)mem[M.area + f] : 0.0f);
  if (tid == 0) eval_blmatrix(mem[M.atry + P.ncoef], mem[M.atry + P.ncoef + 1], blm);
  load_chunk(0);
  store_chunk(0);
  if (nch > 1) load_chunk(kFacetChunk);
  __syncthreads();
  const int p0 = tile * kTilePts;
  const int *__restrict__ perm = reinterpret_cast<const int *>(mem + M.perm);
  if (tid < kTilePts) {
    const int q = min(p0 + tid, P.n_points - 1);
    PointGeom G;
    point_geometry(P, W.order ? perm[q] : q, x, blm, G);
    gst[tid] = G;
  }
  __syncthreads();
  const int pw = 16 * wave + pi;  // the lane's position in the tile
  // B operands of the mu and mu0 products: component g of the point's e and e0, zero for g = 3
  const float be = kg == 0 ? (float)gst[pw].e[0] : (kg == 1 ? (float)gst[pw].e[1] : (kg == 2 ? (float)gst[pw].e[2] : 0.0f));
  const float be0 = kg == 0 ? (float)gst[pw].e0[0] : (kg == 1 ? (float)gst[pw].e0[1] : (kg == 2 ? (float)gst[pw].e0[2] : 0.0f));
  const double cld = exp(x[P.ma - 2]);
  const float cl = (float)cld, cls = (float)x[P.ma - 1], tiny = (float)kTiny;
  float v0 = 0, v1 = 0, v2 = 0, u0 = 0, u1 = 0, u2 = 0, scl = 0, scls = 0;
  v4f acc[NB];
#pragma unroll
  for (int cb = 0; cb < NB; cb++) acc[cb] = v4f{0, 0, 0, 0};
  const int fa = 4 * (pi & 3) + (pi >> 2);  // facet of A row pi within a half-chunk
  const v4f z4 = {0, 0, 0, 0};
  for (int c = 0; c < nch; c++) {
    const int b = c & 1, f0 = c * kFacetChunk;
    if (c + 1 < nch) {
      store_chunk(b ^ 1);  // chunk c+1; its buffer was last read before the previous barrier
      if (c + 2 < nch) load_chunk(f0 + 2 * kFacetChunk);
    }
    const float na0 = ftf[(f0 + fa) * 4 + kg], na1 = ftf[(f0 + 16 + fa) * 4 + kg];
    // the lane's eight facets, read together (one wait for all of them)
    float4 t8[8];
#pragma unroll
    for (int i = 0; i < 8; i++) t8[i] = ftab[f0 + 4 * i + kg];
#pragma unroll
    for (int i = 0; i < 8; i++) asm volatile("" ::"v"(t8[i].x), "v"(t8[i].y), "v"(t8[i].z), "v"(t8[i].w));
    const v4f mua = __builtin_amdgcn_mfma_f32_16x16x4f32(na0, be, z4, 0, 0, 0);
    const v4f mub = __builtin_amdgcn_mfma_f32_16x16x4f32(na1, be, z4, 0, 0, 0);
    const v4f m0a = __builtin_amdgcn_mfma_f32_16x16x4f32(na0, be0, z4, 0, 0, 0);
    const v4f m0b = __builtin_amdgcn_mfma_f32_16x16x4f32(na1, be0, z4, 0, 0, 0);
    float wv[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const float mu = i < 4 ? mua[i & 3] : mub[i & 3], mu0 = i < 4 ? m0a[i & 3] : m0b[i & 3];
      float w = 0.0f;
      if (mu > tiny && mu0 > tiny) {
        // w = Area mu mu0 (cl + cls/(mu+mu0)) and its derivatives
        // dw/dmu = Area mu0 (cl + cls mu0/(mu+mu0)^2), dw/dmu0 likewise; the
        // brightness sum is cl*scl + cls*scls, formed after the loop
        const float4 t = t8[i];
        const float af = t.w;
        const float rd = __builtin_amdgcn_rcpf(mu + mu0);
        const float amu = af * mu, amu0 = af * mu0;
        const float am = amu * mu0;
        w = am * (cl + cls * rd);
        scl += am;
        scls += am * rd;
        const float k = cls * (rd * rd);
        const float ad = amu0 * (cl + k * mu0);
        const float ad0 = amu * (cl + k * mu);
        v0 += ad * t.x;
        v1 += ad * t.y;
        v2 += ad * t.z;
        u0 += ad0 * t.x;
        u1 += ad0 * t.y;
        u2 += ad0 * t.z;
      }
      wv[i] = w;
    }
    const T *d = dch + b * BUF;
    if constexpr (BF) {
      // weight j is facet 4j + g: k-slot 8g + j of the permuted chunk
      unsigned h[4];
#pragma unroll
      for (int q = 0; q < 4; q++) h[q] = bf16_pair(wv[2 * q], wv[2 * q + 1]);
      const v8bf a = __builtin_bit_cast(v8bf, make_uint4(h[0], h[1], h[2], h[3]));
#pragma unroll
      for (int cb = 0; cb < NB; cb++) {
        const v8bf bh = *reinterpret_cast<const v8bf *>(d + (16 * cb + pi) * WS + 8 * kg);
        acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh, acc[cb], 0, 0, 0);
      }
    } else {
      // k-step s takes facet 4s + g
#pragma unroll
      for (int ks = 0; ks < kFacetChunk / 4; ks++) {
#pragma unroll
        for (int cb = 0; cb < NB; cb++) {
          const float bv = d[(16 * cb + pi) * WS + 4 * ks + kg];
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(wv[ks], bv, acc[cb], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // the four facet groups of each point: lanes l, l^16, l^32, l^48 add the
  // same values, so every lane ends with the point's totals
  auto quad = [](float v) __attribute__((always_inline)) {
    v += __shfl_xor(v, 16, 64);
    return v + __shfl_xor(v, 32, 64);
  };
  v0 = quad(v0);
  v1 = quad(v1);
  v2 = quad(v2);
  u0 = quad(u0);
  u1 = quad(u1);
  u2 = quad(u2);
  scl = quad(scl);
  scls = quad(scls);
  const int ld = P.npad;
  double *J = mem + M.J;                     // row ma (ymod) in FP64
  float *Jf = reinterpret_cast<float *>(J);  // rows 0..ma-1 in FP32
  const int p = p0 + pw;                     // position
  if (kg == 0 && p < P.n_points) {
    const double br = cld * (double)scl + x[P.ma - 1] * (double)scls;  // sum of Area mu mu0 (cl + cls/(mu+mu0))
    const PointGeom G = gst[pw];
    const double sc = G.scale;
    const double V[3] = {v0, v1, v2}, U[3] = {u0, u1, u2};
    double r[3];
    rotation_derivs(P, W.order ? perm[p] : p, blm, G, V, U, r);
    for (int k = 0; k < 3; k++) Jf[(size_t)(P.ncoef + k) * ld + p] = (float)(sc * r[k]);
    for (int i = 0; i < 3; i++) Jf[(size_t)(P.ncoef + 3 + i) * ld + p] = (float)(br * G.dphp[i]);
    Jf[(size_t)(P.ncoef + 6) * ld + p] = (float)(sc * (double)scl * cld);
    Jf[(size_t)(P.ncoef + 7) * ld + p] = (float)(sc * (double)scls);
    J[(size_t)P.ma * ld + p] = br * sc;
  }
#pragma unroll
  for (int cb = 0; cb < NB; cb++) {
    const int c = 16 * cb + pi;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int pp = 16 * wave + 4 * kg + r;
      const int q = p0 + pp;
      if (q < P.n_points && c < P.ncoef) Jf[(size_t)c * ld + q] = (float)(gst[pp].scale * (double)acc[cb][r]);
    }
  }
}

// FP64 model with the facet stage and the contraction overlapped (the model
// of period_search/bright.c:14-126, as model_tile).  One workgroup per (slot,
// tile of kTilePts points), blockIdx.x = tile * nslots + slot; wave w owns
// the tile's points 16w..16w+15.
//   Lane l takes point l%16 and the facets f0 + 4k + l/16 (k = 0..7) of each
//   32-facet chunk.  That is the A operand of v_mfma_f64_16x16x4f64's k-step
//   k (A lane l: row l%16, k-group l/16), so the weights W[p][f] = Area_f
//   s_pf stay in registers and never pass through LDS.  While the matrix
//   cores contract chunk c (S[p][j] += sum_f W[p][f] Dsph[f][j], NB blocks of
//   16 columns, B lane l: row k-group l/16, column l%16), the same wave
//   evaluates chunk c+1's weights on the vector units.
//   With ncoef = 16 NB + 1 (R1 = 1: lmax 6 and 8 give 49 and 81) the last
//   column is one FP64 FMA per point-facet pair, not a 16-column block of
//   which 15 would be padding.
//   A group of four facets that no point of the wave sees (neither lit nor
//   seen) has all its weights zero: the wave skips the rest of their facet
//   stage and the k-step's MFMAs (wave-uniform branches; the skipped terms
//   are exact zeros, so the sums are the same).  With the fit's point order
//   (point_order) about 40% of the k-steps are skipped at the 5,000-point,
//   800-facet shape, against 21% in input order.
//   Dsph chunks are double-buffered in LDS (one barrier per chunk) and pass
//   through registers one chunk ahead; the chunks' normals and areas go
//   through a two-chunk LDS ring two chunks ahead of the contraction (the
//   facet stage runs one ahead).  About 51 KB of LDS at NB = 5, so three
//   workgroups share a CU.  D layout: lane l holds S[4r + l/16][l%16].
//   Tile positions are points in the fit's order (W.order) or input order;
//   the Jacobian's columns follow the positions.
struct F64Lds {
  size_t x, blm, ftab, geom, dch, d1, total;
};
// LDS row stride of a staged Dsph chunk, in doubles: an odd multiple of 16,
// so that the four rows one B-operand read touches fall on alternate halves
// of the banks
__host__ __device__ constexpr int f64_dsph_stride(int nb) { return (16 * nb) % 32 == 16 ? 16 * nb : 16 * nb + 16; }
__host__ __device__ inline F64Lds f64_layout(int fpad, int nb, int r1) {
  F64Lds L;
  size_t o = 0;
  L.x = o;
  o += kMaxMa * sizeof(double);
  L.blm = o;
  o += 32 * sizeof(double);
  L.ftab = o;  // [2][kFacetChunk] (nx, ny, nz, Area), zero past the last facet
  o += 2 * (size_t)kFacetChunk * sizeof(double4);
  L.geom = o;
  o = align16(o + kTilePts * sizeof(PointGeom));
  L.dch = o;  // [2][kFacetChunk][stride]
  o += 2 * (size_t)kFacetChunk * f64_dsph_stride(nb) * sizeof(double);
  L.d1 = o;  // [2][kFacetChunk][2]: column 16 nb and the table's zero beside it
  o += r1 ? 2 * (size_t)kFacetChunk * 2 * sizeof(double) : 0;
  L.total = o;
  return L;
}

template <int NB, int R1, bool MULTI>
__global__ void __launch_bounds__(kThreads, NB <= 5 ? 3 : 2) model_f64_kernel(DevProblem P0,
                                                                             const DevProblem *__restrict__ probs,
                                                                             DevWork W) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smb[];
  const int slot = blockIdx.x % W.nslots, tile = blockIdx.x / W.nslots;
  if (!W.st[slot].eval) return;
  const DevProblem &P = pick_problem<MULTI>(P0, probs, W.st[slot].prob);
  if (tile >= P.ntiles64) return;
  constexpr int NC = 16 * NB, DS = f64_dsph_stride(NB);
  constexpr int RP = 8 * (NB + R1);   // 16-byte pieces per row of the table (16 (NB + R1) doubles)
  constexpr int NP = NC / 2 + R1;     // pieces staged per row: the NB blocks, then column NC's piece
  constexpr int kPieces = kFacetChunk * NP;
  constexpr int kStageIt = (kPieces + kThreads - 1) / kThreads;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, pi = lane & 15, kq = lane >> 4;
  const SlotMem M = slot_mem(P.ma, P.mfit, P.npad);
  double *mem = W.mem + (size_t)slot * W.slot_doubles;
  const int F = P.nfac, nch = P.fpad / kFacetChunk;
  const F64Lds L = f64_layout(P.fpad, NB, R1);
  double *x = reinterpret_cast<double *>(smb + L.x), *blm = reinterpret_cast<double *>(smb + L.blm);
  double4 *ftab = reinterpret_cast<double4 *>(smb + L.ftab);
  PointGeom *gst = reinterpret_cast<PointGeom *>(smb + L.geom);
  double *dch = reinterpret_cast<double *>(smb + L.dch);
  double *d1 = reinterpret_cast<double *>(smb + L.d1);
  const v4u *__restrict__ dsq = reinterpret_cast<const v4u *>(P.dsphP);  // [fpad][16 (NB + R1)], zero-padded
  v4u pre[kStageIt];
  // (threads past the last piece repeat it: identical loads and stores)
  auto load_chunk = [&](int c) __attribute__((always_inline)) {
    static_for<0, kStageIt>([&](auto ic) __attribute__((always_inline)) {
      constexpr int it = decltype(ic)::value;
      const int idx = min(tid + it * kThreads, kPieces - 1);
      const int r = idx / NP, pc = idx - r * NP;
      pre[it] = dsq[(size_t)(c * kFacetChunk + r) * RP + pc];
    });
  };
  auto store_chunk = [&](int b) __attribute__((always_inline)) {
    static_for<0, kStageIt>([&](auto ic) __attribute__((always_inline)) {
      constexpr int it = decltype(ic)::value;
      const int idx = min(tid + it * kThreads, kPieces - 1);
      const int r = idx / NP, pc = idx - r * NP;
      double *dst = (R1 != 0 && pc == NC / 2) ? d1 + (b * kFacetChunk + r) * 2 : dch + (b * kFacetChunk + r) * DS + 2 * pc;
      *reinterpret_cast<v4u *>(dst) = pre[it];
    });
  };

  // facet f's normal and area (zero past the last facet), for the ring
  auto facet_entry = [&](int f) __attribute__((always_inline)) {
    double4 t = make_double4(0.0, 0.0, 0.0, 0.0);
    if (f < F) {
      t = P.nor4[f];
      t.w = mem[M.area + f];
    }
    return t;
  };
  for (int l = tid; l < P.ma; l += kThreads) x[l] = mem[M.atry + l];
  if (tid < 2 * kFacetChunk) ftab[tid] = facet_entry(tid);  // chunks 0 and 1
  if (tid == 0) eval_blmatrix(mem[M.atry + P.ncoef], mem[M.atry + P.ncoef + 1], blm);
  load_chunk(0);
  store_chunk(0);
  if (nch > 1) load_chunk(1);
  __syncthreads();
  const int p0 = tile * kTilePts;
  const int *__restrict__ perm = reinterpret_cast<const int *>(mem + M.perm);
  if (tid < kTilePts) {
    const int q = min(p0 + tid, P.n_points - 1);
    PointGeom G;
    point_geometry(P, W.order ? perm[q] : q, x, blm, G);
    gst[tid] = G;
  }
  __syncthreads();
  const int pw = 16 * wave + pi;  // the lane's position in the tile
  const double e_0 = gst[pw].e[0], e_1 = gst[pw].e[1], e_2 = gst[pw].e[2];
  const double s_0 = gst[pw].e0[0], s_1 = gst[pw].e0[1], s_2 = gst[pw].e0[2];
  const double cl = exp(x[P.ma - 2]), cls = x[P.ma - 1];
  double br = 0, v0 = 0, v1 = 0, v2 = 0, u0 = 0, u1 = 0, u2 = 0, scl = 0, scls = 0, s1 = 0;
  v4d acc[NB];
#pragma unroll
  for (int cb = 0; cb < NB; cb++) acc[cb] = v4d{0, 0, 0, 0};
  // weight of facet f at the lane's point, w = Area mu mu0 (cl + cls/(mu+mu0))
  // if the facet is lit and seen (mu, mu0 > kTiny), else 0, and its terms of
  // the partial sums: br, dbr/d(e) = sum Area mu0 (cl + cls mu0/(mu+mu0)^2) n,
  // dbr/d(e0) likewise, and the sums of Area mu mu0 (cl-derivative) and of
  // Area mu mu0/(mu+mu0) (cls-derivative).  A facet that is not visible
  // takes zero area and a unit denominator, so every term is a finite zero.
  // any: some lane of the wave sees its facet (wave-uniform).
  auto weight = [&](int f, bool &any) __attribute__((always_inline)) {
    const double4 t = ftab[f & (2 * kFacetChunk - 1)];  // ring slot of chunk f/32
    const double mu = e_0 * t.x + e_1 * t.y + e_2 * t.z;
    const double mu0 = s_0 * t.x + s_1 * t.y + s_2 * t.z;
    const bool vis = mu > kTiny && mu0 > kTiny;
    any = __ballot(vis) != 0;
    if (!any) return 0.0;
    const double af = vis ? t.w : 0.0;
    const double rd = recip(vis ? mu + mu0 : 1.0);
    const double amu = af * mu, amu0 = af * mu0, am = amu * mu0;
    const double w = am * (cl + cls * rd);
    br += w;
    scl += am;
    scls += am * rd;
    const double k = cls * (rd * rd);
    const double ad = amu0 * (cl + k * mu0), ad0 = amu * (cl + k * mu);
    v0 += ad * t.x;
    v1 += ad * t.y;
    v2 += ad * t.z;
    u0 += ad0 * t.x;
    u1 += ad0 * t.y;
    u2 += ad0 * t.z;
    return w;
  };
  // k-step k of chunk buffer b with the weights a of facets 4k + l/16
  auto contract = [&](int b, int k, double a) __attribute__((always_inline)) {
    const int kr = b * kFacetChunk + 4 * k + kq;
    const double *db = dch + kr * DS + pi;
    double bv[NB];
#pragma unroll
    for (int cb = 0; cb < NB; cb++) bv[cb] = db[16 * cb];
#pragma unroll
    for (int cb = 0; cb < NB; cb++) acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv[cb], acc[cb], 0, 0, 0);
    if constexpr (R1 != 0) s1 = fma(a, d1[kr * 2], s1);
  };
  // wa: the weights of the chunk being contracted, ka: its k-steps with a
  // visible facet (bit k)
  double wa[8];
  unsigned ka = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    bool any;
    wa[k] = weight(4 * k + kq, any);
    ka |= any ? 1u << k : 0u;
  }
  __syncthreads();  // every wave has read chunk 0's ring slot before iteration 0 refills it
  for (int c = 0; c + 1 < nch; c++) {
    const int b = c & 1;
    store_chunk(b ^ 1);  // chunk c+1; its buffer was last read before the previous barrier
    if (c + 2 < nch) load_chunk(c + 2);
    // chunk c+2's facet entries replace chunk c's, whose facet stage ran
    // before the previous barrier; stored before this iteration's barrier
    const bool fring = c + 2 < nch && tid < kFacetChunk;
    double4 fnext;
    if (fring) fnext = facet_entry((c + 2) * kFacetChunk + tid);
    const int fn = (c + 1) * kFacetChunk + kq;
    double wb[8];
    unsigned kb = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (ka & (1u << k)) contract(b, k, wa[k]);
      bool any;
      wb[k] = weight(fn + 4 * k, any);
      kb |= any ? 1u << k : 0u;
    }
    if (fring) ftab[b * kFacetChunk + tid] = fnext;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; k++) wa[k] = wb[k];
    ka = kb;
  }
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (ka & (1u << k)) contract((nch - 1) & 1, k, wa[k]);

  // the four facet groups of each point: lanes l, l^16, l^32, l^48 add the
  // same values, so every lane ends with the point's totals
  auto quad = [](double v) __attribute__((always_inline)) {
    v += __shfl_xor(v, 16, 64);
    return v + __shfl_xor(v, 32, 64);
  };
  br = quad(br);
  v0 = quad(v0);
  v1 = quad(v1);
  v2 = quad(v2);
  u0 = quad(u0);
  u1 = quad(u1);
  u2 = quad(u2);
  scl = quad(scl);
  scls = quad(scls);
  if constexpr (R1 != 0) s1 = quad(s1);
  const int ld = P.npad;
  double *J = mem + M.J;
  const int p = p0 + pw;  // position
  if (kq == 0 && p < P.n_points) {
    const PointGeom G = gst[pw];
    const double sc = G.scale;
    const double V[3] = {v0, v1, v2}, U[3] = {u0, u1, u2};
    double r[3];
    rotation_derivs(P, W.order ? perm[p] : p, blm, G, V, U, r);
    for (int k = 0; k < 3; k++) J[(size_t)(P.ncoef + k) * ld + p] = sc * r[k];
    for (int i = 0; i < 3; i++) J[(size_t)(P.ncoef + 3 + i) * ld + p] = br * G.dphp[i];
    J[(size_t)(P.ncoef + 6) * ld + p] = sc * scl * cl;
    J[(size_t)(P.ncoef + 7) * ld + p] = sc * scls;
    J[(size_t)P.ma * ld + p] = br * sc;
    if (R1 != 0 && NC < P.ncoef) J[(size_t)NC * ld + p] = sc * s1;
  }
#pragma unroll
  for (int cb = 0; cb < NB; cb++) {
    const int c = 16 * cb + pi;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int pp = 16 * wave + 4 * r + kq;
      const int q = p0 + pp;
      if (q < P.n_points && c < P.ncoef) J[(size_t)c * ld + q] = gst[pp].scale * acc[cb][r];
    }
  }
}

// neq_kernel: one workgroup per slot.  Convexity rows, normal equations at
// the trial parameters, then accept/reject and the driver bookkeeping.
// LDS: staged rows (neq_jn_doubles), aux [3*kNeqChunk], dave [ma+1], x [ma],
// scalars [4], dave/ave [ma].
#ifndef PS_NEQ_KS_WAVES
#define PS_NEQ_KS_WAVES 2
#endif
template <int NT, bool MIX, bool MULTI, int KS, bool ABF = false>
__device__ __forceinline__ void neq_body(const DevProblem &P0, const DevProblem *__restrict__ probs, const DevWork &W) {
  extern __shared__ double sm[];
  const int tid = threadIdx.x, slot = blockIdx.x;
  SlotState &S = W.st[slot];
  const int ph = S.phase;
  if (ph != PH_INIT && ph != PH_ITER) return;
  const DevProblem &P = pick_problem<MULTI>(P0, probs, S.prob);
  const int mfit = P.mfit, ma = P.ma, nc = P.ncoef;
  const SlotMem M = slot_mem(ma, mfit, P.npad);
  double *mem = W.mem + (size_t)slot * W.slot_doubles;
  double *wst = sm, *dave = wst + neq_scratch_doubles(mfit, MIX, KS) + 3 * kNeqChunk;
  double *x = dave + kMaxMa + 1, *scal = x + kMaxMa;  // scal[0] = chi2, [1..3] = chck
  double *drat = scal + 4;                            // [ma] dave/ave of the current light curve
  int *fidx = reinterpret_cast<int *>(drat + kMaxMa); // [kFidxPad] free parameter -> index into a[], zero past mfit
  const int eval = S.eval;
  const int sel = S.sel;
  double *atr = mem + (sel ? M.alpha0 : M.alpha1), *btr = mem + (sel ? M.beta0 : M.beta1);
  if (eval) {
    if (tid == 0) atomicAdd(W.done + 1, 1);
    for (int l = tid; l < ma; l += kThreads) x[l] = mem[M.atry + l];
    for (int jf = tid; jf < kFidxPad; jf += kThreads) fidx[jf] = jf < mfit ? P.fidx[jf] : 0;
    conv_rows<MIX>(P, mem + M.area, mem + M.g, mem + M.J, scal + 1, wst);
    __threadfence_block();
    __syncthreads();
    if constexpr (KS > 0 && MIX)
      // the points' brightness, sigma and weight/sigma^2 in the Jacobian's column order
      normal_equations_ks<KS, ABF>(P, mem + M.J, W.order ? mem + M.pys : P.y, W.order ? mem + M.psg : P.sig,
                              W.order ? mem + M.psw : P.swt, W.conw_rs ? W.conw_rs[S.prob] : W.conw_r, wst, dave,
                              drat, fidx, btr, scal, atr);
    else if constexpr (KS > 0)
      normal_equations_ks64<KS>(P, mem + M.J, W.order ? reinterpret_cast<const int *>(mem + M.perm) : nullptr,
                                W.conw_rs ? W.conw_rs[S.prob] : W.conw_r, wst, dave, drat, fidx, btr, scal, atr);
    else
      normal_equations_mfma<NT, MIX>(P, mem + M.J, W.order ? reinterpret_cast<const int *>(mem + M.perm) : nullptr,
                                     W.conw_rs ? W.conw_rs[S.prob] : W.conw_r, wst, dave, drat, fidx, btr, scal, atr);
    __threadfence_block();
    __syncthreads();
  }
  const double conw_r = W.conw_rs ? W.conw_rs[S.prob] : W.conw_r;
  double *a = mem + M.a, *atry = mem + M.atry, *alast = mem + M.alast;
  if (ph == PH_INIT) {
    if (W.mode == 1) {
      const int item = S.item;
      for (int e = tid; e < mfit * mfit; e += kThreads) W.ev_alpha[(size_t)item * mfit * mfit + e] = atr[e];
      for (int j = tid; j < mfit; j += kThreads) W.ev_beta[(size_t)item * mfit + j] = btr[j];
      if (tid == 0) {
        W.ev_chi[item] = scal[0];
        for (int k = 0; k < 3; k++) W.ev_chck[(size_t)item * 3 + k] = scal[1 + k];
        S.phase = PH_EMPTY;
        S.eval = 0;
        atomicAdd(W.done, 1);
      }
      return;
    }
    // first mrqmin call (mrqmin.c:36-51) and the driver's start values
    // (period_search_BOINC.cpp:860-866)
    for (int l = tid; l < nc; l += kThreads) alast[l] = a[l];
    if (tid == 0) {
      S.sel = 1 - sel;
      S.lam = P.alamda_start;
      S.ochisq = scal[0];
      S.chisq = scal[0];
      for (int k = 0; k < 3; k++) S.chck[k] = scal[1 + k];
      S.dev_old = 1e30;
      S.dev_new = 0;
      S.iter_diff = 1e40;
      S.rchisq = -1;
      S.niter = 0;
      S.err = 0;
      S.phase = PH_ITER;
      S.eval = 0;
    }
    return;
  }
  // accept or reject the trial (mrqmin.c:82-107)
  __shared__ int s_accept;
  if (eval) {
    for (int l = tid; l < nc; l += kThreads) alast[l] = x[l];
    if (tid == 0) s_accept = scal[0] < S.ochisq ? 1 : 0;
    __syncthreads();
    if (s_accept)
      for (int l = tid; l < ma; l += kThreads) a[l] = x[l];
    if (tid == 0) {
#pragma clang fp contract(off)
      S.chisq = scal[0];
      if (s_accept) {
        S.lam = S.lam / P.alamda_incr;
        S.sel = 1 - sel;
        for (int k = 0; k < 3; k++) S.chck[k] = scal[1 + k];
      } else {
        S.lam = P.alamda_incr * S.lam;
        S.chisq = S.ochisq;
      }
    }
    __syncthreads();
  }
  // driver bookkeeping (period_search_BOINC.cpp:868-892)
  __shared__ int s_improved;
  if (tid == 0) {
    S.niter = S.niter + 1;
    s_improved = ((S.niter == 1) || (S.chisq < S.ochisq)) ? 1 : 0;
  }
  __syncthreads();
  if (s_improved)
    for (int l = tid; l < nc; l += kThreads) alast[l] = a[l];
  if (tid == 0) {
#pragma clang fp contract(off)
    if (s_improved) {
      S.ochisq = S.chisq;
      const double c0 = S.chck[0], c1 = S.chck[1], c2 = S.chck[2];
      S.rchisq = S.chisq - (c0 * c0 + c1 * c1 + c2 * c2) * (conw_r * conw_r);
    }
    const double dev_new = sqrt(S.rchisq / (P.ndata - 3));
    S.dev_new = dev_new;
    if (S.dev_old - dev_new > 1e-10) {
      S.iter_diff = S.dev_old - dev_new;
      S.dev_old = dev_new;
    }
    S.eval = 0;
  }
  (void)atry;
}

// Waves per SIMD of the normal-equation kernels: the staged kernel 1; the
// split kernels: FP64 2 up to 4 column blocks (at 4, 25 registers spill:
// measured faster than one wave per SIMD), else 1; bf16a up to 4 blocks 3;
// otherwise PS_NEQ_KS_WAVES.
constexpr int neq_waves(bool mix, int ks, bool abf) {
  return ks == 0 ? 1 : (!mix ? (ks <= 4 ? 2 : 1) : ((abf && ks <= 4) ? 3 : PS_NEQ_KS_WAVES));
}

template <int NT, bool MIX, bool MULTI, int KS = 0, bool ABF = false>
__global__ void __launch_bounds__(kThreads, neq_waves(MIX, KS, ABF)) neq_kernel(DevProblem P0, const DevProblem *__restrict__ probs, DevWork W) {
  neq_body<NT, MIX, MULTI, KS, ABF>(P0, probs, W);
}

// iter_kernel: one workgroup per slot, the end of one LM iteration and the
// start of the next in one launch: the normal equations at the trial
// parameters with accept/reject and the driver bookkeeping (neq_body), then
// the loop predicate, the fit summary or refill, and the damped solve with
// the trial parameters and their shape (step_body).  For more than 63 free
// parameters, where both parts run on four waves (the four-wave register
// solver, BR rows per wave, or the LDS solver, BR = 0).  An LM iteration is
// then two launches: this kernel and the model kernel.
template <int NT, bool MIX, bool MULTI, int KS, int BR, bool ABF = false>
__global__ void __launch_bounds__(kThreads, (!MIX && KS > 4) ? 1 : 2) iter_kernel(DevProblem P0, const DevProblem *__restrict__ probs, DevWork W) {
  neq_body<NT, MIX, MULTI, KS, ABF>(P0, probs, W);
  __syncthreads();
  step_body<MULTI, 0, BR>(P0, probs, W);
}

// Device memory of ps_fit_batch_noalloc: counters, slot states and slot
// memory of its fits, in one fixed array per device, held until the process
// ends.  Host code never passes the array's name to a kernel (in host code it
// is only a stand-in for the symbol): it uses the device address that
// hipGetSymbolAddress returns at ps_create (ps_engine::scratch_dev).
constexpr size_t kScratchBytes = 64ull << 20;
constexpr size_t kScratchHeader = 256;  // counters [3 * kMaxGroups], the pole check flag at byte 128
__device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ void clear_flag(int *flag) { *flag = 0; }

__global__ void check_poles(const int *pole, int n, int n_poles, int *flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (pole[i] < 0 || pole[i] >= n_poles)) atomicOr(flag, 1);
}

__global__ void reset_batch(SlotState *st, int nslots, int *queue, int *done) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslots) {
    SlotState s{};
    s.phase = PH_EMPTY;
    s.item = -1;
    st[i] = s;
  }
  if (i == 0) {
    *queue = 0;
    done[0] = 0;
    done[1] = 0;
  }
}

// Best pole per trial period (period_search_BOINC.cpp:913-924): the first
// pole with the strictly smallest dev wins; lambda wrapped to [0,360).
__global__ void reduce_poles(const ps_fit_out *fits, int n_periods, int n_poles, int ndata, int precalc,
                             ps_row *rows) {
#pragma clang fp contract(off)
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_periods) return;
  double dev_best = 1e40, per_best = 0, dark_best = 0, la_best = 0, be_best = 0;
  for (int m = 0; m < n_poles; m++) {
    const ps_fit_out &o = fits[(size_t)n * n_poles + m];
    if (o.dev < dev_best) {
      dev_best = o.dev;
      per_best = o.prd;
      dark_best = o.dark / o.totarea * 100;
      la_best = o.la;
      be_best = o.be;
    }
  }
  if (la_best < 0) la_best += 360;
  if (precalc && isnan(dark_best)) dark_best = 1.0;
  ps_row r;
  r.period_h = 24 * per_best;
  r.rms = dev_best;
  r.chisq = dev_best * dev_best * (ndata - 3);
  r.dark = dark_best;
  r.lambda = la_best;
  r.beta = be_best;
  rows[n] = r;
}

// Proposal features: one workgroup per trial frequency; the points are
// folded at phase frac(24 freq t) into nbins bins whose sums of ynorm are
// kept in 2^16 fixed point (integer LDS atomics: the result does not depend
// on the order of the additions).
__global__ void __launch_bounds__(kThreads) lc_features(DevProblem P, const double *freq, int nbins, float c_alpha,
                                                         float c_aspect, float *out) {
  __shared__ int sum[256], cnt[256];
  const int i = blockIdx.x, tid = threadIdx.x;
  for (int b = tid; b < nbins; b += kThreads) {
    sum[b] = 0;
    cnt[b] = 0;
  }
  __syncthreads();
  const double fd = 24.0 * freq[i];  // rotations per day
  for (int p = tid; p < P.n_points; p += kThreads) {
    const double ph = P.tim[p] * fd;
    const double fr = ph - floor(ph);
    const int b = min((int)(fr * nbins), nbins - 1);
    atomicAdd(&sum[b], __float2int_rn(P.ynorm[p] * 65536.0f));
    atomicAdd(&cnt[b], 1);
  }
  __syncthreads();
  float *o = out + (size_t)i * (nbins + 2);
  for (int b = tid; b < nbins; b += kThreads) o[b] = cnt[b] ? (float)sum[b] / 65536.0f / (float)cnt[b] : 0.0f;
  if (tid == 0) {
    o[nbins] = c_alpha;
    o[nbins + 1] = c_aspect;
  }
}

// Trial frequencies and poles of periods n_from.. (period_search_BOINC.cpp:812, 817).
__global__ void make_items(double freq_start, double freq_step, int n_from, int n_periods, int n_poles,
                           double *freq, int *pole) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_periods * n_poles) return;
  const int n = n_from + i / n_poles;
  freq[i] = freq_start - (n - 1) * freq_step;
  pole[i] = i % n_poles;
}

using ModelFn = void (*)(DevProblem, const DevProblem *, DevWork);

template <bool MULTI>
ModelFn pick_model(int ncp) {
  switch (ncp) {
    case 8: return model_kernel<8, MULTI>;
    case 16: return model_kernel<16, MULTI>;
    case 24: return model_kernel<24, MULTI>;
    case 32: return model_kernel<32, MULTI>;
    case 40: return model_kernel<40, MULTI>;
    case 48: return model_kernel<48, MULTI>;
    case 56: return model_kernel<56, MULTI>;
    case 64: return model_kernel<64, MULTI>;
    case 72: return model_kernel<72, MULTI>;
    case 80: return model_kernel<80, MULTI>;
    case 88: return model_kernel<88, MULTI>;
    case 96: return model_kernel<96, MULTI>;
    case 104: return model_kernel<104, MULTI>;
    case 112: return model_kernel<112, MULTI>;
    case 120: return model_kernel<120, MULTI>;
    case 128: return model_kernel<128, MULTI>;
    default: return nullptr;
  }
}

template <bool MULTI>
ModelFn pick_mfma(int ncoef) {
  switch ((ncoef + 15) / 16) {
    case 1: return model_mfma_kernel<1, MULTI>;
    case 2: return model_mfma_kernel<2, MULTI>;
    case 3: return model_mfma_kernel<3, MULTI>;
    case 4: return model_mfma_kernel<4, MULTI>;
    case 5: return model_mfma_kernel<5, MULTI>;
    case 6: return model_mfma_kernel<6, MULTI>;
    case 7: return model_mfma_kernel<7, MULTI>;
    case 8: return model_mfma_kernel<8, MULTI>;
    default: return nullptr;
  }
}

// FP64 overlapped kernel: ncoef = 16 nb + 1 takes nb blocks and the
// vector column, any other count whole blocks (zero-padded)
template <bool MULTI>
ModelFn pick_f64(int ncoef) {
  if (ncoef % 16 == 1 && ncoef > 16) switch (ncoef / 16) {
      case 1: return model_f64_kernel<1, 1, MULTI>;
      case 2: return model_f64_kernel<2, 1, MULTI>;
      case 3: return model_f64_kernel<3, 1, MULTI>;
      case 4: return model_f64_kernel<4, 1, MULTI>;
      case 5: return model_f64_kernel<5, 1, MULTI>;
      case 6: return model_f64_kernel<6, 1, MULTI>;
      case 7: return model_f64_kernel<7, 1, MULTI>;
      default: return nullptr;
    }
  switch ((ncoef + 15) / 16) {
    case 1: return model_f64_kernel<1, 0, MULTI>;
    case 2: return model_f64_kernel<2, 0, MULTI>;
    case 3: return model_f64_kernel<3, 0, MULTI>;
    case 4: return model_f64_kernel<4, 0, MULTI>;
    case 5: return model_f64_kernel<5, 0, MULTI>;
    case 6: return model_f64_kernel<6, 0, MULTI>;
    case 7: return model_f64_kernel<7, 0, MULTI>;
    case 8: return model_f64_kernel<8, 0, MULTI>;
    default: return nullptr;
  }
}
size_t f64_lds(int nfac, int ncoef) {
  const int fpad = (nfac + kFacetChunk - 1) / kFacetChunk * kFacetChunk;
  const bool r1 = ncoef % 16 == 1 && ncoef > 16;
  return f64_layout(fpad, r1 ? ncoef / 16 : (ncoef + 15) / 16, r1 ? 1 : 0).total;
}

template <int MODE, bool MULTI>
ModelFn pick_mixedw(int ncoef) {
  switch ((ncoef + 15) / 16) {
    case 1: return model_mixedw_kernel<1, MODE, MULTI>;
    case 2: return model_mixedw_kernel<2, MODE, MULTI>;
    case 3: return model_mixedw_kernel<3, MODE, MULTI>;
    case 4: return model_mixedw_kernel<4, MODE, MULTI>;
    case 5: return model_mixedw_kernel<5, MODE, MULTI>;
    case 6: return model_mixedw_kernel<6, MODE, MULTI>;
    case 7: return model_mixedw_kernel<7, MODE, MULTI>;
    case 8: return model_mixedw_kernel<8, MODE, MULTI>;
    default: return nullptr;
  }
}

template <int MODE, bool MULTI>
ModelFn pick_mixed(int ncoef) {
  switch ((ncoef + 15) / 16) {
    case 1: return model_mixed_kernel<1, MODE, MULTI>;
    case 2: return model_mixed_kernel<2, MODE, MULTI>;
    case 3: return model_mixed_kernel<3, MODE, MULTI>;
    case 4: return model_mixed_kernel<4, MODE, MULTI>;
    case 5: return model_mixed_kernel<5, MODE, MULTI>;
    case 6: return model_mixed_kernel<6, MODE, MULTI>;
    case 7: return model_mixed_kernel<7, MODE, MULTI>;
    case 8: return model_mixed_kernel<8, MODE, MULTI>;
    default: return nullptr;
  }
}
size_t mfma_lds(int nfac, int ncoef) {
  const int nc = (ncoef + 15) / 16 * 16;
  size_t d = (size_t)kMaxMa + 32 + nfac + (size_t)kFacetChunk * kTilePts + (size_t)kFacetChunk * nc;
  size_t red = 4 * kTilePts * 9;
  if ((size_t)kFacetChunk * kTilePts + (size_t)kFacetChunk * nc < red) d += red;
  return d * sizeof(double) + kTilePts * sizeof(PointGeom);
}

size_t step_lds(int mfit, int nfac) {  // mfit = 0: a register solver (at most 136 doubles of LDS)
  if (mfit == 0) {
    const size_t fin = 2 * (size_t)nfac, reg = 2 * (kMaxMfit + 1) + 2 * kWaves;
    return (fin > reg ? fin : reg) * sizeof(double);
  }
  const size_t copies = solve_in_place(mfit) ? 1 : 2;
  const size_t solve = mfit ? copies * (size_t)mfit * (mfit + 1) + 4 * kWaves : 0, fin = 2 * (size_t)nfac;
  return (solve > fin ? solve : fin) * sizeof(double);
}
size_t model_lds(int nfac) {
  return ((size_t)kMaxMa + 32 + 2 * (size_t)nfac) * sizeof(double) + kThreads * sizeof(PointGeom);
}
size_t neq_lds(int mfit, bool mix, int ks) {
  return ((size_t)neq_scratch_doubles(mfit, mix, ks) + 3 * kNeqChunk + 3 * kMaxMa + 1 + 8 + kFidxPad / 2) *
         sizeof(double);
}

using StepFn = void (*)(DevProblem, const DevProblem *, DevWork);
// Solver kernel for mfit free parameters: the one-wave register solver with
// the smallest row count that holds them, or (mfit = 0 here) the LDS solver.
template <bool MULTI>
StepFn pick_step(int mfit, bool diag) {
  if (mfit <= 0 || mfit > kBlockSolveMax) return step_kernel<MULTI, 0>;
  if (mfit > kWaveSolveMax) {
    if (mfit <= 64) return diag ? step_kernel<MULTI, 0, 16, true> : step_kernel<MULTI, 0, 16>;
    if (mfit <= 88) return diag ? step_kernel<MULTI, 0, 22, true> : step_kernel<MULTI, 0, 22>;
    if (mfit <= 96) return diag ? step_kernel<MULTI, 0, 24, true> : step_kernel<MULTI, 0, 24>;
    return diag ? step_kernel<MULTI, 0, 32, true> : step_kernel<MULTI, 0, 32>;
  }
  if (mfit <= 16) return diag ? step_kernel<MULTI, 16, 0, true> : step_kernel<MULTI, 16>;
  if (mfit <= 24) return diag ? step_kernel<MULTI, 24, 0, true> : step_kernel<MULTI, 24>;
  if (mfit <= 32) return diag ? step_kernel<MULTI, 32, 0, true> : step_kernel<MULTI, 32>;
  if (mfit <= 40) return diag ? step_kernel<MULTI, 40, 0, true> : step_kernel<MULTI, 40>;
  if (mfit <= 48) return diag ? step_kernel<MULTI, 48, 0, true> : step_kernel<MULTI, 48>;
  if (mfit <= 56) return diag ? step_kernel<MULTI, 56, 0, true> : step_kernel<MULTI, 56>;
  return diag ? step_kernel<MULTI, 64, 0, true> : step_kernel<MULTI, 64>;
}

using NeqFn = void (*)(DevProblem, const DevProblem *, DevWork);
// Column blocks of the K-split normal equations for mfit free parameters
// (every form), or 0: the staged kernel.
inline int neq_ks_blocks(int mfit, bool staged) {
  const int nb = (mfit + 15) / 16;
  return !staged && nb <= 6 ? nb : 0;
}

template <bool MIX, bool MULTI>
NeqFn pick_neq(int mfit, int ks, bool abf) {
  if constexpr (MIX) {
    if (abf) switch (ks) {
        case 1: return neq_kernel<1, true, MULTI, 1, true>;
        case 2: return neq_kernel<1, true, MULTI, 2, true>;
        case 3: return neq_kernel<1, true, MULTI, 3, true>;
        case 4: return neq_kernel<1, true, MULTI, 4, true>;
        case 5: return neq_kernel<1, true, MULTI, 5, true>;
        case 6: return neq_kernel<1, true, MULTI, 6, true>;
        default: break;
      }
  }
  switch (ks) {
    case 1: return neq_kernel<1, MIX, MULTI, 1>;
    case 2: return neq_kernel<1, MIX, MULTI, 2>;
    case 3: return neq_kernel<1, MIX, MULTI, 3>;
    case 4: return neq_kernel<1, MIX, MULTI, 4>;
    case 5: return neq_kernel<1, MIX, MULTI, 5>;
    case 6: return neq_kernel<1, MIX, MULTI, 6>;
    default: break;
  }
  const int nb = (mfit + 15) / 16, ntile = nb * (nb + 1) / 2, nt = (ntile + 3) / 4;
  if (nt <= 1) return neq_kernel<1, MIX, MULTI>;
  if (nt <= 3) return neq_kernel<3, MIX, MULTI>;
  if (nt <= 6) return neq_kernel<6, MIX, MULTI>;
  if (nt <= 9) return neq_kernel<9, MIX, MULTI>;
  return nullptr;
}

// The fused iteration kernel for mfit > 63 free parameters (the solver's
// rows per wave from mfit, the normal equations as pick_neq chooses them).
template <bool MIX, bool MULTI>
NeqFn pick_iter(int mfit, int ks, bool abf) {
  if (mfit <= kWaveSolveMax || mfit > kMaxMfit) return nullptr;
  if (MIX && abf && ks == 4 && mfit <= 64) return iter_kernel<1, MIX, MULTI, 4, 16, true>;
  if (MIX && abf && ks == 5) return iter_kernel<1, MIX, MULTI, 5, 24, true>;
  if (MIX && abf && ks == 6) return iter_kernel<1, MIX, MULTI, 6, 24, true>;
  if (ks == 4 && mfit <= 64) return iter_kernel<1, MIX, MULTI, 4, 16>;
  if (ks == 5) return iter_kernel<1, MIX, MULTI, 5, 24>;
  if (ks == 6) return iter_kernel<1, MIX, MULTI, 6, 24>;
  if (ks == 0) {
    const int nb = (mfit + 15) / 16;
    if (nb == 4) return iter_kernel<3, MIX, MULTI, 0, 16>;
    if (nb <= 6) return iter_kernel<6, MIX, MULTI, 0, 24>;
    if (mfit <= kBlockSolveMax) return iter_kernel<9, MIX, MULTI, 0, 32>;
    return iter_kernel<9, MIX, MULTI, 0, 0>;
  }
  return nullptr;
}


// The kernels of one LM iteration, chosen for a set of problems: their
// largest coefficient count and free-parameter count pick the instances,
// their largest facet/tile counts size the grids.
struct KernelSet {
  ModelFn model = nullptr;
  NeqFn neq = nullptr;
  StepFn step = nullptr;
  NeqFn iter = nullptr;  // neq + step in one launch (more than 63 free parameters), or null
  size_t lds_step = 0, lds_model = 0, lds_neq = 0, lds_iter = 0;
  int order = 0;         // DevWork.order of the set's batches
  int step_threads = kThreads;
  bool use_mfma = true;
  DevProblem G{};  // maxima over the set's problems
};

// Size class of a problem: the kernel instances it needs (model column
// blocks, normal-equation tiles per wave, solver rows).
inline long long size_class(int ncoef, int mfit) {
  const int nb = (ncoef + 15) / 16;
  const int fb = (mfit + 15) / 16, nt = (fb * (fb + 1) / 2 + 3) / 4;
  const int rw = mfit > kBlockSolveMax ? 0
                 : (mfit > kWaveSolveMax ? (mfit <= 64 ? 116 : (mfit <= 96 ? 124 : 132))
                                         : (mfit <= 16 ? 16 : (mfit + 7) / 8 * 8));
  return ((long long)nb * 64 + nt) * 128 + rw;
}

}  // namespace

struct ps_engine {
  int device = 0;
  hipStream_t stream = nullptr;                   // = streams[0]
  hipStream_t streams[kMaxGroups] = {};           // one per slot group
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evf = nullptr, evg[kMaxGroups] = {};
  bool use_graph = true;  // replay each batch's poll from a HIP graph (PSAMD_GRAPH=0: plain launches)
  int groups = 1;  // slot groups (streams); PSAMD_GROUPS=1..4 overrides
  std::string err;
  int precision = PS_PREC_FP64;
  bool have_problem = false;
  DevProblem P{};                 // problem 0 (host copy of the device view)
  DevProblem G{};                 // maxima over the problems of the batch
  const DevProblem *d_probs = nullptr;  // [n_problems] device views
  int n_problems = 0;
  std::vector<int> prob_npoles;   // pole count of each problem
  int *d_iprob = nullptr;         // [item_cap] problem of each item
  double *d_conw = nullptr;       // [conw_cap] convexity weight per problem
  size_t iprob_cap = 0, conw_cap = 0;
  std::vector<void *> prob_allocs;
  int n_poles = 0, cu_count = 0, max_slots = 2048;
  double us_per_tick = 0;  // device wall clock period [us]
  float feat_alpha = 0, feat_aspect = 0;  // mean solar phase angle, mean |Earth latitude| [rad]
  float *d_feat = nullptr;
  size_t feat_cap = 0;
  size_t lds_step = 0, lds_model = 0, lds_neq = 0;
  int step_threads = 256;  // one wave when the solve fits in registers
  // several problems of more than one size class (at most one class per
  // stream): one kernel set per class, each class with its own slots, item
  // range and stream; empty otherwise
  std::vector<KernelSet> cls;
  std::vector<int> prob_cls;
  std::vector<int> cls_start;  // item range of each class in the current batch (items sorted by class)
  ModelFn model = nullptr;
  NeqFn neq = nullptr;
  void (*step)(DevProblem, const DevProblem *, DevWork) = nullptr;
  bool use_mfma = true;   // FP64: matrix-core contraction (default) or the VALU kernel
  bool neq_staged = false;  // mixed forms: the staged normal-equation kernel instead of the K-split one (PSAMD_NEQ=staged)
  bool fuse = false;        // neq + step in one launch where both run on four waves (PSAMD_FUSE=1); measured
                            // slower than the two launches (register pressure), so off by default
  NeqFn iter = nullptr;
  size_t lds_iter = 0;
  bool order_points = true;  // mixed forms: each fit's points ordered for the model kernel (PSAMD_ORDER=0: input order)
  bool diag_pivots = true;   // mixed forms: diagonal pivots in the four-wave solve (PSAMD_PIVOT=full: complete pivoting)
  bool f64_barrier = false;  // FP64: the kernel whose facet stage and contraction alternate across barriers
                             // (PSAMD_MODEL=mfma) instead of the overlapped one
  bool mixed_shared = false;  // FP32 and bf16 forms: the kernel whose waves share all 64 points
                              // (PSAMD_MIXED=shared) instead of the one with wave-owned points
  int order = 0;
  // slot memory
  SlotState *st = nullptr;
  double *mem = nullptr;
  size_t mem_slots = 0;
  long long slot_doubles = 0;
  int *counters = nullptr;      // [0] queue, [1] done
  int *host_done = nullptr;     // pinned
  // item buffers
  double *d_freq = nullptr;
  int *d_pole = nullptr;
  ps_fit_out *d_out = nullptr;
  ps_row *d_rows = nullptr;
  size_t item_cap = 0, row_cap = 0;
  double *d_cg = nullptr, *d_ea = nullptr, *d_eb = nullptr, *d_ec = nullptr, *d_ek = nullptr;
  size_t ev_cap = 0, cg_cap = 0, cgo_cap = 0;
  double *d_cgo = nullptr;
  double last_ms = 0;
  long long last_items = 0, last_iters = 0;
  // optional per-launch timing of the model kernel (HIP events on the
  // engine's stream)
  unsigned char *scratch_dev = nullptr;  // device address of this device's `scratch` array
  bool profile = false;
  // progress callback (ps_set_progress), called from the host poll loop
  ps_progress_fn progress = nullptr;
  void *progress_user = nullptr;
  std::vector<hipEvent_t> pev;
  double model_ms = 0;
  long long model_launches = 0, evals = 0;
};

namespace {

template <typename T>
int upload(ps_engine *e, const T *src, size_t n, const T **dst) {
  void *p = nullptr;
  HIP_TRY(hipMalloc(&p, n * sizeof(T) + 16));
  e->prob_allocs.push_back(p);
  if (n) HIP_TRY(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
  *dst = static_cast<const T *>(p);
  return PS_OK;
}

void free_problem(ps_engine *e) {
  for (void *p : e->prob_allocs) (void)hipFree(p);
  e->prob_allocs.clear();
  e->have_problem = false;
  e->d_probs = nullptr;
  e->n_problems = 0;
  e->cls.clear();
  e->prob_cls.clear();
}

// ncoef and mfit of a problem (as upload_problem counts them), for its size class
void problem_dims(const ps_problem &p, int &ncoef, int &mfit) {
  ncoef = 0;
  for (int m = 0; m <= p.mmax; m++) ncoef += (p.lmax - m + 1) * (m ? 2 : 1);
  mfit = ncoef - (p.onlyrel ? 1 : 0) + 2 + (p.ia_prd ? 1 : 0) + (p.ia_par[0] ? 1 : 0) + (p.ia_par[1] ? 1 : 0) +
         (p.ia_par[2] ? 1 : 0) + (p.ia_cl ? 1 : 0);
}

// Kernel instances, LDS sizes and grids for problems with maxima G.
int build_kernels(ps_engine *e, const DevProblem &G, int nb, bool multi, bool allow_valu, KernelSet &ks) {
  const int F = G.nfac, ncoef = G.ncoef, mfit = G.mfit;
  const bool mixed = e->precision != PS_PREC_FP64;
  const int mode = (e->precision == PS_PREC_BF16 || e->precision == PS_PREC_BF16A) ? 1
                   : (e->precision == PS_PREC_BF16X3 ? 2 : (e->precision == PS_PREC_BF16X6 ? 3 : 0));
  ks.G = G;
  ks.use_mfma = mixed || !allow_valu;
  if (!mixed && !ks.use_mfma)
    ks.model = pick_model<false>(G.ncp);
  else if (!mixed && e->f64_barrier)
    ks.model = multi ? pick_mfma<true>(ncoef) : pick_mfma<false>(ncoef);
  else if (!mixed)
    ks.model = multi ? pick_f64<true>(ncoef) : pick_f64<false>(ncoef);
  else if ((e->mixed_shared || mode >= 2) && multi)
    ks.model = mode == 1 ? pick_mixed<1, true>(ncoef)
               : (mode == 2 ? pick_mixed<2, true>(ncoef) : (mode == 3 ? pick_mixed<3, true>(ncoef) : pick_mixed<0, true>(ncoef)));
  else if (e->mixed_shared || mode >= 2)
    ks.model = mode == 1 ? pick_mixed<1, false>(ncoef)
               : (mode == 2 ? pick_mixed<2, false>(ncoef) : (mode == 3 ? pick_mixed<3, false>(ncoef) : pick_mixed<0, false>(ncoef)));
  else if (multi)
    ks.model = mode == 1 ? pick_mixedw<1, true>(ncoef) : pick_mixedw<0, true>(ncoef);
  else
    ks.model = mode == 1 ? pick_mixedw<1, false>(ncoef) : pick_mixedw<0, false>(ncoef);
  if (!ks.model) {
    e->err = "no model kernel for this coefficient count";
    return PS_ERR_LIMIT;
  }
  const bool wave_solve = mfit <= kWaveSolveMax, reg_solve = mfit <= kBlockSolveMax;
  ks.lds_step = step_lds(reg_solve ? 0 : mfit, F);
  const int fpad = (F + kFacetChunk - 1) / kFacetChunk * kFacetChunk;
  ks.lds_model = mixed ? (e->mixed_shared || mode >= 2 ? mixed_layout(F, nb, mode).total : mixedw_layout(fpad, nb, mode).total)
                       : (!ks.use_mfma ? model_lds(F) : (e->f64_barrier ? mfma_lds(F, ncoef) : f64_lds(F, ncoef)));
  const int neq_blocks = neq_ks_blocks(mfit, e->neq_staged);
  // ordered points: the mixed forms with the split normal equations, and
  // FP64 with the overlapped model kernel
  ks.order = e->order_points && (mixed ? neq_blocks > 0 : ks.use_mfma && !e->f64_barrier) ? 1 : 0;
  ks.lds_neq = neq_lds(mfit, mixed, neq_blocks);
  const size_t lds_max = 160 * 1024 - 1024;  // dynamic LDS per workgroup (static variables take the rest)
  if (ks.lds_step > lds_max || ks.lds_neq > lds_max || ks.lds_model > lds_max) {
    e->err = "problem needs more LDS than a CU has";
    return PS_ERR_LIMIT;
  }
  const bool diag = mixed && e->diag_pivots;  // FP64 keeps gauss_errc's complete pivoting
  const bool abf = e->precision == PS_PREC_BF16A && neq_blocks > 0;  // alpha's products in bf16
  ks.step = multi ? pick_step<true>(mfit, diag) : pick_step<false>(mfit, diag);
  ks.step_threads = wave_solve ? 64 : kThreads;
  if (e->fuse) {
    ks.iter = mixed ? (multi ? pick_iter<true, true>(mfit, neq_blocks, abf) : pick_iter<true, false>(mfit, neq_blocks, abf))
                    : (multi ? pick_iter<false, true>(mfit, neq_blocks, false) : pick_iter<false, false>(mfit, neq_blocks, false));
    ks.lds_iter = std::max(ks.lds_step, ks.lds_neq);
  }
  ks.neq = mixed ? (multi ? pick_neq<true, true>(mfit, neq_blocks, abf) : pick_neq<true, false>(mfit, neq_blocks, abf))
                 : (multi ? pick_neq<false, true>(mfit, neq_blocks, false) : pick_neq<false, false>(mfit, neq_blocks, false));
  if (!ks.neq) {
    e->err = "too many free parameters";
    return PS_ERR_LIMIT;
  }
  HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(ks.step), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)ks.lds_step));
  HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(ks.neq), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)ks.lds_neq));
  HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(ks.model), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)ks.lds_model));
  if (ks.iter)
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(ks.iter), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)ks.lds_iter));
  return PS_OK;
}

void grow(DevProblem &G, const DevProblem &D) {
  G.nfac = std::max(G.nfac, D.nfac);
  G.ncoef = std::max(G.ncoef, D.ncoef);
  G.ncp = std::max(G.ncp, D.ncp);
  G.ma = std::max(G.ma, D.ma);
  G.mfit = std::max(G.mfit, D.mfit);
  G.npad = std::max(G.npad, D.npad);
  G.n_lc = std::max(G.n_lc, D.n_lc);
  G.ntiles = std::max(G.ntiles, D.ntiles);
  G.ntiles64 = std::max(G.ntiles64, D.ntiles64);
  G.n_iter_max = std::max(G.n_iter_max, D.n_iter_max);
  G.n_iter_min = std::max(G.n_iter_min, D.n_iter_min);
}

int ensure_cg(ps_engine *e, size_t n) {
  if (n <= e->cg_cap) return PS_OK;
  if (e->d_cg) (void)hipFree(e->d_cg);
  e->d_cg = nullptr;
  e->cg_cap = 0;
  HIP_TRY(hipMalloc(&e->d_cg, n * kMaxMa * sizeof(double)));
  e->cg_cap = n;
  return PS_OK;
}

int ensure_items(ps_engine *e, size_t n) {
  if (n <= e->item_cap) return PS_OK;
  if (e->d_freq) (void)hipFree(e->d_freq);
  if (e->d_pole) (void)hipFree(e->d_pole);
  if (e->d_out) (void)hipFree(e->d_out);
  e->d_freq = nullptr;
  e->d_pole = nullptr;
  e->d_out = nullptr;
  e->item_cap = 0;
  size_t cap = n < 1024 ? 1024 : n;
  HIP_TRY(hipMalloc(&e->d_freq, cap * sizeof(double)));
  HIP_TRY(hipMalloc(&e->d_pole, cap * sizeof(int)));
  HIP_TRY(hipMalloc(&e->d_out, cap * sizeof(ps_fit_out)));
  e->item_cap = cap;
  return PS_OK;
}

int ensure_rows(ps_engine *e, size_t n) {
  if (n <= e->row_cap) return PS_OK;
  if (e->d_rows) (void)hipFree(e->d_rows);
  e->d_rows = nullptr;
  e->row_cap = 0;
  size_t cap = n < 256 ? 256 : n;
  HIP_TRY(hipMalloc(&e->d_rows, cap * sizeof(ps_row)));
  e->row_cap = cap;
  return PS_OK;
}

int ensure_slots(ps_engine *e, int nslots) {
  const SlotMem M = slot_mem(e->G.ma, e->G.mfit, e->G.npad);  // large enough for every problem's layout
  if ((size_t)nslots <= e->mem_slots && M.total == e->slot_doubles) return PS_OK;
  if (e->mem) (void)hipFree(e->mem);
  if (e->st) (void)hipFree(e->st);
  e->mem = nullptr;
  e->st = nullptr;
  e->mem_slots = 0;
  HIP_TRY(hipMalloc(&e->mem, (size_t)nslots * M.total * sizeof(double)));
  HIP_TRY(hipMalloc(&e->st, (size_t)nslots * sizeof(SlotState)));
  e->mem_slots = nslots;
  e->slot_doubles = M.total;
  return PS_OK;
}

// Run W.n_items items through the slots until all are done.
// Slot state, slot memory and counters taken from caller-provided device
// memory instead of the engine's own allocations (ps_fit_batch_noalloc).
struct FixedSlots {
  SlotState *st;
  double *mem;
  int *counters;
  int max_slots;
  long long slot_doubles;
};

int run_batch(ps_engine *e, DevWork W, const FixedSlots *fixed = nullptr) {
  if (W.n_items <= 0) return PS_OK;
  W.us_per_tick = e->us_per_tick;
  // as few refill rounds as max_slots allows, with the items spread evenly
  // over them (fits of equal length then finish together, no thin tail)
  const int max_slots = fixed ? fixed->max_slots : e->max_slots;
  const int nrounds = (W.n_items + max_slots - 1) / max_slots;
  int *const counters = fixed ? fixed->counters : e->counters;
  const DevProblem &P0 = e->P;  // the problem when there is one
  const DevProblem *probs = e->d_probs;
  // Lanes: each drives its own slots with its own kernel set on its own
  // stream.  Size classes (several problems of different sizes): one lane
  // per class over its contiguous item range, with its own queue and done
  // counters.  Otherwise the slots are split into up to `groups` lanes that
  // share the item queue and the done counters (atomics), so the
  // latency-bound per-slot kernels of one lane overlap the model kernels of
  // the others.
  struct Lane {
    const KernelSet *ks;
    DevWork W;
  };
  KernelSet eks;
  eks.model = e->model;
  eks.neq = e->neq;
  eks.step = e->step;
  eks.iter = e->iter;
  eks.lds_iter = e->lds_iter;
  eks.order = e->order;
  eks.lds_step = e->lds_step;
  eks.lds_model = e->lds_model;
  eks.lds_neq = e->lds_neq;
  eks.step_threads = e->step_threads;
  eks.use_mfma = e->use_mfma;
  eks.G = e->G;
  Lane lane[kMaxGroups];
  int nl = 0, nslots = 0;
  const bool by_class = !e->cls.empty() && W.item_prob && e->cls_start.size() == e->cls.size() + 1;
  if (by_class) {
    for (size_t c = 0; c < e->cls.size(); c++) {
      const int n_c = e->cls_start[c + 1] - e->cls_start[c];
      if (n_c <= 0) continue;
      Lane &L = lane[nl];
      L.ks = &e->cls[c];
      L.W = W;
      L.W.order = e->cls[c].order;
      L.W.n_items = n_c;
      const int start = e->cls_start[c];
      L.W.freq = W.freq + start;
      L.W.pole = W.pole + start;
      L.W.item_prob = W.item_prob + start;
      L.W.out = W.out + start;
      L.W.nslots = (n_c + nrounds - 1) / nrounds;
      L.W.queue = counters + 3 * nl;  // per lane: queue, fits done, evaluations
      L.W.done = counters + 3 * nl + 1;
      nslots += L.W.nslots;
      nl++;
    }
  } else {
    nslots = (W.n_items + nrounds - 1) / nrounds;
    int ng = e->groups;
    while (ng > 1 && nslots < ng * 2 * kWaves) ng--;
    for (int g = 0; g < ng; g++) {
      lane[g].ks = &eks;
      lane[g].W = W;
      lane[g].W.order = eks.order;
      lane[g].W.nslots = nslots / ng + (g < nslots % ng ? 1 : 0);
      lane[g].W.queue = counters;
      lane[g].W.done = counters + 1;
    }
    nl = ng;
  }
  int rc;
  if (!fixed && (rc = ensure_slots(e, nslots))) return rc;
  SlotState *const st0 = fixed ? fixed->st : e->st;
  double *const mem0 = fixed ? fixed->mem : e->mem;
  const long long sd = fixed ? fixed->slot_doubles : e->slot_doubles;
  for (int l = 0, base = 0; l < nl; l++) {
    lane[l].W.st = st0 + base;
    lane[l].W.mem = mem0 + (size_t)base * sd;
    lane[l].W.slot_doubles = sd;
    base += lane[l].W.nslots;
  }
  // clear every lane's slots and counters (shared counters: the same values again)
  for (int l = 0; l < nl; l++)
    hipLaunchKernelGGL(reset_batch, dim3((lane[l].W.nslots + 255) / 256), dim3(256), 0, e->stream, lane[l].W.st,
                       lane[l].W.nslots, lane[l].W.queue, lane[l].W.done);
  HIP_TRY(hipGetLastError());
  hipStream_t *sg = e->streams;
  HIP_TRY(hipEventRecord(e->evg[0], sg[0]));
  for (int l = 1; l < nl; l++) HIP_TRY(hipStreamWaitEvent(sg[l], e->evg[0], 0));
  const DevProblem &P = e->G;  // iteration bound: maxima over the problems
  const int per_fit = (P.n_iter_max > P.n_iter_min ? P.n_iter_max : P.n_iter_min) + 3;
  const long long rounds = nrounds + 1;
  const long long guard = rounds * per_fit * 2 + 64;
  const int poll = 8;
  // the fixed-memory path allocates nothing: no graph, no timing events
  const bool profile = e->profile && !fixed;
  if (profile && e->pev.size() < (size_t)(2 * kMaxGroups * poll)) {
    for (auto ev : e->pev) (void)hipEventDestroy(ev);
    e->pev.assign(2 * kMaxGroups * poll, nullptr);
    for (auto &ev : e->pev) HIP_TRY(hipEventCreate(&ev));
  }
  // One poll = `poll` LM iterations of every lane, then the lanes join
  // stream 0.  Without profiling events the poll is captured once per batch
  // into a HIP graph (its kernel arguments do not change between polls) and
  // replayed; profiling records events between the launches instead.
  auto enqueue_poll = [&](bool prof) -> int {
    for (int k = 0; k < poll; k++) {
      for (int l = 0; l < nl; l++) {
        const DevWork &Wk = lane[l].W;
        const KernelSet &K = *lane[l].ks;
        const int ns = Wk.nslots;
        // step, model, neq; or (fused) neq + step, then model
        if (K.iter)
          hipLaunchKernelGGL(K.iter, dim3(ns), dim3(kThreads), K.lds_iter, sg[l], P0, probs, Wk);
        else
          hipLaunchKernelGGL(K.step, dim3(ns), dim3(K.step_threads), K.lds_step, sg[l], P0, probs, Wk);
        if (prof) HIP_TRY(hipEventRecord(e->pev[2 * (kMaxGroups * k + l)], sg[l]));
        hipLaunchKernelGGL(K.model, dim3(ns * (K.use_mfma ? K.G.ntiles64 : K.G.ntiles)), dim3(kThreads),
                           K.lds_model, sg[l], P0, probs, Wk);
        if (prof) HIP_TRY(hipEventRecord(e->pev[2 * (kMaxGroups * k + l) + 1], sg[l]));
        if (!K.iter) hipLaunchKernelGGL(K.neq, dim3(ns), dim3(kThreads), K.lds_neq, sg[l], P0, probs, Wk);
      }
    }
    HIP_TRY(hipGetLastError());
    for (int l = 1; l < nl; l++) {
      HIP_TRY(hipEventRecord(e->evg[l], sg[l]));
      HIP_TRY(hipStreamWaitEvent(sg[0], e->evg[l], 0));
    }
    return PS_OK;
  };
  hipGraphExec_t gexec = nullptr;
  // Profiling times each model launch with HIP events between plain launches:
  // this runtime cannot time events captured into a graph
  // (hipEventElapsedTime on them returns hipErrorInvalidHandle), and the
  // kernels' durations do not depend on how they are launched.
  if (e->use_graph && !fixed && !profile) {
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(sg[0], hipStreamCaptureModeThreadLocal));
    // every call made while capturing is checked, but the capture is always
    // ended before an error is returned, so stream 0 never stays in capture
    // mode
    auto capture_body = [&]() -> int {
      HIP_TRY(hipEventRecord(e->evf, sg[0]));  // fork the other lanes' streams into the capture
      for (int l = 1; l < nl; l++) HIP_TRY(hipStreamWaitEvent(sg[l], e->evf, 0));
      return enqueue_poll(false);
    };
    const int rc_cap = capture_body();
    hipError_t err_end = hipStreamEndCapture(sg[0], &graph);
    if (rc_cap) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc_cap;
    }
    HIP_TRY(err_end);
    hipError_t err_inst = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(err_inst);
  }
  struct GraphGuard {
    hipGraphExec_t &g;
    ~GraphGuard() {
      if (g) (void)hipGraphExecDestroy(g);
    }
  } guard_graph{gexec};
  const int ncnt = by_class ? 3 * nl : 3;
  HIP_TRY(hipEventRecord(e->ev0, e->stream));
  long long it = 0;
  while (true) {
    if (gexec) {
      HIP_TRY(hipGraphLaunch(gexec, sg[0]));
    } else {
      const int rc_p = enqueue_poll(profile);
      if (rc_p) return rc_p;
    }
    it += poll;
    HIP_TRY(hipMemcpyAsync(e->host_done, counters, ncnt * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (profile) {
      for (int k = 0; k < poll; k++)
        for (int l = 0; l < nl; l++) {
          float ms = 0;
          HIP_TRY(hipEventElapsedTime(&ms, e->pev[2 * (kMaxGroups * k + l)], e->pev[2 * (kMaxGroups * k + l) + 1]));
          e->model_ms += ms;
        }
      e->model_launches += poll * nl;
    }
    // host_done: queue, fits done, evaluations (per lane with size classes)
    bool all = true;
    long long fits_done = 0;
    if (by_class) {
      for (int l = 0; l < nl; l++) {
        all = all && e->host_done[3 * l + 1] >= lane[l].W.n_items;
        fits_done += std::min(e->host_done[3 * l + 1], lane[l].W.n_items);
      }
    } else {
      all = e->host_done[1] >= W.n_items;
      fits_done = std::min(e->host_done[1], W.n_items);
    }
    if (e->progress) e->progress(e->progress_user, fits_done, W.n_items);
    if (all) break;
    if (it > guard) {
      e->err = "fit batch did not finish within its iteration bound";
      return PS_ERR_STATE;
    }
  }
  HIP_TRY(hipEventRecord(e->ev1, e->stream));
  HIP_TRY(hipEventSynchronize(e->ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
  e->last_ms = ms;
  e->last_items = W.n_items;
  e->last_iters = it;
  for (int l = 0; l < (by_class ? nl : 1); l++) e->evals += e->host_done[3 * l + 2];
  return PS_OK;
}

int check_ready(ps_engine *e, bool single = true) {
  if (!e) return PS_ERR_ARG;
  if (!e->have_problem) {
    e->err = "no problem set (ps_set_problem)";
    return PS_ERR_STATE;
  }
  if (single && e->n_problems != 1) {
    e->err = "this call works on one problem; several are set (ps_set_problems): use ps_fit_batch_multi";
    return PS_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(e->device));
  return PS_OK;
}

// Validates one problem and uploads its tables; D receives the device view.
// nb_tab: column blocks of 16 for the transposed Dsph tables (the largest
// problem of a batch, so one kernel instance serves every problem).
int upload_problem(ps_engine *e, const ps_problem *p, int nb_tab, DevProblem &D, float *feat) {
  if (!p) return PS_ERR_ARG;
  const int N = p->n_points, F = p->n_fac;
  if (N < 1 || p->n_lc < 1 || F < 1 || F > kMaxFacets || p->lmax < 0 || p->mmax < 0 || p->mmax > p->lmax ||
      p->n_poles < 1 || !p->lc_points || !p->lc_inrel || !p->tim || !p->brightness || !p->sig || !p->weight ||
      !p->ee || !p->ee0 || !p->nor || !p->darea || !p->pleg || !p->fc || !p->fs || !p->dsph || !p->cg_first ||
      !p->beta_pole || !p->lambda_pole) {
    e->err = "invalid problem";
    return PS_ERR_ARG;
  }
  int ncoef = 0, ngrp = 0;
  std::vector<int> grp_m;
  for (int m = 0; m <= p->mmax; m++)
    for (int l = m; l <= p->lmax; l++) {
      ncoef += (m != 0) ? 2 : 1;
      grp_m.push_back(m);
      ngrp++;
    }
  if (ncoef > kMaxNcoef) {
    e->err = "too many shape coefficients";
    return PS_ERR_LIMIT;
  }
  long long sum = 0;
  for (int i = 0; i < p->n_lc; i++) {
    if (p->lc_points[i] < 1) {
      e->err = "empty light curve";
      return PS_ERR_ARG;
    }
    sum += p->lc_points[i];
  }
  if (sum != N) {
    e->err = "light-curve point counts do not add up to n_points";
    return PS_ERR_ARG;
  }
  const int ma = ncoef + 8;
  // free-parameter mask (period_search_BOINC.cpp:789-802, 838-840)
  std::vector<int> ia(ma, 1), fidx;
  if (p->onlyrel) ia[0] = 0;
  ia[ncoef + 2] = p->ia_prd ? 1 : 0;
  for (int i = 0; i < 3; i++) ia[ncoef + 3 + i] = p->ia_par[i] ? 1 : 0;
  ia[ncoef + 6] = p->ia_cl ? 1 : 0;
  ia[ncoef + 7] = 0;
  for (int l = 0; l < ma; l++)
    if (ia[l]) fidx.push_back(l);
  const int mfit = (int)fidx.size();
  if (mfit < 1 || mfit > 128) {
    e->err = "the number of free parameters must be 1..128";
    return PS_ERR_LIMIT;
  }
  const int ncp = (ncoef + 7) / 8 * 8;
  const int nb = nb_tab;
  // host-side SoA staging
  std::vector<double> ex(N), ey(N), ez(N), sx(N), sy(N), sz(N);
  for (int i = 0; i < N; i++) {
    ex[i] = p->ee[3 * i];
    ey[i] = p->ee[3 * i + 1];
    ez[i] = p->ee[3 * i + 2];
    sx[i] = p->ee0[3 * i];
    sy[i] = p->ee0[3 * i + 1];
    sz[i] = p->ee0[3 * i + 2];
  }
  std::vector<int> lc_start(p->n_lc + 2), lc_inrel(p->n_lc + 1);
  lc_start[0] = 0;
  for (int i = 0; i < p->n_lc; i++) {
    lc_start[i + 1] = lc_start[i] + p->lc_points[i];
    lc_inrel[i] = p->lc_inrel[i] ? 1 : 0;
  }
  lc_start[p->n_lc + 1] = N + 3;
  lc_inrel[p->n_lc] = 0;
  std::vector<float> ynorm(N);
  double sa = 0, sb = 0;
  for (int i = 0; i < p->n_lc; i++) {
    double mean = 0;
    for (int q = lc_start[i]; q < lc_start[i + 1]; q++) mean += p->brightness[q];
    mean /= p->lc_points[i];
    for (int q = lc_start[i]; q < lc_start[i + 1]; q++) ynorm[q] = (float)(p->brightness[q] / mean - 1.0);
  }
  for (int q = 0; q < N; q++) {
    const double *ee = p->ee + 3 * q, *e0 = p->ee0 + 3 * q;
    double c = ee[0] * e0[0] + ee[1] * e0[1] + ee[2] * e0[2];
    c = c > 1 ? 1 : (c < -1 ? -1 : c);
    sa += acos(c);
    sb += fabs(asin(ee[2] > 1 ? 1 : (ee[2] < -1 ? -1 : ee[2])));
  }
  feat[0] = (float)(sa / N);
  feat[1] = (float)(sb / N);
  std::vector<double> nx(F), ny(F), nz(F), dsph((size_t)F * ncp, 0.0);
  std::vector<double4> nor4(F);
  std::vector<float> nor3f;
  const int fpad = (F + kFacetChunk - 1) / kFacetChunk * kFacetChunk, ncT = 16 * nb;
  std::vector<float> dsphT32((size_t)ncT * fpad, 0.0f);
  std::vector<unsigned short> dsphT16((size_t)ncT * fpad, 0), dsphT16lo((size_t)ncT * fpad, 0),
      dsphT16lo2((size_t)ncT * fpad, 0), dsphQ16((size_t)ncT * fpad, 0);
  std::vector<double> dsphP((size_t)fpad * ncT, 0.0);
  auto to_bf16 = [](float v) {
    unsigned u;
    memcpy(&u, &v, 4);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  };
  for (int f = 0; f < F; f++) {
    nx[f] = p->nor[3 * f];
    ny[f] = p->nor[3 * f + 1];
    nz[f] = p->nor[3 * f + 2];
    const int fi = f % kFacetChunk, fq = f - fi + 8 * (fi % 4) + fi / 4;  // position of f in the permuted chunk
    for (int c = 0; c < ncoef; c++) {
      const double v = p->dsph[(size_t)f * ncoef + c];
      dsph[(size_t)f * ncp + c] = v;
      dsphP[(size_t)f * ncT + c] = v;
      const float vf = (float)v;
      dsphT32[(size_t)c * fpad + f] = vf;
      const unsigned short hi = to_bf16(vf);
      const unsigned hb = (unsigned)hi << 16;
      float hf;
      memcpy(&hf, &hb, 4);
      dsphT16[(size_t)c * fpad + f] = hi;
      const unsigned short lo = to_bf16(vf - hf);
      dsphT16lo[(size_t)c * fpad + f] = lo;
      const unsigned lb = (unsigned)lo << 16;
      float lf;
      memcpy(&lf, &lb, 4);
      dsphT16lo2[(size_t)c * fpad + f] = to_bf16((vf - hf) - lf);
      dsphQ16[(size_t)c * fpad + fq] = hi;
    }
    nor4[f] = make_double4(nx[f], ny[f], nz[f], 0.0);
    nor3f.push_back((float)nx[f]);
    nor3f.push_back((float)ny[f]);
    nor3f.push_back((float)nz[f]);
  }
  const int L1 = p->lmax + 1, M1 = p->mmax + 1;
  std::vector<double> cpleg((size_t)ngrp * F), cfc((size_t)M1 * F), cfs((size_t)M1 * F);
  for (int f = 0; f < F; f++) {
    int q = 0;
    for (int m = 0; m <= p->mmax; m++)
      for (int l = m; l <= p->lmax; l++) cpleg[(size_t)(q++) * F + f] = p->pleg[((size_t)f * L1 + l) * L1 + m];
    for (int m = 0; m < M1; m++) {
      cfc[(size_t)m * F + f] = p->fc[(size_t)f * M1 + m];
      cfs[(size_t)m * F + f] = p->fs[(size_t)f * M1 + m];
    }
  }
  D = DevProblem{};
  int rc = PS_OK;
#define UP(src, n, dst) \
  if ((rc = upload(e, src, n, &D.dst)) != PS_OK) return rc;
  UP(p->tim, N, tim);
  UP(p->brightness, N, y);
  UP(p->sig, N, sig);
  UP(p->weight, N, w);
  {
    std::vector<double> swt(N);
    for (int q = 0; q < N; q++) swt[q] = (1 / (p->sig[q] * p->sig[q])) * p->weight[q];
    UP(swt.data(), N, swt);
  }
  UP(ex.data(), N, ex);
  UP(ey.data(), N, ey);
  UP(ez.data(), N, ez);
  UP(sx.data(), N, sx);
  UP(sy.data(), N, sy);
  UP(sz.data(), N, sz);
  UP(lc_start.data(), lc_start.size(), lc_start);
  UP(ynorm.data(), N, ynorm);
  UP(lc_inrel.data(), lc_inrel.size(), lc_inrel);
  UP(nx.data(), F, nx);
  UP(ny.data(), F, ny);
  UP(nz.data(), F, nz);
  UP(p->darea, F, darea);
  UP(dsph.data(), dsph.size(), dsph);
  UP(dsphP.data(), dsphP.size(), dsphP);
  UP(nor4.data(), F, nor4);
  nor3f.resize((size_t)3 * fpad, 0.0f);
  UP(nor3f.data(), nor3f.size(), nor3f);
  UP(dsphT32.data(), dsphT32.size(), dsphT32);
  UP(dsphT16.data(), dsphT16.size(), dsphT16);
  UP(dsphT16lo.data(), dsphT16lo.size(), dsphT16lo);
  UP(dsphT16lo2.data(), dsphT16lo2.size(), dsphT16lo2);
  UP(dsphQ16.data(), dsphQ16.size(), dsphQ16);
  UP(cpleg.data(), cpleg.size(), cpleg);
  UP(cfc.data(), cfc.size(), cfc);
  UP(cfs.data(), cfs.size(), cfs);
  UP(grp_m.data(), grp_m.size(), grp_m);
  UP(fidx.data(), fidx.size(), fidx);
  UP(p->cg_first, ncoef, cg_first);
  UP(p->beta_pole, p->n_poles, pole_beta);
  UP(p->lambda_pole, p->n_poles, pole_lambda);
#undef UP
  D.n_points = N;
  D.n_lc = p->n_lc;
  D.nfac = F;
  D.ngrp = ngrp;
  D.ncoef = ncoef;
  D.ncp = ncp;
  D.fpad = fpad;
  D.ma = ma;
  D.mfit = mfit;
  D.phi0 = p->phi0;
  D.alamda_start = p->alamda_start;
  D.alamda_incr = p->alamda_incr;
  D.log_cl = log(p->cl);
  D.par0 = p->par[0];
  D.par1 = p->par[1];
  D.par2 = p->par[2];
  // stop rule (period_search_BOINC.cpp:846-866)
  D.n_iter_max = 0;
  D.iter_diff_max = -1;
  D.n_iter_min = p->n_iter_min;
  if (p->stop_condition > 1) {
    D.n_iter_max = (int)p->stop_condition;
    D.iter_diff_max = 0;
    D.n_iter_min = 0;
  }
  if (p->stop_condition < 1) {
    D.n_iter_max = 100;
    D.iter_diff_max = p->stop_condition;
  }
  D.ndata = N + 3;
  D.npad = (N + 3 + 63) / 64 * 64;
  D.ntiles = (N + kThreads - 1) / kThreads;
  D.ntiles64 = (N + kTilePts - 1) / kTilePts;
  return PS_OK;
}

}  // namespace

void *psamd_engine_stream(ps_engine *e) { return e->stream; }
int psamd_engine_device(const ps_engine *e) { return e->device; }
void psamd_engine_error(ps_engine *e, const std::string &msg) { e->err = msg; }

extern "C" {

int ps_create(int device, ps_engine **out) {
  if (!out) return PS_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return PS_ERR_HIP;
  auto *e = new ps_engine();
  e->device = device;
  bool ok = hipSetDevice(device) == hipSuccess;
  for (int g = 0; g < kMaxGroups; g++)
    ok = ok && hipStreamCreateWithFlags(&e->streams[g], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&e->evg[g], hipEventDisableTiming) == hipSuccess;
  e->stream = e->streams[0];
  if (!ok || hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&e->evf, hipEventDisableTiming) != hipSuccess ||
      hipMalloc(&e->counters, 3 * kMaxGroups * sizeof(int)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&e->host_done), 3 * kMaxGroups * sizeof(int),
                    hipHostMallocDefault) != hipSuccess) {
    delete e;
    return PS_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete e;
    return PS_ERR_HIP;
  }
  e->cu_count = prop.multiProcessorCount;
  int wclk_khz = 0;
  if (hipDeviceGetAttribute(&wclk_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || wclk_khz <= 0) {
    delete e;
    return PS_ERR_HIP;
  }
  e->us_per_tick = 1e3 / wclk_khz;
  // the device address of the no-alloc scratch array, looked up once per
  // device (with that device current) and kept for every engine on it
  {
    static std::mutex mu;
    static std::vector<unsigned char *> per_device;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)per_device.size() <= device) per_device.resize(device + 1, nullptr);
    if (!per_device[device]) {
      void *ptr = nullptr;
      if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        (void)hipGetLastError();
        delete e;
        return PS_ERR_HIP;
      }
      per_device[device] = static_cast<unsigned char *>(ptr);
    }
    e->scratch_dev = per_device[device];
  }
  if (const char *gr = getenv("PSAMD_GRAPH")) e->use_graph = atoi(gr) != 0;
  if (const char *nq = getenv("PSAMD_NEQ")) e->neq_staged = strcmp(nq, "staged") == 0;
  if (const char *fu = getenv("PSAMD_FUSE")) e->fuse = atoi(fu) != 0;
  if (const char *po = getenv("PSAMD_ORDER")) e->order_points = atoi(po) != 0;
  if (const char *pv = getenv("PSAMD_PIVOT")) e->diag_pivots = strcmp(pv, "full") != 0;
  if (const char *mk = getenv("PSAMD_MODEL")) e->f64_barrier = strcmp(mk, "mfma") == 0;
  if (const char *mx = getenv("PSAMD_MIXED")) e->mixed_shared = strcmp(mx, "shared") == 0;
  if (const char *g = getenv("PSAMD_GROUPS")) {
    const int v = atoi(g);
    e->groups = v < 1 ? 1 : (v > kMaxGroups ? kMaxGroups : v);
  }
  *out = e;
  return PS_OK;
}

int ps_destroy(ps_engine *e) {
  if (!e) return PS_ERR_ARG;
  (void)hipSetDevice(e->device);
  for (int g = 0; g < kMaxGroups; g++)
    if (e->streams[g]) (void)hipStreamSynchronize(e->streams[g]);
  free_problem(e);
  void *bufs[] = {e->mem, e->st, e->counters, e->d_freq, e->d_pole, e->d_out, e->d_rows, e->d_feat, e->d_cgo,
                  e->d_iprob, e->d_conw,
                  e->d_cg, e->d_ea, e->d_eb, e->d_ec, e->d_ek};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  if (e->host_done) (void)hipHostFree(e->host_done);
  for (auto ev : e->pev) (void)hipEventDestroy(ev);
  if (e->ev0) (void)hipEventDestroy(e->ev0);
  if (e->ev1) (void)hipEventDestroy(e->ev1);
  if (e->evf) (void)hipEventDestroy(e->evf);
  for (int g = 0; g < kMaxGroups; g++) {
    if (e->evg[g]) (void)hipEventDestroy(e->evg[g]);
    if (e->streams[g]) (void)hipStreamDestroy(e->streams[g]);
  }
  delete e;
  return PS_OK;
}

const char *ps_last_error(const ps_engine *e) { return e ? e->err.c_str() : "null engine"; }

int ps_set_precision(ps_engine *e, int precision) {
  if (!e) return PS_ERR_ARG;
  if (precision < PS_PREC_FP64 || precision > PS_PREC_BF16A) {
    e->err = "unknown precision";
    return PS_ERR_ARG;
  }
  e->precision = precision;
  return PS_OK;
}

int ps_set_groups(ps_engine *e, int groups) {
  if (!e || groups < 1 || groups > kMaxGroups) return PS_ERR_ARG;
  e->groups = groups;
  return PS_OK;
}

int ps_set_slots(ps_engine *e, int max_slots) {
  if (!e || max_slots < 1) return PS_ERR_ARG;
  e->max_slots = max_slots;
  return PS_OK;
}

int ps_set_problems(ps_engine *e, const ps_problem *p, int n) {
  if (!e || !p || n < 1) return PS_ERR_ARG;
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipStreamSynchronize(e->stream));
  free_problem(e);
  // size classes: with 2..kMaxGroups of them each class gets its own kernel
  // instances (its problems' tables built for them); otherwise the largest
  // coefficient count decides the instances and the tables of every problem
  std::vector<long long> key(n);
  std::vector<int> pnc(n), pmf(n);
  int nc_max = 0;
  for (int i = 0; i < n; i++) {
    if (p[i].lmax < 0 || p[i].mmax < 0 || p[i].mmax > p[i].lmax) {
      e->err = "invalid problem";
      return PS_ERR_ARG;
    }
    problem_dims(p[i], pnc[i], pmf[i]);
    key[i] = size_class(pnc[i], pmf[i]);
    nc_max = std::max(nc_max, pnc[i]);
  }
  std::vector<long long> keys(key);
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const bool by_class = n > 1 && keys.size() > 1 && keys.size() <= (size_t)kMaxGroups && getenv("PSAMD_NOCLASS") == nullptr;
  std::vector<int> cls(n, 0);
  for (int i = 0; i < n; i++) cls[i] = (int)(std::lower_bound(keys.begin(), keys.end(), key[i]) - keys.begin());
  const int nb_all = (nc_max + 15) / 16;
  std::vector<DevProblem> hp(n);
  DevProblem G{};  // maxima over the batch (grids, slot size, iteration bound)
  std::vector<DevProblem> Gc(keys.size());
  std::vector<bool> seen(keys.size(), false);
  int rc;
  for (int i = 0; i < n; i++) {
    float feat[2];
    const int nb_tab = by_class ? (pnc[i] + 15) / 16 : nb_all;
    if ((rc = upload_problem(e, p + i, nb_tab, hp[i], feat))) {
      free_problem(e);
      return rc;
    }
    if (i == 0) {
      e->feat_alpha = feat[0];
      e->feat_aspect = feat[1];
      G = hp[0];
    }
    grow(G, hp[i]);
    const int c = cls[i];
    if (!seen[c]) {
      Gc[c] = hp[i];
      seen[c] = true;
    }
    grow(Gc[c], hp[i]);
  }
  const DevProblem *dp = nullptr;
  if ((rc = upload(e, hp.data(), (size_t)n, &dp)) != PS_OK) {
    free_problem(e);
    return rc;
  }
  e->d_probs = dp;
  e->P = hp[0];
  e->G = G;
  e->n_problems = n;
  e->n_poles = p[0].n_poles;
  e->prob_npoles.resize(n);
  for (int i = 0; i < n; i++) e->prob_npoles[i] = p[i].n_poles;
  const char *kv = getenv("PSAMD_MODEL");  // "valu" selects the VALU FP64 kernel (comparison runs, one problem)
  const bool allow_valu = n == 1 && kv && !strcmp(kv, "valu");
  KernelSet ks;
  if (by_class) {
    e->cls.resize(keys.size());
    for (size_t c = 0; c < keys.size(); c++)
      if ((rc = build_kernels(e, Gc[c], (Gc[c].ncoef + 15) / 16, true, false, e->cls[c]))) {
        free_problem(e);
        return rc;
      }
    e->prob_cls = cls;
    ks = e->cls.back();  // the engine-wide set is not used for several classes; keep a valid one
  } else if ((rc = build_kernels(e, G, nb_all, n > 1, allow_valu, ks))) {
    free_problem(e);
    return rc;
  }
  e->model = ks.model;
  e->neq = ks.neq;
  e->step = ks.step;
  e->iter = ks.iter;
  e->lds_iter = ks.lds_iter;
  e->order = ks.order;
  e->lds_step = ks.lds_step;
  e->lds_model = ks.lds_model;
  e->lds_neq = ks.lds_neq;
  e->step_threads = ks.step_threads;
  e->use_mfma = ks.use_mfma;
  e->have_problem = true;
  return PS_OK;
}

int ps_set_problem(ps_engine *e, const ps_problem *p) {
  if (!e || !p) return PS_ERR_ARG;
  return ps_set_problems(e, p, 1);
}

int ps_dims(const ps_engine *e, int *ncoef, int *ma, int *mfit) {
  if (!e || !e->have_problem) return PS_ERR_STATE;
  if (ncoef) *ncoef = e->P.ncoef;
  if (ma) *ma = e->P.ma;
  if (mfit) *mfit = e->P.mfit;
  return PS_OK;
}

int ps_last_kernel_ms(const ps_engine *e, double *ms, long long *fits) {
  if (!e) return PS_ERR_ARG;
  if (ms) *ms = e->last_ms;
  if (fits) *fits = e->last_items;
  return PS_OK;
}

int ps_last_batch(const ps_engine *e, double *ms, long long *items, long long *iterations) {
  if (!e) return PS_ERR_ARG;
  if (ms) *ms = e->last_ms;
  if (items) *items = e->last_items;
  if (iterations) *iterations = e->last_iters;
  return PS_OK;
}

int ps_set_progress(ps_engine *e, ps_progress_fn fn, void *user) {
  if (!e) return PS_ERR_ARG;
  e->progress = fn;
  e->progress_user = user;
  return PS_OK;
}

int ps_set_profiling(ps_engine *e, int on) {
  if (!e) return PS_ERR_ARG;
  e->profile = on != 0;
  e->model_ms = 0;
  e->model_launches = 0;
  e->evals = 0;
  return PS_OK;
}

int ps_kernel_stats(const ps_engine *e, double *model_ms, long long *model_launches, long long *evals) {
  if (!e) return PS_ERR_ARG;
  if (model_ms) *model_ms = e->model_ms;
  if (model_launches) *model_launches = e->model_launches;
  if (evals) *evals = e->evals;
  return PS_OK;
}

int ps_fit_batch(ps_engine *e, const double *freq, const int *pole, int n, double conw_r, ps_fit_out *out) {
  int rc = check_ready(e);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!freq || !pole || !out))) return PS_ERR_ARG;
  if (n == 0) return PS_OK;
  for (int i = 0; i < n; i++)
    if (pole[i] < 0 || pole[i] >= e->n_poles) {
      e->err = "pole index out of range";
      return PS_ERR_ARG;
    }
  if ((rc = ensure_items(e, n))) return rc;
  HIP_TRY(hipMemcpyAsync(e->d_freq, freq, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipMemcpyAsync(e->d_pole, pole, n * sizeof(int), hipMemcpyHostToDevice, e->stream));
  DevWork W{};
  W.n_items = n;
  W.mode = 0;
  W.freq = e->d_freq;
  W.pole = e->d_pole;
  W.conw_r = conw_r;
  W.out = e->d_out;
  if ((rc = run_batch(e, W))) return rc;
  HIP_TRY(hipMemcpyAsync(out, e->d_out, n * sizeof(ps_fit_out), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return PS_OK;
}

int ps_fit_batch_noalloc(ps_engine *e, const double *freq, const int *pole, int n, double conw_r, ps_fit_out *out) {
  int rc = check_ready(e);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!freq || !pole || !out))) return PS_ERR_ARG;
  if (n == 0) return PS_OK;
  if (!e->scratch_dev) {
    e->err = "no scratch memory on this device";
    return PS_ERR_STATE;
  }
  const long long sd = slot_mem(e->G.ma, e->G.mfit, e->G.npad).total;
  const size_t st_bytes = (sizeof(SlotState) + 255) / 256 * 256;  // per slot, state rounded up
  const size_t per_slot = st_bytes + (size_t)sd * sizeof(double);
  const long long fit_slots = (long long)((kScratchBytes - kScratchHeader) / per_slot);
  if (fit_slots < 1) {
    e->err = "ps_fit_batch_noalloc: one fit of this problem needs " + std::to_string(per_slot >> 10) +
             " KiB of slot memory, more than the " + std::to_string(kScratchBytes >> 20) + " MiB scratch holds";
    return PS_ERR_LIMIT;
  }
  const int max_slots = (int)std::min<long long>(fit_slots, e->max_slots);
  unsigned char *base = e->scratch_dev;
  int *counters = reinterpret_cast<int *>(base);
  int *flag = reinterpret_cast<int *>(base + 128);
  FixedSlots fx;
  fx.counters = counters;
  fx.st = reinterpret_cast<SlotState *>(base + kScratchHeader);
  fx.mem = reinterpret_cast<double *>(base + kScratchHeader + (size_t)max_slots * st_bytes);
  fx.max_slots = max_slots;
  fx.slot_doubles = sd;
  // pole indices are read by the kernels: check them on the device first
  hipLaunchKernelGGL(clear_flag, dim3(1), dim3(1), 0, e->stream, flag);
  hipLaunchKernelGGL(check_poles, dim3((n + 255) / 256), dim3(256), 0, e->stream, pole, n, e->n_poles, flag);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(e->host_done, flag, sizeof(int), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (e->host_done[0]) {
    e->err = "pole index out of range";
    return PS_ERR_ARG;
  }
  DevWork W{};
  W.n_items = n;
  W.mode = 0;
  W.freq = freq;
  W.pole = pole;
  W.conw_r = conw_r;
  W.out = out;
  if ((rc = run_batch(e, W, &fx))) return rc;
  HIP_TRY(hipStreamSynchronize(e->stream));
  return PS_OK;
}

// Fits of periods n_from..n_to and their pole reduction, rows left on the device.
static int run_periods(ps_engine *e, double freq_start, double freq_step, int n_from, int n_to, double conw_r,
                       int precalc) {
  const int np = n_to - n_from + 1;
  const size_t items = (size_t)np * e->n_poles;
  int rc;
  if ((rc = ensure_items(e, items))) return rc;
  if ((rc = ensure_rows(e, np))) return rc;
  hipLaunchKernelGGL(make_items, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, e->stream, freq_start,
                     freq_step, n_from, np, e->n_poles, e->d_freq, e->d_pole);
  DevWork W{};
  W.n_items = (int)items;
  W.mode = 0;
  W.freq = e->d_freq;
  W.pole = e->d_pole;
  W.conw_r = conw_r;
  W.out = e->d_out;
  if ((rc = run_batch(e, W))) return rc;
  hipLaunchKernelGGL(reduce_poles, dim3((np + 127) / 128), dim3(128), 0, e->stream, e->d_out, np, e->n_poles,
                     e->P.ndata, precalc, e->d_rows);
  HIP_TRY(hipGetLastError());
  return PS_OK;
}

int ps_scan(ps_engine *e, double freq_start, double freq_step, int n_from, int n_to, double conw_r, ps_row *rows) {
  int rc = check_ready(e);
  if (rc) return rc;
  if (n_from < 1 || n_to < n_from || !rows) return PS_ERR_ARG;
  if ((rc = run_periods(e, freq_start, freq_step, n_from, n_to, conw_r, 0))) return rc;
  HIP_TRY(hipMemcpyAsync(rows, e->d_rows, (size_t)(n_to - n_from + 1) * sizeof(ps_row), hipMemcpyDeviceToHost,
                         e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return PS_OK;
}

int ps_precalc(ps_engine *e, double freq_start, double freq_end, double freq_step, double *conw_r, int *new_conw) {
  return ps_precalc_resume(e, freq_start, freq_end, freq_step, 0, 0.0, conw_r, new_conw);
}

int ps_precalc_resume(ps_engine *e, double freq_start, double freq_end, double freq_step, int done_periods,
                      double sum_dark_done, double *conw_r, int *new_conw) {
  int rc = check_ready(e);
  if (rc) return rc;
  if (!conw_r || !new_conw || done_periods < 0) return PS_ERR_ARG;
  // period_search_BOINC.cpp:571-712; a round resumed after done_periods
  // periods continues from their sum
  // (period_search_optimization_cpu/period_search/period_search_BOINC.cpp:627-640, 772)
  int max_test_periods = 10;
  int n_iter = (int)((freq_start - freq_end) / freq_step) + 1;
  if (n_iter < max_test_periods) max_test_periods = n_iter;
  if (max_test_periods < 1 || done_periods > max_test_periods) return PS_ERR_ARG;
  std::vector<ps_row> rows(max_test_periods);
  if (done_periods < max_test_periods) {
    if ((rc = run_periods(e, freq_start, freq_step, done_periods + 1, max_test_periods, *conw_r, 1))) return rc;
    HIP_TRY(hipMemcpyAsync(rows.data(), e->d_rows, (max_test_periods - done_periods) * sizeof(ps_row),
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
  }
  double sum_dark = sum_dark_done;
  for (int n = 0; n < max_test_periods - done_periods; n++) sum_dark = sum_dark + rows[n].dark;
  double ave_dark = sum_dark / max_test_periods;
  *new_conw = 0;
  if (ave_dark < 1.0) *new_conw = 1;
  if (ave_dark >= 1.0) *conw_r = *conw_r * 2;
  return PS_OK;
}

int ps_fit_batch_multi(ps_engine *e, const int *prob, const double *freq, const int *pole, int n,
                       const double *conw_r, ps_fit_out *out) {
  int rc = check_ready(e, false);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!prob || !freq || !pole || !conw_r || !out))) return PS_ERR_ARG;
  if (n == 0) return PS_OK;
  for (int i = 0; i < n; i++)
    if (prob[i] < 0 || prob[i] >= e->n_problems || pole[i] < 0 || pole[i] >= e->prob_npoles[prob[i]]) {
      e->err = "problem or pole index out of range";
      return PS_ERR_ARG;
    }
  if ((rc = ensure_items(e, n))) return rc;
  if ((size_t)n > e->iprob_cap) {
    if (e->d_iprob) (void)hipFree(e->d_iprob);
    e->d_iprob = nullptr;
    e->iprob_cap = 0;
    HIP_TRY(hipMalloc(&e->d_iprob, (size_t)n * sizeof(int)));
    e->iprob_cap = n;
  }
  if ((size_t)e->n_problems > e->conw_cap) {
    if (e->d_conw) (void)hipFree(e->d_conw);
    e->d_conw = nullptr;
    e->conw_cap = 0;
    HIP_TRY(hipMalloc(&e->d_conw, (size_t)e->n_problems * sizeof(double)));
    e->conw_cap = e->n_problems;
  }
  // with size classes the items run grouped by class (each class its own
  // lane); the results go back to the caller's order
  std::vector<int> order;
  std::vector<double> freq_s;
  std::vector<int> pole_s, prob_s;
  e->cls_start.clear();
  if (!e->cls.empty()) {
    const int nc = (int)e->cls.size();
    e->cls_start.assign(nc + 1, 0);
    for (int i = 0; i < n; i++) e->cls_start[e->prob_cls[prob[i]] + 1]++;
    for (int c = 0; c < nc; c++) e->cls_start[c + 1] += e->cls_start[c];
    std::vector<int> fill(e->cls_start.begin(), e->cls_start.end() - 1);
    order.resize(n);
    for (int i = 0; i < n; i++) order[fill[e->prob_cls[prob[i]]]++] = i;
    freq_s.resize(n);
    pole_s.resize(n);
    prob_s.resize(n);
    for (int k = 0; k < n; k++) {
      freq_s[k] = freq[order[k]];
      pole_s[k] = pole[order[k]];
      prob_s[k] = prob[order[k]];
    }
    freq = freq_s.data();
    pole = pole_s.data();
    prob = prob_s.data();
  }
  HIP_TRY(hipMemcpyAsync(e->d_freq, freq, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipMemcpyAsync(e->d_pole, pole, n * sizeof(int), hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipMemcpyAsync(e->d_iprob, prob, n * sizeof(int), hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipMemcpyAsync(e->d_conw, conw_r, e->n_problems * sizeof(double), hipMemcpyHostToDevice, e->stream));
  DevWork W{};
  W.n_items = n;
  W.mode = 0;
  W.freq = e->d_freq;
  W.pole = e->d_pole;
  W.item_prob = e->d_iprob;
  W.conw_rs = e->d_conw;
  W.out = e->d_out;
  if ((rc = run_batch(e, W))) return rc;
  if (order.empty()) {
    HIP_TRY(hipMemcpyAsync(out, e->d_out, n * sizeof(ps_fit_out), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
  } else {
    std::vector<ps_fit_out> tmp(n);
    HIP_TRY(hipMemcpyAsync(tmp.data(), e->d_out, n * sizeof(ps_fit_out), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (int k = 0; k < n; k++) out[order[k]] = tmp[k];
  }
  return PS_OK;
}

int ps_lc_features(ps_engine *e, const double *freq, int n, int nbins, float *out) {
  int rc = check_ready(e);
  if (rc) return rc;
  if (n < 0 || nbins < 1 || nbins > 256 || (n > 0 && (!freq || !out))) return PS_ERR_ARG;
  if (n == 0) return PS_OK;
  if ((rc = ensure_items(e, n))) return rc;
  const size_t need = (size_t)n * (nbins + 2);
  if (need > e->feat_cap) {
    if (e->d_feat) (void)hipFree(e->d_feat);
    e->d_feat = nullptr;
    e->feat_cap = 0;
    HIP_TRY(hipMalloc(&e->d_feat, need * sizeof(float)));
    e->feat_cap = need;
  }
  HIP_TRY(hipMemcpyAsync(e->d_freq, freq, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
  hipLaunchKernelGGL(lc_features, dim3(n), dim3(kThreads), 0, e->stream, e->P, e->d_freq, nbins, e->feat_alpha,
                     e->feat_aspect, e->d_feat);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, e->d_feat, need * sizeof(float), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return PS_OK;
}

int ps_fit_batch_from(ps_engine *e, const double *cg, int n, double conw_r, ps_fit_out *out, double *cg_out) {
  int rc = check_ready(e);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!cg || !out))) return PS_ERR_ARG;
  if (n == 0) return PS_OK;
  if ((rc = ensure_items(e, n))) return rc;
  if ((rc = ensure_cg(e, n))) return rc;
  if (cg_out && (size_t)n > e->cgo_cap) {
    if (e->d_cgo) (void)hipFree(e->d_cgo);
    e->d_cgo = nullptr;
    e->cgo_cap = 0;
    HIP_TRY(hipMalloc(&e->d_cgo, (size_t)n * kMaxMa * sizeof(double)));
    e->cgo_cap = n;
  }
  HIP_TRY(hipMemcpyAsync(e->d_cg, cg, (size_t)n * e->P.ma * sizeof(double), hipMemcpyHostToDevice, e->stream));
  DevWork W{};
  W.n_items = n;
  W.mode = 2;
  W.cg_in = e->d_cg;
  W.conw_r = conw_r;
  W.out = e->d_out;
  W.cg_out = cg_out ? e->d_cgo : nullptr;
  if ((rc = run_batch(e, W))) return rc;
  HIP_TRY(hipMemcpyAsync(out, e->d_out, n * sizeof(ps_fit_out), hipMemcpyDeviceToHost, e->stream));
  if (cg_out)
    HIP_TRY(hipMemcpyAsync(cg_out, e->d_cgo, (size_t)n * e->P.ma * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return PS_OK;
}

int ps_eval_mrqcof(ps_engine *e, const double *cg, int n, double conw_r, double *alpha, double *beta, double *chisq,
                   double *chck) {
  int rc = check_ready(e);
  if (rc) return rc;
  if (n <= 0 || !cg || !alpha || !beta || !chisq || !chck) return PS_ERR_ARG;
  const int ma = e->P.ma, mfit = e->P.mfit;
  if ((rc = ensure_cg(e, n))) return rc;
  if ((size_t)n > e->ev_cap) {
    void *bufs[] = {e->d_ea, e->d_eb, e->d_ec, e->d_ek};
    for (void *b : bufs)
      if (b) (void)hipFree(b);
    e->d_ea = e->d_eb = e->d_ec = e->d_ek = nullptr;
    e->ev_cap = 0;
    HIP_TRY(hipMalloc(&e->d_ea, (size_t)n * kMaxMfit * kMaxMfit * sizeof(double)));
    HIP_TRY(hipMalloc(&e->d_eb, (size_t)n * kMaxMfit * sizeof(double)));
    HIP_TRY(hipMalloc(&e->d_ec, (size_t)n * sizeof(double)));
    HIP_TRY(hipMalloc(&e->d_ek, (size_t)n * 3 * sizeof(double)));
    e->ev_cap = n;
  }
  HIP_TRY(hipMemcpyAsync(e->d_cg, cg, (size_t)n * ma * sizeof(double), hipMemcpyHostToDevice, e->stream));
  DevWork W{};
  W.n_items = n;
  W.mode = 1;
  W.cg_in = e->d_cg;
  W.conw_r = conw_r;
  W.ev_alpha = e->d_ea;
  W.ev_beta = e->d_eb;
  W.ev_chi = e->d_ec;
  W.ev_chck = e->d_ek;
  if ((rc = run_batch(e, W))) return rc;
  HIP_TRY(hipMemcpyAsync(alpha, e->d_ea, (size_t)n * mfit * mfit * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipMemcpyAsync(beta, e->d_eb, (size_t)n * mfit * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipMemcpyAsync(chisq, e->d_ec, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipMemcpyAsync(chck, e->d_ek, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return PS_OK;
}

}  // extern "C"
