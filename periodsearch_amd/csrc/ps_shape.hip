// ps_shape.hip -- shape setup for many asteroids on the GPU.
//
// Per asteroid the reference builds, once per run, the facet normals and
// areas of the triangulated Gaussian sphere (period_search/areanorm.c:17-55),
// the Laplace-series basis at the facet normals (period_search/sphfunc.c:
// 17-86) and the ellipsoid's starting coefficients, a least-squares fit of
// its log-curvature (period_search/ellfit.c:24-82, with the Crout LU of
// ludcmp.c:9-69 and lubksb.c:3-26).  For a batch of asteroids this file does
// that in six launches:
//
//   shape_facets    one thread per (shape, facet): vertex directions
//                   (period_search_BOINC.cpp:507-542), normal, area,
//                   Legendre functions, Dsph
//   shape_normal    one thread per (shape, i, j): fm^T fm of the ellipsoid fit
//   shape_logcurv   one thread per (asteroid, facet): the ellipsoid's
//                   log-curvature at the facet normal
//   shape_rhs       one thread per (asteroid, coefficient): fm^T er
//   shape_lu        one wave per shape: Crout LU with implicit partial
//                   pivoting, in LDS
//   shape_subst     one wave per asteroid: forward and back substitution
//
// Everything but the log-curvature depends only on (nrows, lmax, mmax), so
// the facet tables, the normal matrix and its LU are built once per distinct
// shape.  The fit matrix fm of ellfit equals Dsph entry for entry (both are
// P_l^m times cos(m af) or sin(m af), in the same column order), so Dsph is
// used for it.  The integer triangulation (trifac.c) is made on the host.
//
// The arithmetic keeps the host restatement's operation order (ps_host.cpp)
// with FMA contraction off; the tables differ from it only where the device
// libm's sin/cos/acos/atan2/log round differently from the host's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "periodsearch.h"
#include "ps_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kMaxLm = 10;     // period_search/constants.h:8
constexpr int kMaxFac = 1000;  // period_search/constants.h:5
constexpr int kMaxL = kMaxLm + 1;

// One distinct (nrows, lmax, mmax): its device tables.
struct ShapeDev {
  int nrows, lmax, mmax, nfac, ncoef;
  const int *ifp;  // [F][3] 1-based vertex ids
  double *at, *af;  // [F] facet normal directions
  double *nor, *darea, *pleg, *fc, *fs, *dsph;  // as ps_shape_basis: [F][3], [F], [F][L][L], [F][M] x2, [F][ncoef]
  double *lu;  // [ncoef][ncoef] ellipsoid normal matrix, then its LU factors
  int *indx;   // [ncoef] pivot rows
};

// One asteroid: its shape and ellipsoid.
struct JobDev {
  int shape;
  double a, b, c;
  double *er;  // [F] log-curvature
  double *cg;  // [ncoef] right-hand side, then the coefficients
};

// Direction of vertex v (0-based) of the Gaussian-sphere grid with nrows
// rows per hemisphere, as the host's vertex list: the north pole, rows
// i = 1..nrows of 4i vertices at theta = i dth, rows i = nrows-1..1 at
// theta = pi - i dth, the south pole (period_search_BOINC.cpp:507-542).
__device__ void vertex_xyz(int nrows, int v, double &x, double &y, double &z) {
  const double dth = kPi / (2 * nrows);
  const int north = 1 + 2 * nrows * (nrows + 1);  // first vertex of the southern rows
  double th = 0, ph = 0;
  if (v == 0) {
    th = 0;
    ph = 0;
  } else if (v < north) {
    int i = 1;
    while (v > 2 * i * (i + 1)) i++;  // row i holds vertices 1 + 2i(i-1) .. 2i(i+1)
    const int j = v - (1 + 2 * i * (i - 1));
    const double dph = kPi / (2 * i);
    th = i * dth;
    ph = j * dph;
  } else {
    int u = v - north, i = nrows - 1;
    while (i >= 1 && u >= 4 * i) {
      u -= 4 * i;
      i--;
    }
    if (i >= 1) {
      const double dph = kPi / (2 * i);
      th = kPi - i * dth;
      ph = u * dph;
    } else {
      th = kPi;
      ph = 0;
    }
  }
  const double st = sin(th);
  x = st * cos(ph);
  y = st * sin(ph);
  z = cos(th);
}

// Facet tables of one shape (blockIdx.y), one thread per facet.
__global__ void __launch_bounds__(64) shape_facets(const ShapeDev *__restrict__ shapes) {
  const ShapeDev S = shapes[blockIdx.y];
  const int L = S.lmax + 1, M = S.mmax + 1;
  // coefficients of P_l^m as a polynomial in cos(theta) times sin^m(theta)
  // (sphfunc.c:37-62), built by one thread: aleg[(n L + l) L + m]
  __shared__ double aleg[kMaxL * kMaxL * kMaxL];
  for (int q = threadIdx.x; q < L * L * L; q += blockDim.x) aleg[q] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
#define AL(n, l, m) aleg[((n) * L + (l)) * L + (m)]
    AL(0, 0, 0) = 1;
    if (S.lmax >= 1) AL(1, 1, 0) = 1;
    for (int l = 1; l <= S.lmax; l++) AL(0, l, l) = AL(0, l - 1, l - 1) * (2 * l - 1);
    for (int m = 0; m <= S.mmax; m++)
      for (int l = m + 1; l <= S.lmax; l++) {
        int ibot;
        if ((l - m) % 2 == 0) {
          AL(0, l, m) = -(l + m - 1) * AL(0, l - 2, m) / (1 * (l - m));
          ibot = 2;
        } else {
          ibot = 1;
        }
        if (l != 1)
          for (int n = ibot; n <= l - m; n += 2)
            AL(n, l, m) = ((2 * l - 1) * AL(n - 1, l - 1, m) - (l + m - 1) * AL(n, l - 2, m)) / (1 * (l - m));
      }
  }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.nfac) return;
  // normal and area (areanorm.c:17-55)
  double x[3], y[3], z[3];
  for (int k = 0; k < 3; k++) vertex_xyz(S.nrows, S.ifp[3 * i + k] - 1, x[k], y[k], z[k]);
  const double ax = x[1] - x[0], ay = y[1] - y[0], az = z[1] - z[0];
  const double bx = x[2] - x[0], by = y[2] - y[0], bz = z[2] - z[0];
  const double c1 = ay * bz - by * az;
  const double c2 = az * bx - bz * ax;
  const double c3 = ax * by - bx * ay;
  const double len = sqrt(c1 * c1 + c2 * c2 + c3 * c3);
  const double n0 = c1 / len, n1 = c2 / len, n2 = c3 / len;
  S.nor[3 * i] = n0;
  S.nor[3 * i + 1] = n1;
  S.nor[3 * i + 2] = n2;
  const double at = acos(n2), af = atan2(n1, n0);
  S.at[i] = at;
  S.af[i] = af;
  S.darea[i] = 0.5 * len;
  // Legendre functions and the basis at the normal (sphfunc.c:17-35, 64-86)
  double ts[kMaxL], tc[kMaxL];
  ts[0] = 1;
  tc[0] = 1;
  if (S.lmax >= 1) {
    ts[1] = sin(at);
    tc[1] = cos(at);
  }
  for (int j = 2; j <= S.lmax; j++) {
    ts[j] = ts[1] * ts[j - 1];
    tc[j] = tc[1] * tc[j - 1];
  }
  double *fs = S.fs + (size_t)i * M, *fc = S.fc + (size_t)i * M;
  fs[0] = 0;
  fc[0] = 1;
  for (int j = 1; j <= S.mmax; j++) {
    fs[j] = sin(j * af);
    fc[j] = cos(j * af);
  }
  double *pl = S.pleg + (size_t)i * L * L;
  for (int q = 0; q < L * L; q++) pl[q] = 0;
  double *ds = S.dsph + (size_t)i * S.ncoef;
  int k = 0;
  for (int m = 0; m <= S.mmax; m++)
    for (int l = m; l <= S.lmax; l++) {
      double p = 0;
      for (int n = (l - m) % 2; n <= l - m; n += 2) p = p + AL(n, l, m) * tc[n] * ts[m];
      pl[l * L + m] = p;
      ds[k++] = fc[m] * p;
      if (m != 0) ds[k++] = fs[m] * p;
    }
#undef AL
}

// Facets per LDS chunk of the fit matrix in shape_normal and shape_rhs.
constexpr int kFmChunk = 32;
constexpr int kMaxNc = 128;  // >= 121 coefficients of lmax = mmax = 10

// Stage fit-matrix rows k0 .. k0+cnt-1 (contiguous in fm) into LDS.
__device__ inline void stage_rows(double *ch, const double *fm, int k0, int cnt, int nc) {
  for (int q = threadIdx.x; q < cnt * nc; q += blockDim.x) ch[q] = fm[(size_t)k0 * nc + q];
}

// Normal matrix of the ellipsoid fit, fm^T fm summed over facets in order
// (ellfit.c:57-67); one thread per entry, shape blockIdx.y.  The rows of fm
// pass through LDS in chunks of 32 facets.
__global__ void __launch_bounds__(256) shape_normal(const ShapeDev *__restrict__ shapes) {
  const ShapeDev S = shapes[blockIdx.y];
  const int nc = S.ncoef;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x * blockDim.x >= nc * nc) return;  // whole block idle (uniform)
  const bool active = t < nc * nc;
  const int i = active ? t / nc : 0, j = active ? t - (t / nc) * nc : 0;
  __shared__ double ch[kFmChunk * kMaxNc];
  double s = 0;
  for (int k0 = 0; k0 < S.nfac; k0 += kFmChunk) {
    const int cnt = min(kFmChunk, S.nfac - k0);
    __syncthreads();
    stage_rows(ch, S.dsph, k0, cnt, nc);
    __syncthreads();
    if (active)
      for (int k = 0; k < cnt; k++) s = s + ch[k * nc + i] * ch[k * nc + j];
  }
  if (active) S.lu[t] = s;
}

// log-curvature of asteroid blockIdx.y's ellipsoid at each facet normal
// (ellfit.c:29-36).
__global__ void __launch_bounds__(64) shape_logcurv(const ShapeDev *__restrict__ shapes,
                                                    const JobDev *__restrict__ jobs) {
  const JobDev J = jobs[blockIdx.y];
  const ShapeDev &S = shapes[J.shape];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.nfac) return;
  double st, ct, sa, ca;
  sincos(S.at[i], &st, &ct);
  sincos(S.af[i], &sa, &ca);
  const double u = J.a * st * ca, v = J.b * st * sa, w = J.c * ct;
  const double sum = u * u + v * v + w * w;
  J.er[i] = 2 * (log(J.a * J.b * J.c) - log(sum));
}

// Right-hand side fm^T er of asteroid blockIdx.y (ellfit.c:68-71), one
// thread per coefficient; fm rows and er pass through LDS.
__global__ void __launch_bounds__(kMaxNc) shape_rhs(const ShapeDev *__restrict__ shapes,
                                                    const JobDev *__restrict__ jobs) {
  const JobDev J = jobs[blockIdx.x];
  const ShapeDev &S = shapes[J.shape];
  const int nc = S.ncoef, i = threadIdx.x;
  __shared__ double ch[kFmChunk * kMaxNc];
  __shared__ double ec[kFmChunk];
  double s = 0;
  for (int k0 = 0; k0 < S.nfac; k0 += kFmChunk) {
    const int cnt = min(kFmChunk, S.nfac - k0);
    __syncthreads();
    stage_rows(ch, S.dsph, k0, cnt, nc);
    if (i < cnt) ec[i] = J.er[k0 + i];
    __syncthreads();
    if (i < nc)
      for (int k = 0; k < cnt; k++) s = s + ch[k * nc + i] * ec[k];
  }
  if (i < nc) J.cg[i] = s;
}

// Broadcast lane `lane` (wave-uniform) of a double through scalar registers.
__device__ inline double lane_value(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Crout LU with implicit partial pivoting (ludcmp.c:9-69) of one shape's
// normal matrix, one wave.  Lane r holds rows r and r + 64 of the current
// column.  Column j's entries a[i][j] = a[i][j] - sum_{k < min(i, j)}
// a[i][k] a[k][j] are accumulated right-looking: at step k the final
// a[k][j] is broadcast and every row below k subtracts its term, so each
// entry's sum runs over k in increasing order, as in the sequential loops.
// The pivot is the last row of largest v[i] |a[i][j]| (the reference's
// `>=`).  LDS: the matrix and the row scales.
__global__ void __launch_bounds__(64) shape_lu(const ShapeDev *__restrict__ shapes) {
  const ShapeDev S = shapes[blockIdx.x];
  const int n = S.ncoef, lane = threadIdx.x;
  extern __shared__ double lds[];
  double *A = lds, *vv = lds + n * n;
  for (int q = lane; q < n * n; q += 64) A[q] = S.lu[q];
  __syncthreads();
  for (int r = lane; r < n; r += 64) {
    double big = 0.0;
    for (int j = 0; j < n; j++) {
      const double t = fabs(A[r * n + j]);
      if (t > big) big = t;
    }
    vv[r] = 1.0 / big;
  }
  __syncthreads();
  const int r0 = lane, r1 = lane + 64;
  for (int j = 0; j < n; j++) {
    double acc0 = r0 < n ? A[r0 * n + j] : 0.0;
    double acc1 = r1 < n ? A[r1 * n + j] : 0.0;
    int k = 0;
    for (; k + 4 <= j; k += 4) {  // the L entries of four steps loaded ahead of the dependent chain
      double l0[4], l1[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        l0[q] = r0 < n ? A[r0 * n + k + q] : 0.0;
        l1[q] = r1 < n ? A[r1 * n + k + q] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int kk = k + q;
        const double u = kk < 64 ? lane_value(acc0, kk) : lane_value(acc1, kk - 64);
        if (r0 > kk && r0 < n) acc0 = acc0 - l0[q] * u;
        if (r1 > kk && r1 < n) acc1 = acc1 - l1[q] * u;
      }
    }
    for (; k < j; k++) {
      const double u = k < 64 ? lane_value(acc0, k) : lane_value(acc1, k - 64);
      if (r0 > k && r0 < n) acc0 = acc0 - A[r0 * n + k] * u;
      if (r1 > k && r1 < n) acc1 = acc1 - A[r1 * n + k] * u;
    }
    if (r0 < n) A[r0 * n + j] = acc0;
    if (r1 < n) A[r1 * n + j] = acc1;
    // pivot: largest scaled entry among rows >= j, ties to the last row
    double best = -1.0;
    int imax = -1;
    for (int r = (r0 >= j ? r0 : (r1 >= j ? r1 : n)); r < n; r += 64) {
      const double dum = vv[r] * fabs(r == r0 ? acc0 : acc1);
      if (dum >= best) {  // NaN never wins
        best = dum;
        imax = r;
      }
    }
    for (int off = 32; off >= 1; off >>= 1) {
      const double ob = __shfl_xor(best, off);
      const int oi = __shfl_xor(imax, off);
      if (ob > best || (ob == best && oi > imax)) {
        best = ob;
        imax = oi;
      }
    }
    if (imax < 0) imax = j;  // every candidate NaN: keep the row
    __syncthreads();
    if (imax != j) {
      for (int c = lane; c < n; c += 64) {
        const double t = A[imax * n + c];
        A[imax * n + c] = A[j * n + c];
        A[j * n + c] = t;
      }
      if (lane == 0) vv[imax] = vv[j];
    }
    if (lane == 0) S.indx[j] = imax;
    __syncthreads();
    if (lane == 0 && A[j * n + j] == 0.0) A[j * n + j] = 1.0e-20;
    __syncthreads();
    if (j != n - 1) {
      const double dum = 1.0 / A[j * n + j];
      for (int i = j + 1 + lane; i < n; i += 64) A[i * n + j] *= dum;
    }
    __syncthreads();
  }
  for (int q = lane; q < n * n; q += 64) S.lu[q] = A[q];
}

// Forward and back substitution of each asteroid's right-hand side with its
// shape's LU factors (lubksb.c:3-26), one wave per asteroid: the wave stages
// the factors in LDS, lane 0 runs the substitution (its sums are sequential
// by definition).
__global__ void __launch_bounds__(64) shape_subst(const ShapeDev *__restrict__ shapes,
                                                  const JobDev *__restrict__ jobs) {
  const JobDev J = jobs[blockIdx.x];
  const ShapeDev &S = shapes[J.shape];
  const int n = S.ncoef;
  extern __shared__ double lds[];
  double *a = lds, *b = lds + n * n;
  for (int q = threadIdx.x; q < n * n; q += 64) a[q] = S.lu[q];
  for (int q = threadIdx.x; q < n; q += 64) b[q] = J.cg[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    int ii = -1;
    for (int i = 0; i < n; i++) {
      const int ip = S.indx[i];
      double sum = b[ip];
      b[ip] = b[i];
      if (ii >= 0)
        for (int j = ii; j <= i - 1; j++) sum -= a[i * n + j] * b[j];
      else if (sum)
        ii = i;
      b[i] = sum;
    }
    for (int i = n - 1; i >= 0; i--) {
      double sum = b[i];
      for (int j = i + 1; j < n; j++) sum -= a[i * n + j] * b[j];
      b[i] = sum / a[i * n + i];
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += 64) J.cg[q] = b[q];
}

int count_coefs(int lmax, int mmax) {
  int n = 0;
  for (int m = 0; m <= mmax; m++)
    for (int l = m; l <= lmax; l++) n += (m != 0) ? 2 : 1;
  return n;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

extern "C" int ps_shape_basis_many(ps_engine *e, int n, const ps_shape_job *jobs) {
  if (!e) return PS_ERR_ARG;
  if (n == 0) return PS_OK;
  if (n < 0 || !jobs) {
    psamd_engine_error(e, "ps_shape_basis_many: bad job list");
    return PS_ERR_ARG;
  }
  // distinct shapes and their table sizes (doubles)
  std::map<std::tuple<int, int, int>, int> key;
  std::vector<ShapeDev> sh;
  std::vector<int> job_shape(n);
  for (int j = 0; j < n; j++) {
    const ps_shape_job &q = jobs[j];
    if (q.nrows < 1 || q.lmax < 0 || q.mmax < 0 || q.mmax > q.lmax || q.lmax > kMaxLm || !q.nor || !q.darea ||
        !q.pleg || !q.fc || !q.fs || !q.dsph || !q.cg_first) {
      psamd_engine_error(e, "ps_shape_basis_many: job " + std::to_string(j) + " is invalid");
      return PS_ERR_ARG;
    }
    if (8 * q.nrows * q.nrows > kMaxFac) {
      psamd_engine_error(e, "ps_shape_basis_many: job " + std::to_string(j) + " has more than 1000 facets");
      return PS_ERR_LIMIT;
    }
    auto k = std::make_tuple(q.nrows, q.lmax, q.mmax);
    auto it = key.find(k);
    if (it == key.end()) {
      ShapeDev s{};
      s.nrows = q.nrows;
      s.lmax = q.lmax;
      s.mmax = q.mmax;
      s.nfac = 8 * q.nrows * q.nrows;
      s.ncoef = count_coefs(q.lmax, q.mmax);
      it = key.emplace(k, (int)sh.size()).first;
      sh.push_back(s);
    }
    job_shape[j] = it->second;
  }
  const int ns = (int)sh.size();
  // arena layout: per shape the output tables (contiguous, copied back in
  // one piece), the work arrays and the triangulation; per job er and cg;
  // the descriptors last
  struct Off {
    size_t out, out_bytes, at, af, lu, indx, ifp;
  };
  std::vector<Off> so(ns);
  size_t off = 0;
  for (int s = 0; s < ns; s++) {
    const ShapeDev &S = sh[s];
    const size_t F = S.nfac, L = S.lmax + 1, M = S.mmax + 1, nc = S.ncoef;
    so[s].out = off;
    so[s].out_bytes = (F * 3 + F + F * L * L + 2 * F * M + F * nc) * sizeof(double);
    off = align256(off + so[s].out_bytes);
    so[s].at = off;
    off = align256(off + F * sizeof(double));
    so[s].af = off;
    off = align256(off + F * sizeof(double));
    so[s].lu = off;
    off = align256(off + nc * nc * sizeof(double));
    so[s].indx = off;
    off = align256(off + nc * sizeof(int));
    so[s].ifp = off;
    off = align256(off + F * 3 * sizeof(int));
  }
  std::vector<size_t> jer(n), jcg(n);
  for (int j = 0; j < n; j++) {
    jer[j] = off;
    off = align256(off + sh[job_shape[j]].nfac * sizeof(double));
  }
  const size_t cg_off = off;  // every job's coefficients, packed: one download
  for (int j = 0; j < n; j++) {
    jcg[j] = off;
    off += sh[job_shape[j]].ncoef * sizeof(double);
  }
  const size_t cg_bytes = off - cg_off;
  off = align256(off);
  const size_t shapes_off = off;
  off = align256(off + ns * sizeof(ShapeDev));
  const size_t jobs_off = off;
  off = align256(off + n * sizeof(JobDev));
  const size_t total = off;

  hipStream_t stream = static_cast<hipStream_t>(psamd_engine_stream(e));
  const int dev = psamd_engine_device(e);
  char *base = nullptr;
  auto fail = [&](const char *what, hipError_t err) {
    psamd_engine_error(e, std::string("ps_shape_basis_many: ") + what + ": " + hipGetErrorString(err));
    if (base) (void)hipFree(base);
    return PS_ERR_HIP;
  };
  hipError_t err = hipSetDevice(dev);
  if (err != hipSuccess) return fail("hipSetDevice", err);
  err = hipMalloc(&base, total);
  if (err != hipSuccess) return fail("hipMalloc", err);

  // host image of the inputs: triangulations and descriptors
  std::vector<char> img(total, 0);
  for (int s = 0; s < ns; s++) {
    ShapeDev &S = sh[s];
    const size_t F = S.nfac, L = S.lmax + 1, M = S.mmax + 1;
    psamd_triangulate(S.nrows, reinterpret_cast<int *>(img.data() + so[s].ifp));
    double *o = reinterpret_cast<double *>(base + so[s].out);
    S.nor = o;
    S.darea = S.nor + F * 3;
    S.pleg = S.darea + F;
    S.fc = S.pleg + F * L * L;
    S.fs = S.fc + F * M;
    S.dsph = S.fs + F * M;
    S.at = reinterpret_cast<double *>(base + so[s].at);
    S.af = reinterpret_cast<double *>(base + so[s].af);
    S.lu = reinterpret_cast<double *>(base + so[s].lu);
    S.indx = reinterpret_cast<int *>(base + so[s].indx);
    S.ifp = reinterpret_cast<const int *>(base + so[s].ifp);
  }
  std::vector<JobDev> jd(n);
  for (int j = 0; j < n; j++) {
    jd[j].shape = job_shape[j];
    jd[j].a = jobs[j].a;
    jd[j].b = jobs[j].b;
    jd[j].c = jobs[j].c;
    jd[j].er = reinterpret_cast<double *>(base + jer[j]);
    jd[j].cg = reinterpret_cast<double *>(base + jcg[j]);
  }
  std::memcpy(img.data() + shapes_off, sh.data(), ns * sizeof(ShapeDev));
  std::memcpy(img.data() + jobs_off, jd.data(), n * sizeof(JobDev));
  // only the triangulations and the descriptors need to travel
  for (int s = 0; s < ns; s++) {
    err = hipMemcpyAsync(base + so[s].ifp, img.data() + so[s].ifp, sh[s].nfac * 3 * sizeof(int),
                         hipMemcpyHostToDevice, stream);
    if (err != hipSuccess) return fail("upload", err);
  }
  err = hipMemcpyAsync(base + shapes_off, img.data() + shapes_off, total - shapes_off, hipMemcpyHostToDevice,
                       stream);
  if (err != hipSuccess) return fail("upload", err);
  const ShapeDev *d_sh = reinterpret_cast<const ShapeDev *>(base + shapes_off);
  const JobDev *d_jobs = reinterpret_cast<const JobDev *>(base + jobs_off);

  int fmax = 0, ncmax = 0;
  for (const ShapeDev &S : sh) {
    fmax = std::max(fmax, S.nfac);
    ncmax = std::max(ncmax, S.ncoef);
  }
  const size_t lu_lds = (size_t)ncmax * (ncmax + 1) * sizeof(double);
  err = hipFuncSetAttribute(reinterpret_cast<const void *>(shape_lu), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lu_lds);
  if (err != hipSuccess) return fail("shape_lu LDS", err);
  shape_facets<<<dim3((fmax + 63) / 64, ns), 64, 0, stream>>>(d_sh);
  err = hipFuncSetAttribute(reinterpret_cast<const void *>(shape_subst), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lu_lds);
  if (err != hipSuccess) return fail("shape_subst LDS", err);
  shape_normal<<<dim3((ncmax * ncmax + 255) / 256, ns), 256, 0, stream>>>(d_sh);
  shape_logcurv<<<dim3((fmax + 63) / 64, n), 64, 0, stream>>>(d_sh, d_jobs);
  shape_rhs<<<n, kMaxNc, 0, stream>>>(d_sh, d_jobs);
  shape_lu<<<ns, 64, lu_lds, stream>>>(d_sh);
  shape_subst<<<n, 64, lu_lds, stream>>>(d_sh, d_jobs);
  err = hipGetLastError();
  if (err != hipSuccess) return fail("launch", err);

  // results back: each shape's tables once, each job's coefficients
  std::vector<std::vector<double>> tabs(ns);
  for (int s = 0; s < ns; s++) {
    tabs[s].resize(so[s].out_bytes / sizeof(double));
    err = hipMemcpyAsync(tabs[s].data(), base + so[s].out, so[s].out_bytes, hipMemcpyDeviceToHost, stream);
    if (err != hipSuccess) return fail("download", err);
  }
  std::vector<double> cgs(cg_bytes / sizeof(double));
  err = hipMemcpyAsync(cgs.data(), base + cg_off, cg_bytes, hipMemcpyDeviceToHost, stream);
  if (err != hipSuccess) return fail("download", err);
  err = hipStreamSynchronize(stream);
  if (err != hipSuccess) return fail("shape setup", err);
  (void)hipFree(base);
  for (int j = 0; j < n; j++) {
    const ShapeDev &S = sh[job_shape[j]];
    const size_t F = S.nfac, L = S.lmax + 1, M = S.mmax + 1, nc = S.ncoef;
    const double *t = tabs[job_shape[j]].data();
    const ps_shape_job &q = jobs[j];
    std::memcpy(q.nor, t, F * 3 * sizeof(double));
    t += F * 3;
    std::memcpy(q.darea, t, F * sizeof(double));
    t += F;
    std::memcpy(q.pleg, t, F * L * L * sizeof(double));
    t += F * L * L;
    std::memcpy(q.fc, t, F * M * sizeof(double));
    t += F * M;
    std::memcpy(q.fs, t, F * M * sizeof(double));
    t += F * M;
    std::memcpy(q.dsph, t, F * nc * sizeof(double));
    std::memcpy(q.cg_first, cgs.data() + (jcg[j] - cg_off) / sizeof(double), nc * sizeof(double));
  }
  return PS_OK;
}
