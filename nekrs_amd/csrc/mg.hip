// mg.hip -- p-multigrid preconditioner: level construction, Chebyshev smoothing
// around Jacobi or overlapping Schwarz (fdm.hip), intergrid transfers, FP64
// coarse-grid PCG and the V-cycle.
//
// Reference behaviour followed (src/solvers/elliptic/MG/):
//   V-cycle                  MGSolver.cpp:167-193
//   smooth / residual        ellipticMultiGridLevel.cpp:37-41,85-277
//   coarsen / prolongate     ellipticMultiGridLevel.cpp:43-83
//   Chebyshev bounds         ellipticMultiGridLevelSetup.cpp:109-180 (lambda_max by a
//                            10-step Arnoldi on S A, :292-453)
//   smoother choice          ellipticMultiGridLevel.cpp:85-111,
//                            ellipticMultiGridLevelSetup.cpp:109-180
//   level schedule           determineMGLevels.cpp:58-95
// The coarse level is solved by Jacobi-PCG in FP64 on the GPU (the reference uses
// BoomerAMG on an assembled FEM matrix on the host).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <type_traits>

#include "amg.hpp"
#include "fdm.hpp"
#include "kernels.hpp"
#include "mesh.hpp"
#include "mg.hpp"
#include "tune.hpp"

namespace nrs {

// ------------------------------------------------------------------ level mesh helpers

std::vector<double> interp_coords(int Nf, int E, const double *xf, int Nc)
{
  std::vector<double> zf, wf, zc, wc;
  gll_nodes(Nf, zf, wf);
  gll_nodes(Nc, zc, wc);
  const auto I = interp_matrix(zf, zc);   // [NqC][NqF]
  const int NqF = Nf + 1, NqC = Nc + 1, NpF = NqF * NqF * NqF, NpC = NqC * NqC * NqC;
  std::vector<double> out((size_t)E * NpC), a(NqC * NqF * NqF), b(NqC * NqC * NqF);
  for (int64_t e = 0; e < E; ++e) {
    const double *u = xf + e * NpF;
    for (int k = 0; k < NqF; ++k)
      for (int j = 0; j < NqF; ++j)
        for (int ic = 0; ic < NqC; ++ic) {
          double s = 0;
          for (int i = 0; i < NqF; ++i) s += I[ic * NqF + i] * u[i + NqF * (j + NqF * k)];
          a[ic + NqC * (j + NqF * k)] = s;
        }
    for (int k = 0; k < NqF; ++k)
      for (int jc = 0; jc < NqC; ++jc)
        for (int ic = 0; ic < NqC; ++ic) {
          double s = 0;
          for (int j = 0; j < NqF; ++j) s += I[jc * NqF + j] * a[ic + NqC * (j + NqF * k)];
          b[ic + NqC * (jc + NqC * k)] = s;
        }
    for (int kc = 0; kc < NqC; ++kc)
      for (int jc = 0; jc < NqC; ++jc)
        for (int ic = 0; ic < NqC; ++ic) {
          double s = 0;
          for (int k = 0; k < NqF; ++k) s += I[kc * NqF + k] * b[ic + NqC * (jc + NqC * k)];
          out[e * NpC + ic + NqC * (jc + NqC * kc)] = s;
        }
  }
  return out;
}

std::vector<long long> coarse_global_ids(int Nf, int E, const long long *gidF, int Nc)
{
  const int NqF = Nf + 1, NqC = Nc + 1, NpF = NqF * NqF * NqF, NpC = NqC * NqC * NqC;
  using Key = std::array<long long, 5>;
  std::vector<Key> keys((size_t)E * NpC);
  auto fid = [&](int64_t e, int i, int j, int k) { return gidF[e * NpF + i + NqF * (j + NqF * k)]; };
  for (int64_t e = 0; e < E; ++e)
    for (int kc = 0; kc < NqC; ++kc)
      for (int jc = 0; jc < NqC; ++jc)
        for (int ic = 0; ic < NqC; ++ic) {
          const int c[3] = {ic, jc, kc};
          bool bnd[3];
          int nb = 0;
          for (int a = 0; a < 3; ++a) { bnd[a] = (c[a] == 0 || c[a] == Nc); nb += bnd[a]; }
          auto fineAt = [&](const int f[3]) { return fid(e, f[0], f[1], f[2]); };
          Key key{};
          if (nb == 3) {
            int f[3];
            for (int a = 0; a < 3; ++a) f[a] = c[a] ? Nf : 0;
            key = {0, fineAt(f), 0, 0, 0};
          } else if (nb == 2) {
            int ax = 0;
            for (int a = 0; a < 3; ++a) if (!bnd[a]) ax = a;
            int f0[3], f1[3];
            for (int a = 0; a < 3; ++a) { f0[a] = c[a] ? Nf : 0; f1[a] = f0[a]; }
            f0[ax] = 0; f1[ax] = Nf;
            const long long A = fineAt(f0), B = fineAt(f1);
            const int t = c[ax];
            key = (A < B) ? Key{1, A, B, t, 0} : Key{1, B, A, Nc - t, 0};
          } else if (nb == 1) {
            int fx = 0;
            for (int a = 0; a < 3; ++a) if (bnd[a]) fx = a;
            const int p = (fx == 0) ? 1 : 0, q = (fx == 2) ? 1 : 2;   // the two free axes
            // face corners in cyclic order (0,0),(1,0),(1,1),(0,1) of (p,q)
            const int cp[4] = {0, 1, 1, 0}, cq[4] = {0, 0, 1, 1};
            long long cid[4];
            for (int m = 0; m < 4; ++m) {
              int f[3];
              f[fx] = c[fx] ? Nf : 0;
              f[p] = cp[m] ? Nf : 0;
              f[q] = cq[m] ? Nf : 0;
              cid[m] = fineAt(f);
            }
            int m0 = 0;
            for (int m = 1; m < 4; ++m) if (cid[m] < cid[m0]) m0 = m;
            const int n1 = (m0 + 1) % 4, n3 = (m0 + 3) % 4;
            const int mu = (cid[n1] < cid[n3]) ? n1 : n3, mv = (mu == n1) ? n3 : n1;
            const int a = c[p], b = c[q];
            const int ox = cp[m0] * Nc, oy = cq[m0] * Nc;
            const int ux = (cp[mu] - cp[m0]), uy = (cq[mu] - cq[m0]);
            const int vx = (cp[mv] - cp[m0]), vy = (cq[mv] - cq[m0]);
            const int u = (a - ox) * ux + (b - oy) * uy;
            const int v = (a - ox) * vx + (b - oy) * vy;
            key = {2, cid[m0], cid[mu], cid[mv], (long long)u * (Nc + 1) + v};
          } else {
            key = {3, e, ic + NqC * (jc + NqC * kc), 0, 0};
          }
          keys[e * NpC + ic + NqC * (jc + NqC * kc)] = key;
        }
  std::vector<Key> uniq(keys);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  std::vector<long long> out(keys.size());
  for (size_t n = 0; n < keys.size(); ++n)
    out[n] = 1 + (long long)(std::lower_bound(uniq.begin(), uniq.end(), keys[n]) - uniq.begin());
  return out;
}

// ------------------------------------------------------------------ small kernels

namespace {

constexpr int KB = 256;
inline unsigned kgrid(int64_t n) { return grid_for(n, KB, 16384); }

// Dirichlet mask bit of entry i of a level's E-vector (mb: [E][ceil(np/32)]
// bits, np nodes per element; the vectors stay below 2^32 entries)
__device__ __forceinline__ bool masked_at(const uint32_t *__restrict__ mb, int64_t i, int np)
{
  const uint32_t iu = (uint32_t)i, e = iu / (uint32_t)np, nn = iu - e * (uint32_t)np;
  return (mb[(int64_t)e * ((np + 31) >> 5) + (nn >> 5)] >> (nn & 31)) & 1u;
}

// updateChebyshev (kernels/elliptic/updateChebyshev.okl:7-30).  JAC: the Jacobi
// smoother applied on the fly, SAd = w .* Ad with the product rounded on its own
// (no contraction into the subtraction), equal to a separate w .* Ad pass.
// LAST: the last step of the smoothing, fused with the x += d that follows it:
// x = (x + d) + d_new in that order, and d, r (no longer needed) not stored.
// X0: x is zero on entry and not read (the first step of a zero start).
// mb (LAST, optional): the level's mask bits, x stored as zero where set (the
// mask that follows the smoothing), np nodes per element.
template <typename T, bool JAC, bool LAST, bool X0 = false>
__global__ void cheby1_kernel(int64_t n, T dC, T rC, const T *__restrict__ w, const T *__restrict__ Ad,
                              T *__restrict__ d, T *__restrict__ r, T *__restrict__ x,
                              const uint32_t *__restrict__ mb = nullptr, int np = 0)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T sad;
    if constexpr (JAC) {
#pragma clang fp contract(off)
      sad = w[i] * Ad[i];
    } else {
      sad = Ad[i];
    }
    const T dn = d[i];
    const T rn = r[i] - sad;
    const T x1 = (X0 ? T(0) : x[i]) + dn;
    const T dnew = dC * dn + rC * rn;
    if constexpr (LAST) {
      x[i] = (mb && masked_at(mb, i, np)) ? T(0) : x1 + dnew;
    } else {
      x[i] = x1;
      d[i] = dnew;
      r[i] = rn;
    }
  }
}

// First step of the Jacobi-smoothed Chebyshev: r = w .* t, d = invTheta r with
// t = rhs - res (SUB, after the operator wrote A x into res) or t = rhs (x = 0);
// the same roundings as the separate passes.
template <typename T, bool SUB>
__global__ void jacobi_first_kernel(int64_t n, T invTheta, const T *__restrict__ w,
                                    const T *__restrict__ rhs, const T *__restrict__ res,
                                    T *__restrict__ r, T *__restrict__ d)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T t = SUB ? rhs[i] - res[i] : rhs[i];
    T ri;
    {
#pragma clang fp contract(off)
      ri = w[i] * t;
    }
    r[i] = ri;
    d[i] = invTheta * ri;
  }
}

// updateFourthKindChebyshev (kernels/elliptic/updateFourthKindChebyshev.okl:5-25)
// split in two: its residual update r -= A d runs before the smoother (which
// reads r), and its x += beta d is deferred into the direction update after
// the smoother, which reads d anyway.  X0: r = rin - A d with rin the
// right-hand side (a zero start, where r is not yet formed).
template <typename T, bool X0>
__global__ void cheby4_res_kernel(int64_t n, const T *__restrict__ Ad, const T *__restrict__ rin,
                                  T *__restrict__ r)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r[i] = (X0 ? rin[i] : r[i]) - Ad[i];
}

// After the smoother: x = x + beta d (the deferred half of the update; X0: x
// taken as zero, not read), then d = rC SAd + dC d -- or, on the last step
// (LAST), x = beta_last (rC SAd + dC d) + x with d not stored: the roundings
// of the separate update and axpby passes, in their order.
template <typename T, bool X0, bool LAST>
__global__ void cheby4_dir_kernel(int64_t n, T beta, T rC, T dC, T betaLast, const T *__restrict__ SAd,
                                  T *__restrict__ d, T *__restrict__ x, const uint32_t *__restrict__ mb = nullptr,
                                  int np = 0)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T dn = d[i];
    const T x1 = (X0 ? T(0) : x[i]) + beta * dn;
    const T dp = rC * SAd[i] + dC * dn;
    if constexpr (LAST) {
      x[i] = (mb && masked_at(mb, i, np)) ? T(0) : betaLast * dp + T(1) * x1;
    } else {
      x[i] = x1;
      d[i] = dp;
    }
  }
}

// y = a x (+ b y when b != 0)
template <typename T>
__global__ void axpby_kernel(int64_t n, T a, const T *__restrict__ x, T b, T *__restrict__ y)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (b == T(0)) ? a * x[i] : a * x[i] + b * y[i];
}

// y = w .* x (in place when y == x)
template <typename T>
__global__ void wmul_kernel(int64_t n, const T *__restrict__ w, const T *x, T *y)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = w[i] * x[i];
}

template <typename A, typename B>
__global__ void cast_kernel(int64_t n, const A *__restrict__ x, B *__restrict__ y)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (B)x[i];
}

template <typename T> void k_axpby(int64_t n, T a, const T *x, T b, T *y, hipStream_t s)
{
  axpby_kernel<T><<<kgrid(n), KB, 0, s>>>(n, a, x, b, y);
  NRS_LAUNCH_CHECK();
}
template <typename T> void k_wmul(int64_t n, const T *w, const T *x, T *y, hipStream_t s)
{
  wmul_kernel<T><<<kgrid(n), KB, 0, s>>>(n, w, x, y);
  NRS_LAUNCH_CHECK();
}
template <typename A, typename B> void k_cast(int64_t n, const A *x, B *y, hipStream_t s)
{
  if ((const void *)x == (const void *)y) return;
  cast_kernel<A, B><<<kgrid(n), KB, 0, s>>>(n, x, y);
  NRS_LAUNCH_CHECK();
}

// splitmix64: deterministic uniform [0,1) start vector for the Arnoldi process
// (the reference seeds std::default_random_engine from random_device,
// src/utils/randomVector.hpp:10-32, which is not reproducible).
struct SplitMix {
  uint64_t s;
  double next()
  {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
};

// spectral radius of the k x k Hessenberg matrix (column-major) by power
// iteration with the growth rate averaged over the last sweeps; stands in for
// max |eig(H)| from dgeev (ellipticMultiGridLevelSetup.cpp:435-447).
double spectral_radius(const std::vector<double> &H, int k)
{
  std::vector<double> v(k, 1.0), w(k);
  double logsum = 0.0;
  const int iters = 4000, tail = 1000;
  for (int it = 0; it < iters; ++it) {
    for (int r = 0; r < k; ++r) {
      double s = 0;
      for (int c = 0; c < k; ++c) s += H[r + c * k] * v[c];
      w[r] = s;
    }
    double nrm = 0;
    for (double a : w) nrm += a * a;
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) return 0.0;
    for (int r = 0; r < k; ++r) v[r] = w[r] / nrm;
    if (it >= iters - tail) logsum += std::log(nrm);
  }
  return std::exp(logsum / tail);
}

}  // namespace

// ------------------------------------------------------------------ levels

template <typename T>
struct Level {
  int N = 0, Nq = 0, Np = 0, E = 0;
  int64_t L = 0;
  bool isCoarse = false;
  DevBuf<T> ggeoOwn, DOwn, lam0, lam1;
  const T *ggeo = nullptr, *D = nullptr;
  nrs_ogs *ogsFull = nullptr, *ogs = nullptr;
  bool ownOgs = false;
  int64_t Nmasked = 0;
  DevBuf<int> maskIdsOwn;
  DevBuf<uint32_t> maskBitsOwn;
  const int *maskIds = nullptr;
  const uint32_t *maskBits = nullptr;
  DevBuf<T> invDegree;      // T copy of the masked ogs invDegree (o_invDegree of the level)
  DevBuf<T> invDiagA;
  DevBuf<T> R;              // [Nq][NqFiner] transfer from the finer level
  DevBuf<T> zw;             // GLL nodes and weights [2][Nq] (corner-geometry operator)
  const T *corners = nullptr;   // element corners [E][3][8] when the operator forms its geometry
  DevBuf<T> rhsOwn, xOwn, res;
  T *rhs = nullptr, *x = nullptr;
  double rho = 0, lmin = 0, lmax = 0;
  int degDown = 0, degUp = 0;        // before restriction (x = 0) / after prolongation
  std::vector<T> betasDown, betasUp;  // optimal 4th-kind Chebyshev weights of those degrees
  std::unique_ptr<Schwarz<T>> schwarz;   // ASM / RAS smoother (else Jacobi)
  ~Level()
  {
    if (ownOgs) { delete ogs; delete ogsFull; }
  }
};

// Jacobi-PCG in FP64 on the coarsest level.
struct CoarsePCG {
  int N = 1, E = 0;
  int64_t L = 0;
  DevBuf<double> ggeo, D, lam0, lam1, invDiag, r, x, z, p, Ap, scal;
  const nrs_ogs *ogs = nullptr;
  const uint32_t *maskBits = nullptr;
  bool poisson = true;
  double tol = 1e-2;
  int maxit = 200;
  int lastIters = 0;
  std::unique_ptr<Reducer> red;
  double *scalHost = nullptr;
  // A chunk of iterations is captured once into a HIP graph and replayed: the
  // chunk is ~70 small launches whose arguments never change (every scalar
  // lives in device memory), so replaying removes the per-launch overhead.
  bool useGraphs = true;
  hipStream_t capStream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  static constexpr int kChunk = 8;
  // AMG mode: PCG on the assembled matrix of the unmasked unique nodes
  // (rows numbered by first appearance), preconditioned by one
  // smoothed-aggregation V-cycle; E-vector <-> rows through rowOf / first.
  bool amgMode = false;
  bool deviceOnly = false;   // AMG with tol <= 0: exactly maxit iterations, no host round trip
  // all-Neumann problem: the coarse matrix is singular with the constant
  // vector as null space; the right-hand side is made consistent (its mean over
  // the unique nodes removed) and the AMG's dense level is pseudo-inverted
  // (the reference tells BoomerAMG the same, coarseLevel.cpp:113-146 ->
  // hypreWrapper.cpp:143-147)
  bool nullSpace = false;
  double uniqueNodes = 0.0;  // sum of invDegree = unique unmasked nodes (E-vector mode)
  int n = 0;
  Amg amg;
  DevBuf<int> rowOf, first, arp, aci;
  DevBuf<double> av, ones, ru, xu, zu, pu, Apu, tmp;
  static constexpr int kChunkAmg = 2;
  ~CoarsePCG()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capStream) (void)hipStreamDestroy(capStream);
    if (scalHost) (void)hipHostFree(scalHost);
  }

  void chunk(hipStream_t s)
  {
    if (amgMode) {
      for (int c = 0; c < kChunkAmg; ++c) {
        amg.apply(ru.get(), zu.get(), tmp.get(), s);
        pcg_precon_dots(*red, n, ones.get(), nullptr, ru.get(), zu.get(), Apu.get(), false, false, false,
                        scal.get(), s);
        pcg_update_p(n, zu.get(), pu.get(), scal.get(), s);
        csr_spmv(SPMV_Y, n, arp.get(), aci.get(), av.get(), pu.get(), Apu.get(), nullptr, nullptr, 0.0, s);
        pcg_alpha(*red, n, ones.get(), pu.get(), Apu.get(), scal.get(), s);
        pcg_update_rx(*red, n, ones.get(), Apu.get(), pu.get(), ru.get(), xu.get(), 1.0, scal.get(), s);
      }
      return;
    }
    const double *w = ogs->invDegree.get();
    for (int c = 0; c < kChunk; ++c) {
      pcg_precon_dots(*red, L, w, invDiag.get(), r.get(), z.get(), Ap.get(), true, false, false,
                      scal.get(), s);
      pcg_update_p(L, z.get(), p.get(), scal.get(), s);
      op(p.get(), Ap.get(), s);
      pcg_alpha(*red, L, w, p.get(), Ap.get(), scal.get(), s);
      pcg_update_rx(*red, L, w, Ap.get(), p.get(), r.get(), x.get(), 1.0, scal.get(), s);
    }
  }

  void run_chunk(hipStream_t s)
  {
    if (!useGraphs) { chunk(s); return; }
    if (!exec) {
      NRS_HIP(hipStreamCreateWithFlags(&capStream, hipStreamNonBlocking));
      NRS_HIP(hipStreamBeginCapture(capStream, hipStreamCaptureModeThreadLocal));
      chunk(capStream);
      NRS_HIP(hipStreamEndCapture(capStream, &graph));
      NRS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    NRS_HIP(hipGraphLaunch(exec, s));
  }

  void op(const double *q, double *Aq, hipStream_t s)
  {
    launch_ax<double>(N, E, nullptr, ggeo.get(), D.get(), lam0.get(), lam1.get(), poisson, false, q, Aq,
                      maskBits, s);
    ogs_gather_scatter<double>(ogs, Aq, 1, L, NRS_ADD, s);
  }

  // x = A^{-1} rhs approximately: PCG to relative tolerance `tol`, at most maxit.
  template <typename T>
  void solve(const T *rhs, T *xout, hipStream_t s)
  {
    if (amgMode) { solve_amg(rhs, xout, s); return; }
    const double *w = ogs->invDegree.get();
    k_cast<T, double>(L, rhs, r.get(), s);
    if (nullSpace) remove_mean(*red, L, w, uniqueNodes, r.get(), scal.get(), S_TMP1, s);
    fill<double>(L, 0.0, x.get(), s);
    fill<double>(L, 0.0, p.get(), s);
    wnorm(*red, L, w, r.get(), 1.0, scal.get(), S_TMP0, s);
    NRS_HIP(hipMemcpyAsync(scalHost, scal.get(), S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
    NRS_HIP(hipStreamSynchronize(s));
    const double r0 = scalHost[S_TMP0];
    lastIters = 0;
    if (r0 > 0.0 && !std::isnan(r0)) {
      pcg_init(scal.get(), r0, tol * r0, maxit, s);
      for (int done = 0; done < maxit; done += kChunk) {
        run_chunk(s);
        NRS_HIP(hipMemcpyAsync(scalHost, scal.get(), S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
        NRS_HIP(hipStreamSynchronize(s));
        if (scalHost[S_ACTIVE] == 0.0) break;
      }
      lastIters = (int)scalHost[S_ITER];
    }
    k_cast<double, T>(L, x.get(), xout, s);
  }

  template <typename T>
  void solve_amg(const T *rhs, T *xout, hipStream_t s)
  {
    if (deviceOnly) {
      // a fixed number of AMG-preconditioned PCG iterations, all on the
      // device (capturable): the reference's coarse solve is likewise a fixed
      // BoomerAMG cycle count (coarseLevel.cpp:80-93)
      rows_gather<T>(n, first.get(), rhs, ru.get(), s);
      if (nullSpace) remove_mean(*red, n, nullptr, (double)n, ru.get(), scal.get(), S_TMP1, s);
      fill<double>(n, 0.0, xu.get(), s);
      fill<double>(n, 0.0, pu.get(), s);
      wnorm(*red, n, ones.get(), ru.get(), 1.0, scal.get(), S_RDOTR, s);
      pcg_init_device(scal.get(), 0.0, false, maxit, s);
      for (int it = 0; it < maxit; ++it) {
        amg.apply(ru.get(), zu.get(), tmp.get(), s);
        pcg_precon_dots(*red, n, ones.get(), nullptr, ru.get(), zu.get(), Apu.get(), false, false, false,
                        scal.get(), s);
        pcg_update_p(n, zu.get(), pu.get(), scal.get(), s);
        csr_spmv(SPMV_Y, n, arp.get(), aci.get(), av.get(), pu.get(), Apu.get(), nullptr, nullptr, 0.0, s);
        pcg_alpha(*red, n, ones.get(), pu.get(), Apu.get(), scal.get(), s);
        pcg_update_rx(*red, n, ones.get(), Apu.get(), pu.get(), ru.get(), xu.get(), 1.0, scal.get(), s);
      }
      lastIters = maxit;
      rows_scatter<T>(L, rowOf.get(), xu.get(), xout, s);
      return;
    }
    rows_gather<T>(n, first.get(), rhs, ru.get(), s);
    if (nullSpace) remove_mean(*red, n, nullptr, (double)n, ru.get(), scal.get(), S_TMP1, s);
    fill<double>(n, 0.0, xu.get(), s);
    fill<double>(n, 0.0, pu.get(), s);
    wnorm(*red, n, ones.get(), ru.get(), 1.0, scal.get(), S_TMP0, s);
    NRS_HIP(hipMemcpyAsync(scalHost, scal.get(), S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
    NRS_HIP(hipStreamSynchronize(s));
    const double r0 = scalHost[S_TMP0];
    lastIters = 0;
    if (r0 > 0.0 && !std::isnan(r0)) {
      pcg_init(scal.get(), r0, tol * r0, maxit, s);
      for (int done = 0; done < maxit; done += kChunkAmg) {
        run_chunk(s);
        NRS_HIP(hipMemcpyAsync(scalHost, scal.get(), S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
        NRS_HIP(hipStreamSynchronize(s));
        if (scalHost[S_ACTIVE] == 0.0) break;
      }
      lastIters = (int)scalHost[S_ITER];
    }
    rows_scatter<T>(L, rowOf.get(), xu.get(), xout, s);
  }

  // Assemble the level operator on the unique unmasked nodes (element
  // matrices = the operator applied to unit vectors, one per element node)
  // and build the AMG hierarchy on it.
  void setup_amg(const std::vector<long long> &maskedIds, hipStream_t s)
  {
    const int Np = (N + 1) * (N + 1) * (N + 1);
    std::vector<int> rows(L, -1), firstH;
    {
      std::vector<std::pair<long long, int64_t>> ids;
      ids.reserve(L);
      for (int64_t i = 0; i < L; ++i)
        if (maskedIds[i] != 0) ids.emplace_back(maskedIds[i], i);
      std::sort(ids.begin(), ids.end());
      // rows by first appearance: the smallest local index of each id
      std::vector<std::pair<int64_t, long long>> firsts;
      for (size_t a = 0; a < ids.size();) {
        size_t b = a;
        while (b < ids.size() && ids[b].first == ids[a].first) ++b;
        firsts.emplace_back(ids[a].second, ids[a].first);
        a = b;
      }
      std::sort(firsts.begin(), firsts.end());
      std::vector<std::pair<long long, int>> rowOfId(firsts.size());
      for (size_t r = 0; r < firsts.size(); ++r) rowOfId[r] = {firsts[r].second, (int)r};
      std::sort(rowOfId.begin(), rowOfId.end());
      for (auto &pr : ids) {
        auto it = std::lower_bound(rowOfId.begin(), rowOfId.end(), std::make_pair(pr.first, -1));
        rows[pr.second] = it->second;
      }
      firstH.resize(firsts.size());
      for (size_t r = 0; r < firsts.size(); ++r) firstH[r] = (int)firsts[r].first;
    }
    n = (int)firstH.size();
    NRS_CHECK(n > 0, "coarse AMG: no unknowns on the coarse level");
    // element matrices: column b of every element = operator on unit vectors at node b
    std::vector<std::vector<double>> cols(Np);
    {
      DevBuf<double> u(L), Au(L);
      std::vector<double> uh(L);
      for (int b = 0; b < Np; ++b) {
        for (int64_t i = 0; i < L; ++i) uh[i] = ((i % Np) == b) ? 1.0 : 0.0;
        u.upload(uh);
        launch_ax<double>(N, E, nullptr, ggeo.get(), D.get(), lam0.get(), lam1.get(), poisson, false,
                          u.get(), Au.get(), nullptr, s);
        NRS_HIP(hipStreamSynchronize(s));
        cols[b] = Au.download();
      }
    }
    // assemble rows (columns ascending, duplicates summed in element order)
    std::vector<std::vector<std::pair<int, double>>> rowsAcc(n);
    for (int e = 0; e < E; ++e)
      for (int b = 0; b < Np; ++b) {
        const int rb = rows[(int64_t)e * Np + b];
        if (rb < 0) continue;
        for (int a = 0; a < Np; ++a) {
          const int ra = rows[(int64_t)e * Np + a];
          if (ra < 0) continue;
          const double v = cols[b][(int64_t)e * Np + a];
          auto &rw = rowsAcc[ra];
          bool found = false;
          for (auto &pr : rw)
            if (pr.first == rb) { pr.second += v; found = true; break; }
          if (!found) rw.emplace_back(rb, v);
        }
      }
    CsrHost A;
    A.n = A.m = n;
    A.rp.assign(n + 1, 0);
    for (int r = 0; r < n; ++r) {
      auto &rw = rowsAcc[r];
      std::sort(rw.begin(), rw.end());
      for (auto &pr : rw) { A.ci.push_back(pr.first); A.v.push_back(pr.second); }
      A.rp[r + 1] = (int)A.ci.size();
    }
    rowOf.upload(rows);
    first.upload(firstH);
    arp.upload(A.rp); aci.upload(A.ci); av.upload(A.v);
    ones.upload(std::vector<double>(n, 1.0));
    ru.alloc(n); xu.alloc(n); zu.alloc(n); pu.alloc(n); Apu.alloc(n); tmp.alloc(n);
    AmgParams ap;
    ap.nullSpace = nullSpace;
    amg.setup(A, ap);
    amgMode = true;
    deviceOnly = !(tol > 0.0);
    NRS_CHECK(!deviceOnly || (maxit >= 1 && maxit <= 64),
              "coarse AMG with coarseTol <= 0 runs exactly coarseMaxIter (1..64) iterations");
  }
};

template <typename T>
struct MGImpl final : MGSolver {
  nrs_elliptic *el = nullptr;
  std::vector<std::unique_ptr<Level<T>>> levels;
  DevBuf<T> sRes, sAd, sD, sTmp;  // shared smoother work vectors (finest size)
  DevBuf<T> rP, zP;               // level-0 FP32 copies of r and z
  DevBuf<T> cornersT;             // element corners [E][3][8] of a trilinear mesh (else empty)
  std::unique_ptr<CoarsePCG> coarse;
  enum Cheby { NO_CHEBY, CHEBY1, CHEBY_OPT4 };
  int cheby = CHEBY1;
  int schwarzKind = 0;            // 0 Jacobi, 1 ASM, 2 RAS
  bool coarseSmoothing = false;      // coarse level smoothed (degree 3), no coarse solve
  bool coarseJacobi = false;         // ... with Chebyshev-Jacobi whatever the finer levels use
  bool coarseUser = false;           // the caller's coarse solve (el->userCoarse)
  // Replay of the whole V-cycle as one HIP graph (every argument lives in
  // the handle: level vectors, device scalars, host-fixed Chebyshev
  // coefficients), used when the cycle never waits on the host
  bool cycleGraphs = false;
  DevBuf<double> rIn, zOut;          // fixed FP64 cycle input / output (FP64 multigrid)
  hipStream_t cycCap = nullptr;
  hipGraph_t cycGraph = nullptr;
  hipGraphExec_t cycExec = nullptr;
  ~MGImpl() override
  {
    if (cycExec) (void)hipGraphExecDestroy(cycExec);
    if (cycGraph) (void)hipGraphDestroy(cycGraph);
    if (cycCap) (void)hipStreamDestroy(cycCap);
  }

  // ---- level operators (pMGLevel::Ax / residual / smoother)
  // the level operator; the finest level shares the solver's chunked handle,
  // so its Ax overlaps its gather-scatter as the outer operator's does
  void op(Level<T> &l, const T *q, T *Aq, hipStream_t s)
  {
    const int k = (l.ogs->nChunks == nrs_elliptic::kOpChunks) ? operator_chunks(l.N, dtype_of<T>::value, l.E) : 1;
    chunked_operator<T>(el, l.N, l.E, l.ggeo, l.D, l.lam0.get(), l.lam1.get(), el->poisson, l.maskBits,
                        l.ogs, q, Aq, k, s, false, nullptr, false, l.corners, l.corners ? l.zw.get() : nullptr);
  }

  // pMGLevel::smoother (ellipticMultiGridLevel.cpp:104-111); out != in
  void smootherS(Level<T> &l, const T *in, T *out, hipStream_t s)
  {
    if (l.schwarz) l.schwarz->apply(in, out, s);
    else k_wmul<T>(l.L, l.invDiagA.get(), in, out, s);
  }

  // smoothChebyshev (ellipticMultiGridLevel.cpp:144-210).  The reference
  // smooths res and Ad in place; here S writes to a spare vector whose role is
  // swapped with the input's.  Scalars are formed in FP64 and rounded to T.
  // mask bits for the last Chebyshev update to store x's masked nodes as zero
  // (null: no masked node, or a level too long for the kernels' 32-bit index)
  static const uint32_t *last_mask(const Level<T> &l)
  {
    return (l.Nmasked > 0 && l.maskBits && l.L <= (int64_t)UINT32_MAX) ? l.maskBits : nullptr;
  }

  void smooth_cheby1(Level<T> &l, const T *rhs, T *x, bool xIsZero, hipStream_t s)
  {
    const int deg = xIsZero ? l.degDown : l.degUp;
    if (deg == 0) return;
    const uint32_t *lm = last_mask(l);
    const T theta = T(0.5 * (l.lmax + l.lmin));
    const T delta = T(0.5 * (l.lmax - l.lmin));
    const T invTheta = T(1.0 / (double)theta);
    const T sigma = theta / delta;
    T rho_n = T(1.0 / (double)sigma);
    T *res = sRes.get(), *Ad = sAd.get(), *d = sD.get(), *spare = sTmp.get();
    // Jacobi smoothing (no Schwarz): S is applied inside the update kernels, and
    // the residual rhs - A x inside the first of them
    const bool jac = !l.schwarz;
    // x = 0 with at least one update step: that step takes x as zero (no fill)
    const bool x0 = xIsZero && deg > 1;
    if (xIsZero) {
      if (!x0) fill<T>(l.L, T(0), x, s);
      if (!jac) k_axpby<T>(l.L, T(1), rhs, T(0), res, s);   // res = rhs
    } else {
      op(l, x, res, s);
      if (!jac) k_axpby<T>(l.L, T(1), rhs, T(-1), res, s);
    }
    if (jac) {
      if (xIsZero)
        jacobi_first_kernel<T, false><<<kgrid(l.L), KB, 0, s>>>(l.L, invTheta, l.invDiagA.get(), rhs,
                                                                nullptr, spare, d);
      else
        jacobi_first_kernel<T, true><<<kgrid(l.L), KB, 0, s>>>(l.L, invTheta, l.invDiagA.get(), rhs, res,
                                                               spare, d);
      NRS_LAUNCH_CHECK();
      std::swap(res, spare);
    } else {
      smootherS(l, res, spare, s);
      std::swap(res, spare);
      k_axpby<T>(l.L, invTheta, res, T(0), d, s);
    }
    for (int k = 1; k < deg; ++k) {
      op(l, d, Ad, s);
      const T rhoSave = rho_n;
      rho_n = T(1.0 / (2.0 * (double)sigma - (double)rho_n));
      const T rC = T(2.0 * (double)rho_n / (double)delta);
      const T dC = rho_n * rhoSave;
      const bool last = k + 1 == deg;   // fused with the final x += d
      if (!jac) {
        smootherS(l, Ad, spare, s);
        std::swap(Ad, spare);
      }
      const T *w = l.invDiagA.get();
      const unsigned g = kgrid(l.L);
      const bool first0 = x0 && k == 1;
#define NRS_CHEBY1(J_, L_)                                                                          \
  do {                                                                                            \
    if (first0) cheby1_kernel<T, J_, L_, true><<<g, KB, 0, s>>>(l.L, dC, rC, w, Ad, d, res, x, lm, l.Np); \
    else cheby1_kernel<T, J_, L_, false><<<g, KB, 0, s>>>(l.L, dC, rC, w, Ad, d, res, x, lm, l.Np);       \
  } while (0)
      if (jac && last) NRS_CHEBY1(true, true);
      else if (jac) NRS_CHEBY1(true, false);
      else if (last) NRS_CHEBY1(false, true);
      else NRS_CHEBY1(false, false);
#undef NRS_CHEBY1
      NRS_LAUNCH_CHECK();
    }
    if (deg == 1) k_axpby<T>(l.L, T(1), d, T(1), x, s);
    if (deg == 1 || !lm) launch_mask<T>(l.Nmasked, l.maskIds, x, s);   // else in the last update
  }

  // smoothFourthKindChebyshev (ellipticMultiGridLevel.cpp:212-270)
  void smooth_cheby4(Level<T> &l, const T *rhs, T *x, bool xIsZero, hipStream_t s)
  {
    const int deg = xIsZero ? l.degDown : l.degUp;
    if (deg == 0) return;
    const auto &betas = xIsZero ? l.betasDown : l.betasUp;
    const double rho = l.lmax;
    const uint32_t *lm = last_mask(l);
    T *res = sRes.get(), *Ad = sAd.get(), *d = sD.get();
    // x = 0 with at least one update step: the first step treats x as zero and
    // takes the residual from rhs, so neither is written beforehand
    const bool x0 = xIsZero && deg > 1;
    if (xIsZero) {
      if (!x0) {
        fill<T>(l.L, T(0), x, s);
        k_axpby<T>(l.L, T(1), rhs, T(0), res, s);   // res = rhs
      }
    } else {
      op(l, x, res, s);
      k_axpby<T>(l.L, T(1), rhs, T(-1), res, s);
    }
    smootherS(l, x0 ? rhs : res, Ad, s);
    const T coeff = T(4.0 / (3.0 * rho));
    k_axpby<T>(l.L, coeff, Ad, T(0), d, s);
    for (int k = 1; k < deg; ++k) {
      op(l, d, Ad, s);
      const bool first0 = k == 1 && x0, last = k + 1 == deg;
      const unsigned g = kgrid(l.L);
      if (first0) cheby4_res_kernel<T, true><<<g, KB, 0, s>>>(l.L, Ad, rhs, res);
      else cheby4_res_kernel<T, false><<<g, KB, 0, s>>>(l.L, Ad, res, res);
      NRS_LAUNCH_CHECK();
      smootherS(l, res, Ad, s);
      const T dC = T((2.0 * k - 1.0) / (2.0 * k + 3.0));
      const T rC = T((8.0 * k + 4.0) / ((2.0 * k + 3.0) * rho));
      const T b = betas[k - 1], bl = betas.back();
      // x += beta_(k-1) d deferred to here, with the direction update (and, on
      // the last step, the final x += beta_last d)
      if (first0 && last) cheby4_dir_kernel<T, true, true><<<g, KB, 0, s>>>(l.L, b, rC, dC, bl, Ad, d, x, lm, l.Np);
      else if (first0) cheby4_dir_kernel<T, true, false><<<g, KB, 0, s>>>(l.L, b, rC, dC, bl, Ad, d, x);
      else if (last) cheby4_dir_kernel<T, false, true><<<g, KB, 0, s>>>(l.L, b, rC, dC, bl, Ad, d, x, lm, l.Np);
      else cheby4_dir_kernel<T, false, false><<<g, KB, 0, s>>>(l.L, b, rC, dC, bl, Ad, d, x);
      NRS_LAUNCH_CHECK();
    }
    if (deg == 1) k_axpby<T>(l.L, betas.back(), d, T(1), x, s);
    if (deg == 1 || !lm) launch_mask<T>(l.Nmasked, l.maskIds, x, s);   // else in the last update
  }

  // pMGLevel::smooth (ellipticMultiGridLevel.cpp:85-103): pure ASM / RAS
  // smooths only on the way down (x = 0)
  void smooth(Level<T> &l, const T *rhs, T *x, bool xIsZero, hipStream_t s)
  {
    if (cheby == NO_CHEBY) {
      if (xIsZero) smootherS(l, rhs, x, s);
      return;
    }
    if (cheby == CHEBY_OPT4) smooth_cheby4(l, rhs, x, xIsZero, s);
    else smooth_cheby1(l, rhs, x, xIsZero, s);
  }

  void vcycle(size_t k, hipStream_t s)
  {
    Level<T> &l = *levels[k];
    if (l.isCoarse) {
      if (coarse) coarse->solve<T>(l.rhs, l.x, s);
      else if (coarseUser) el->userCoarse(el->userCoarseCtx, l.rhs, l.x, (void *)s);   // solvePtr
      else smooth(l, l.rhs, l.x, true, s);   // MULTIGRID COARSE SOLVE = FALSE (ellipticMultiGridSetup.cpp:315-320)
      return;
    }
    Level<T> &c = *levels[k + 1];
    smooth(l, l.rhs, l.x, true, s);
    // residual (ellipticMultiGridLevel.cpp:37-41) and coarsen (:43-68):
    // invDegreeFine .* (rhs - A x) formed as the contraction loads A x (the
    // residual itself is not needed afterwards), contract, gs, mask.  The
    // coarse level's masked nodes are in no row of its gather-scatter, so the
    // contraction stores them as zero and the mask needs no pass of its own.
    op(l, l.x, l.res.get(), s);
    launch_coarsen<T>(l.N, c.N, l.E, c.R.get(), l.res.get(), c.rhs, s, l.invDegree.get(), l.rhs,
                      c.Nmasked ? c.maskBits : nullptr);
    ogs_gather_scatter<T>(c.ogs, c.rhs, 1, c.L, NRS_ADD, s);
    vcycle(k + 1, s);
    // prolongate (:70-83): x += P xC
    launch_prolongate<T>(l.N, c.N, l.E, c.R.get(), c.x, l.x, s);
    smooth(l, l.rhs, l.x, false, s);
  }

  void apply(const double *r, double *z, hipStream_t s) override
  {
    NRS_CHECK(!coarseUser || el->userCoarse, "multigrid: coarse solver USER without a coarse solve");
    Level<T> &l0 = *levels[0];
    hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
    NRS_HIP(hipStreamIsCapturing(s, &cst));
    if (cycleGraphs && device_only() && cst == hipStreamCaptureStatusNone) {
      // input and output through fixed buffers, the cycle itself replayed
      if (std::is_same<T, double>::value) {
        k_axpby<double>(l0.L, 1.0, r, 0.0, rIn.get(), s);
        l0.rhs = (T *)rIn.get();
        l0.x = (T *)zOut.get();
      } else {
        k_cast<double, T>(l0.L, r, rP.get(), s);
        l0.rhs = rP.get();
        l0.x = zP.get();
      }
      if (!cycExec) {
        if (!cycCap) NRS_HIP(hipStreamCreateWithFlags(&cycCap, hipStreamNonBlocking));
        NRS_HIP(hipStreamBeginCapture(cycCap, hipStreamCaptureModeThreadLocal));
        vcycle(0, cycCap);
        NRS_HIP(hipStreamEndCapture(cycCap, &cycGraph));
        NRS_HIP(hipGraphInstantiate(&cycExec, cycGraph, nullptr, nullptr, 0));
      }
      NRS_HIP(hipGraphLaunch(cycExec, s));
      k_cast<T, double>(l0.L, l0.x, z, s);
      return;
    }
    if (std::is_same<T, double>::value) {
      l0.rhs = (T *)r;   // read-only inside the cycle
      l0.x = (T *)z;
    } else {
      k_cast<double, T>(l0.L, r, rP.get(), s);
      l0.rhs = rP.get();
      l0.x = zP.get();
    }
    vcycle(0, s);
    k_cast<T, double>(l0.L, l0.x, z, s);
  }

  std::vector<MGLevelInfo> info() const override
  {
    std::vector<MGLevelInfo> out;
    for (auto &l : levels) out.push_back({l->N, l->L, l->rho, l->degDown});
    return out;
  }
  int coarse_iterations() const override { return coarse ? coarse->lastIters : 0; }
  bool device_only() const override { return !coarseUser && (coarse == nullptr || coarse->deviceOnly); }
  void level_smoother(int level, const void *in, void *out, hipStream_t s) override
  {
    NRS_CHECK(level >= 0 && level < (int)levels.size() && (level + 1 < (int)levels.size() || coarseSmoothing),
              "mg smoother: not a smoothed level");
    NRS_CHECK(in != out, "mg smoother: input and output must differ");
    smootherS(*levels[level], (const T *)in, (T *)out, s);
  }

  void level_op(int level, int opk, void *in, const void *xv, void *outv, bool xIsZero,
                hipStream_t s) override
  {
    const int nl = (int)levels.size();
    NRS_CHECK(level >= 0 && level < nl, "mg level: no such level");
    NRS_CHECK(in && outv, "mg level: null vector");
    Level<T> &l = *levels[level];
    T *a = (T *)in, *out = (T *)outv;
    switch (opk) {
      case NRS_MG_AX:   // pMGLevel::Ax (ellipticMultiGridLevel.cpp:32-35)
        NRS_CHECK(a != out, "mg level Ax: input and output must differ");
        op(l, a, out, s);
        break;
      case NRS_MG_RESIDUAL:   // :37-41, out = in - A x
        NRS_CHECK(xv && out != a && out != xv, "mg level residual: x needed, out distinct");
        op(l, (const T *)xv, out, s);
        k_axpby<T>(l.L, T(1), a, T(-1), out, s);
        break;
      case NRS_MG_SMOOTH:     // :85-103, rhs = in, x = out
        NRS_CHECK(!l.isCoarse || coarseSmoothing, "mg level smooth: the coarsest level is solved, not smoothed");
        NRS_CHECK(a != out, "mg level smooth: rhs and x must differ");
        smooth(l, a, out, xIsZero, s);
        break;
      case NRS_MG_COARSEN: {  // :43-68 on the coarse level: in (level-1) .*= invDegree, out = R in
        NRS_CHECK(level >= 1, "mg level coarsen: level is the coarse level (>= 1)");
        Level<T> &f = *levels[level - 1];
        k_wmul<T>(f.L, f.invDegree.get(), a, a, s);
        launch_coarsen<T>(f.N, l.N, f.E, l.R.get(), a, out, s);
        ogs_gather_scatter<T>(l.ogs, out, 1, l.L, NRS_ADD, s);
        launch_mask<T>(l.Nmasked, l.maskIds, out, s);
        break;
      }
      case NRS_MG_PROLONGATE: {   // :70-83 on the coarse level: out (level-1) += P in
        NRS_CHECK(level >= 1, "mg level prolongate: level is the coarse level (>= 1)");
        Level<T> &f = *levels[level - 1];
        launch_prolongate<T>(f.N, l.N, f.E, l.R.get(), a, out, s);
        break;
      }
      default:
        throw Error(NRS_ERR_INVALID, "mg level: unknown operation");
    }
  }

  // ---- setup helpers

  // lambda_max(S A) by a k-step Arnoldi in FP64 with the level operator in T
  // (pMGLevel::maxEigSmoothAx, ellipticMultiGridLevelSetup.cpp:292-453).
  double max_eig(Level<T> &l, uint64_t seed, hipStream_t s)
  {
    const int k = (int)std::min<int64_t>(10, l.L);
    const int64_t L = l.L;
    std::vector<double> Vx(L);
    SplitMix rng{seed + 0x5eedull * (uint64_t)(l.N + 1)};
    for (auto &v : Vx) v = rng.next();
    std::vector<DevBuf<double>> V(k + 1);
    for (auto &b : V) b.alloc(L);
    DevBuf<double> vx;
    vx.upload(Vx);
    ogs_gather_scatter<double>(l.ogsFull, vx.get(), 1, L, NRS_ADD, s);
    launch_mask<double>(l.Nmasked, l.maskIds, vx.get(), s);
    DevBuf<T> vT(L), avT(L);
    DevBuf<double> sc(4);
    Reducer red;
    std::vector<double> h(4);
    auto dot = [&](const double *a, const double *b) {
      wdot(red, L, l.ogs->invDegree.get(), a, b, sc.get(), 0, s);
      NRS_HIP(hipMemcpyAsync(h.data(), sc.get(), sizeof(double), hipMemcpyDeviceToHost, s));
      NRS_HIP(hipStreamSynchronize(s));
      return h[0];
    };
    const double nv = std::sqrt(dot(vx.get(), vx.get()));
    // a level without unknowns (every node masked): nothing to smooth, any
    // positive bound serves (the reference divides by the zero norm)
    if (!(nv > 0.0)) return 1.0;
    axpby<double>(L, 1.0 / nv, vx.get(), 0.0, V[0].get(), s);
    std::vector<double> H((size_t)k * k, 0.0);
    int m = k;   // dimension of the Krylov space reached
    for (int j = 0; j < k; ++j) {
      k_cast<double, T>(L, V[j].get(), vT.get(), s);
      op(l, vT.get(), avT.get(), s);
      smootherS(l, avT.get(), vT.get(), s);
      k_cast<T, double>(L, vT.get(), V[j + 1].get(), s);
      double hn = 0.0;
      for (int i = 0; i <= j; ++i) {
        const double hij = dot(V[i].get(), V[j + 1].get());
        axpby<double>(L, -hij, V[i].get(), 1.0, V[j + 1].get(), s);
        H[i + j * k] = hij;
        hn += hij * hij;
      }
      if (j + 1 < k) {
        const double nj = std::sqrt(dot(V[j + 1].get(), V[j + 1].get()));
        // breakdown: the space is invariant (a level with fewer unknowns than
        // steps), its leading block holds exact eigenvalues; the reference
        // normalises the rounding residue instead (NaN or a spurious rho)
        if (nj <= 1e-8 * std::sqrt(hn + nj * nj)) { m = j + 1; break; }
        axpby<double>(L, 1.0 / nj, V[j + 1].get(), 0.0, V[j + 1].get(), s);
        H[j + 1 + j * k] = nj;
      }
    }
    if (m < k) {
      std::vector<double> Hm((size_t)m * m);
      for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) Hm[i + j * m] = H[i + j * k];
      return spectral_radius(Hm, m);
    }
    return spectral_radius(H, k);
  }

  void build_level_mesh(Level<T> &l, const std::vector<double> &x, const std::vector<double> &y,
                        const std::vector<double> &z, const std::vector<long long> &gid, const int *EToB,
                        hipStream_t s)
  {
    std::vector<double> zz, ww;
    gll_nodes(l.N, zz, ww);
    const auto Dh = dmatrix(zz);
    DevBuf<double> D64, w64, xd, yd, zd, g64;
    D64.upload(Dh);
    w64.upload(ww);
    xd.upload(x); yd.upload(y); zd.upload(z);
    g64.alloc((size_t)l.L * 7);
    launch_geofacs(l.N, l.E, D64.get(), w64.get(), xd.get(), yd.get(), zd.get(), g64.get(), s);
    l.ggeoOwn.alloc((size_t)l.L * 7);
    l.DOwn.alloc(Dh.size());
    k_cast<double, T>((int64_t)l.L * 7, g64.get(), l.ggeoOwn.get(), s);
    k_cast<double, T>((int64_t)Dh.size(), D64.get(), l.DOwn.get(), s);
    l.ggeo = l.ggeoOwn.get();
    l.D = l.DOwn.get();
    l.ogsFull = ogs_setup(l.L, gid.data());
    const auto mids = dirichlet_mask(l.N, l.E, EToB, l.ogsFull, s);
    l.Nmasked = (int64_t)mids.size();
    l.maskIdsOwn.upload(mids);
    l.maskBitsOwn.upload(mask_bits(l.Np, l.E, mids));
    l.maskIds = l.maskIdsOwn.get();
    l.maskBits = l.maskBitsOwn.get();
    std::vector<long long> mg(gid);
    for (int id : mids) mg[id] = 0;
    l.ogs = ogs_setup(l.L, mg.data());
    l.ownOgs = true;
    if (l.isCoarse && !coarseSmoothing && !coarseUser) {   // FP64 operator for the coarse PCG
      coarse = std::make_unique<CoarsePCG>();
      CoarsePCG &c = *coarse;
      c.N = l.N; c.E = l.E; c.L = l.L;
      c.ggeo = std::move(g64);
      c.D = std::move(D64);
      c.lam0.upload(std::vector<double>{el->lambda0});
      c.lam1.upload(std::vector<double>{el->lambda1});
      c.poisson = el->poisson;
      c.ogs = l.ogs;
      c.maskBits = l.maskBits;
      c.tol = el->opts.coarseTol;
      c.maxit = el->opts.coarseMaxIter;
      c.useGraphs = el->opts.useGraphs != 0;
      c.invDiag.alloc(l.L);
      launch_build_diag<double>(c.N, c.E, c.ggeo.get(), c.D.get(), c.lam0.get(), c.lam1.get(), c.poisson,
                                false, c.invDiag.get(), s);
      ogs_gather_scatter<double>(c.ogs, c.invDiag.get(), 1, c.L, NRS_ADD, s);
      reciprocal<double>(c.L, c.invDiag.get(), s);
      c.r.alloc(l.L); c.x.alloc(l.L); c.z.alloc(l.L); c.p.alloc(l.L); c.Ap.alloc(l.L);
      c.scal.alloc(S_COUNT);
      NRS_HIP(hipMemset(c.scal.get(), 0, S_COUNT * sizeof(double)));
      NRS_HIP(hipHostMalloc((void **)&c.scalHost, S_COUNT * sizeof(double), hipHostMallocDefault));
      c.red = std::make_unique<Reducer>();
      c.nullSpace = el->allNeumann;
      if (c.nullSpace) {
        const auto wh = l.ogs->invDegree.download();
        for (double v : wh) c.uniqueNodes += v;
      }
      if (el->opts.coarseSolver == NRS_COARSE_AMG) c.setup_amg(mg, s);
    }
    NRS_HIP(hipStreamSynchronize(s));
  }

  void setup(nrs_elliptic *e, const nrs_elliptic_desc *d, hipStream_t s)
  {
    el = e;
    const auto &o = e->opts;
    switch (o.smoother) {
      case NRS_SMOOTHER_CHEBYSHEV_JACOBI: cheby = CHEBY1; schwarzKind = 0; break;
      case NRS_SMOOTHER_OPT4_CHEBYSHEV_JACOBI: cheby = CHEBY_OPT4; schwarzKind = 0; break;
      case NRS_SMOOTHER_CHEBYSHEV_ASM: cheby = CHEBY1; schwarzKind = 1; break;
      case NRS_SMOOTHER_OPT4_CHEBYSHEV_ASM: cheby = CHEBY_OPT4; schwarzKind = 1; break;
      case NRS_SMOOTHER_CHEBYSHEV_RAS: cheby = CHEBY1; schwarzKind = 2; break;
      case NRS_SMOOTHER_OPT4_CHEBYSHEV_RAS: cheby = CHEBY_OPT4; schwarzKind = 2; break;
      case NRS_SMOOTHER_ASM: cheby = NO_CHEBY; schwarzKind = 1; break;
      case NRS_SMOOTHER_RAS: cheby = NO_CHEBY; schwarzKind = 2; break;
      default: throw Error(NRS_ERR_INVALID, "multigrid: unknown smoother");
    }
    NRS_CHECK(o.coarseSolver == NRS_COARSE_PCG_F64 || o.coarseSolver == NRS_COARSE_SMOOTH ||
                  o.coarseSolver == NRS_COARSE_SMOOTH_JACOBI || o.coarseSolver == NRS_COARSE_AMG ||
                  o.coarseSolver == NRS_COARSE_USER,
              "multigrid: unknown coarse solver");
    coarseUser = o.coarseSolver == NRS_COARSE_USER;
    coarseSmoothing = o.coarseSolver == NRS_COARSE_SMOOTH || o.coarseSolver == NRS_COARSE_SMOOTH_JACOBI;
    coarseJacobi = o.coarseSolver == NRS_COARSE_SMOOTH_JACOBI;
    NRS_CHECK(!(coarseJacobi && cheby == NO_CHEBY),
              "multigrid: a Jacobi-smoothed coarse level needs a Chebyshev smoother");
    NRS_CHECK(o.fdmPrecision == NRS_FDM_F32 || o.fdmPrecision == NRS_FDM_BF16 ||
                  o.fdmPrecision == NRS_FDM_F16,
              "multigrid: unknown FDM precision");
    // level schedule (determineMGLevels.cpp:58-95)
    std::vector<int> orders;
    if (o.nLevels > 0) {
      for (int i = 0; i < o.nLevels && i < 8; ++i) orders.push_back(o.levels[i]);
    } else if (schwarzKind) {
      static const std::map<int, std::vector<int>> table = {
          {1, {1}}, {2, {2, 1}}, {3, {3, 1}}, {4, {4, 2, 1}}, {5, {5, 3, 1}},
          {6, {6, 3, 1}}, {7, {7, 3, 1}}, {8, {8, 5, 1}}, {9, {9, 5, 1}}};
      orders = table.at(e->N);
    } else {
      static const std::map<int, std::vector<int>> table = {
          {1, {1}}, {2, {2, 1}}, {3, {3, 1}}, {4, {4, 2, 1}}, {5, {5, 3, 1}},
          {6, {6, 4, 2, 1}}, {7, {7, 5, 3, 1}}, {8, {8, 6, 4, 1}}, {9, {9, 7, 5, 1}}};
      orders = table.at(e->N);
    }
    NRS_CHECK(!orders.empty() && orders[0] == e->N, "multigrid: the first level must be the mesh order");
    for (size_t i = 1; i < orders.size(); ++i)
      NRS_CHECK(orders[i] >= 1 && orders[i] < orders[i - 1], "multigrid: orders must decrease");
    NRS_CHECK(orders.size() >= 2, "multigrid: at least two levels are required");
    NRS_CHECK(o.chebyDegree >= 1 && o.chebyDegree <= 16, "multigrid: Chebyshev degree must be 1..16");
    for (int i = 0; i < 8; ++i)
      NRS_CHECK(o.preDegree[i] >= 0 && o.preDegree[i] <= 16 && o.postDegree[i] >= 0 && o.postDegree[i] <= 16,
                "multigrid: per-level Chebyshev degrees must be 0 (default) or 1..16");

    std::vector<double> xf(d->x, d->x + e->Nlocal), yf(d->y, d->y + e->Nlocal), zf(d->z, d->z + e->Nlocal);
    std::vector<long long> gf(d->globalIds, d->globalIds + e->Nlocal);
    // trilinear elements: the level operators form their geometric factors from
    // the corners (the same map on every level: coarse coordinates interpolate it)
    {
      std::vector<double> cc;
      if (o.mgGeometry != 0 && trilinear_corners(e->N, e->E, d->x, d->y, d->z, cc)) {
        std::vector<T> ct(cc.begin(), cc.end());
        cornersT.upload(ct);
      }
    }
    // Schwarz: face links of the mesh and element lengths of the finest level
    // (pMGLevel::build passes the finest elliptic to compute_element_lengths)
    std::vector<FaceLink> links;
    std::vector<double> lengths;
    if (schwarzKind) {
      links = face_links(e->N, e->E, gf.data());
      lengths = element_lengths(e->N, e->E, xf.data(), yf.data(), zf.data(), links);
    }
    int Nprev = e->N;
    for (size_t li = 0; li < orders.size(); ++li) {
      auto lp = std::make_unique<Level<T>>();
      Level<T> &l = *lp;
      l.N = orders[li]; l.Nq = l.N + 1; l.Np = l.Nq * l.Nq * l.Nq; l.E = e->E;
      l.L = (int64_t)l.E * l.Np;
      l.isCoarse = (li + 1 == orders.size());
      l.lam0.upload(std::vector<T>{(T)e->lambda0});
      l.lam1.upload(std::vector<T>{(T)e->lambda1});
      if (li == 0) {
        // finest level shares the solver's mesh, ogs and mask (ellipticBuildMultigridLevelFine.cpp)
        if (std::is_same<T, double>::value) {
          l.ggeo = (const T *)e->ggeo.get();
          l.D = (const T *)e->D.get();
        } else {
          l.ggeoOwn.alloc((size_t)l.L * 7);
          l.DOwn.alloc(e->Dh.size());
          k_cast<double, T>((int64_t)l.L * 7, e->ggeo.get(), l.ggeoOwn.get(), s);
          k_cast<double, T>((int64_t)e->Dh.size(), e->D.get(), l.DOwn.get(), s);
          l.ggeo = l.ggeoOwn.get();
          l.D = l.DOwn.get();
        }
        l.ogsFull = e->ogsFull;
        l.ogs = e->ogs;
        l.Nmasked = e->Nmasked;
        l.maskIds = e->maskIds.get();
        l.maskBits = e->maskBits.get();
      } else {
        xf = interp_coords(Nprev, l.E, xf.data(), l.N);
        yf = interp_coords(Nprev, l.E, yf.data(), l.N);
        zf = interp_coords(Nprev, l.E, zf.data(), l.N);
        gf = coarse_global_ids(Nprev, l.E, gf.data(), l.N);
        build_level_mesh(l, xf, yf, zf, gf, d->EToB, s);
        // R = transpose(InterpolationMatrix1D(Nc -> Nf)) (ellipticMultiGridLevelSetup.cpp:238-258)
        std::vector<double> zc, wc, zF, wF;
        gll_nodes(l.N, zc, wc);
        gll_nodes(Nprev, zF, wF);
        const auto I = interp_matrix(zc, zF);   // [NqF][NqC]
        const int NqF = Nprev + 1, NqC = l.N + 1;
        std::vector<T> R((size_t)NqC * NqF);
        for (int i = 0; i < NqC; ++i)
          for (int j = 0; j < NqF; ++j) R[i * NqF + j] = (T)I[j * NqC + i];
        l.R.upload(R);
      }
      // corner geometry: FP32 at every order, FP64 up to order 7 (beyond, its registers spill)
      if (cornersT.size() && (!std::is_same<T, double>::value || l.N <= 7)) {
        std::vector<double> zl, wl;
        gll_nodes(l.N, zl, wl);
        std::vector<T> zwh(2 * (size_t)l.Nq);
        for (int q = 0; q < l.Nq; ++q) { zwh[q] = (T)zl[q]; zwh[l.Nq + q] = (T)wl[q]; }
        l.zw.upload(zwh);
        l.corners = cornersT.get();
      }
      l.invDegree.alloc(l.L);
      k_cast<double, T>(l.L, l.ogs->invDegree.get(), l.invDegree.get(), s);
      l.res.alloc(l.L);
      if (li > 0) {
        l.rhsOwn.alloc(l.L); l.xOwn.alloc(l.L);
        l.rhs = l.rhsOwn.get(); l.x = l.xOwn.get();
      }
      if (!l.isCoarse || coarseSmoothing) {
        if (schwarzKind && !(l.isCoarse && coarseJacobi)) {
          // Schwarz smoother (pMGLevel::build, ellipticMultiGridSchwarz.cpp:837-1054)
          l.schwarz = std::make_unique<Schwarz<T>>();
          l.schwarz->setup(l.N, l.E, schwarzKind == 2, o.fdmPrecision, lengths, d->EToB, links,
                           l.ogs, l.invDegree.get(), l.Nmasked, l.maskIds, l.maskBits, el->poisson,
                           el->lambda0, el->lambda1, s);
        } else {
          // Jacobi smoother: 1/diag(A), assembled (ellipticUpdateJacobi.cpp:36-61)
          l.invDiagA.alloc(l.L);
          launch_build_diag<T>(l.N, l.E, l.ggeo, l.D, l.lam0.get(), l.lam1.get(), el->poisson, false,
                               l.invDiagA.get(), s);
          ogs_gather_scatter<T>(l.ogs, l.invDiagA.get(), 1, l.L, NRS_ADD, s);
          reciprocal<T>(l.L, l.invDiagA.get(), s);
        }
        // the coarsest level, when smoothed, uses degree 3 (ellipticMultiGridLevelSetup.cpp:143-150);
        // the others MULTIGRID CHEBYSHEV DEGREE or the schedule's (:154-167)
        if (cheby == NO_CHEBY) {
          l.degDown = l.degUp = 1;
        } else if (l.isCoarse) {
          l.degDown = l.degUp = 3;
        } else {
          l.degDown = o.preDegree[li] > 0 ? o.preDegree[li] : o.chebyDegree;
          l.degUp = o.postDegree[li] > 0 ? o.postDegree[li] : o.chebyDegree;
        }
        if (cheby == CHEBY_OPT4) {
          l.betasDown = opt4_betas<T>(l.degDown);
          l.betasUp = opt4_betas<T>(l.degUp);
        }
        if (o.autotune && l.schwarz) l.schwarz->autotune(s);
      }
      // level operator launch variants (registerEllipticPreconditionerKernels.cpp:52-66)
      if (o.autotune) {
        autotune_ax<T>(l.N, l.E, l.ggeo, l.D, l.lam0.get(), l.lam1.get(), el->poisson, s);
        if (l.ogs->nChunks == nrs_elliptic::kOpChunks)
          autotune_operator<T>(el, l.N, l.E, l.ggeo, l.D, l.lam0.get(), l.lam1.get(), el->poisson,
                               l.maskBits, l.ogs, s);
      }
      Nprev = l.N;
      levels.push_back(std::move(lp));
    }
    const int64_t L0 = levels[0]->L;
    sRes.alloc(L0); sAd.alloc(L0); sD.alloc(L0); sTmp.alloc(L0);
    cycleGraphs = o.useGraphs != 0;
    if (cycleGraphs && std::is_same<T, double>::value) { rIn.alloc(L0); zOut.alloc(L0); }
    if (!std::is_same<T, double>::value) { rP.alloc(L0); zP.alloc(L0); }
    for (auto &lp : levels) {
      Level<T> &l = *lp;
      if ((l.isCoarse && !coarseSmoothing) || cheby == NO_CHEBY) continue;
      l.rho = max_eig(l, o.seed, s);
      l.lmax = o.chebyMaxFactor * l.rho;
      l.lmin = o.chebyMinFactor * l.rho;
    }
    NRS_HIP(hipStreamSynchronize(s));
  }

  // optimal fourth-kind Chebyshev betas (MG/optimalCoeffs.cpp:4-215; Lottes,
  // arXiv:2202.08830), degrees 1..8 tabulated here
  template <typename U> static std::vector<U> opt4_betas(int deg)
  {
    static const std::vector<std::vector<double>> t = {
        {1.12500000000000},
        {1.02387287570313, 1.26408905371085},
        {1.00842544782028, 1.08867839208730, 1.33753125909618},
        {1.00391310427285, 1.04035811188593, 1.14863498546254, 1.38268869241000},
        {1.00212930146164, 1.02173711549260, 1.07872433192603, 1.19810065292663, 1.41322542791682},
        {1.00128517255940, 1.01304293035233, 1.04678215124113, 1.11616489419675, 1.23829020218444, 1.43524297106744},
        {1.00083464397912, 1.00843949430122, 1.03008707768713, 1.07408384092003, 1.15036186707366, 1.27116474046139, 1.45186658649364},
        {1.00057246631197, 1.00577427662415, 1.02050187922941, 1.05019803444565, 1.10115572984941, 1.18086042806856, 1.29838585382576, 1.46486073151099}};
    NRS_CHECK(deg >= 1 && deg <= (int)t.size(), "opt. 4th-kind Chebyshev: degree 1..8 built");
    std::vector<U> b;
    for (double v : t[deg - 1]) b.push_back((U)v);
    return b;
  }
};

std::unique_ptr<MGSolver> mg_create(nrs_elliptic *el, const nrs_elliptic_desc *d, hipStream_t s)
{
  if (el->opts.mgPrecision == NRS_MG_F64) {
    auto m = std::make_unique<MGImpl<double>>();
    m->setup(el, d, s);
    return m;
  }
  auto m = std::make_unique<MGImpl<float>>();
  m->setup(el, d, s);
  return m;
}

}  // namespace nrs
