// amg.hip -- smoothed-aggregation AMG for the order-1 coarse problem (host
// setup, device V-cycle).  Stands in for the reference's BoomerAMG coarse
// solve (src/solvers/elliptic/MG/coarseLevel.cpp:47-222): hypre is a separate
// build with MPI, so the hierarchy here is our own, built the textbook way
// (Vanek, Mandel, Brezina 1996):
//   strength      j strong for i if |a_ij| >= theta sqrt(|a_ii a_jj|)
//   aggregation   three passes in row order (columns ascending): a node whose
//                 strong neighbours are all free forms an aggregate with them;
//                 a free node joins the pass-1 aggregate of its first strong
//                 neighbour that has one; a free node forms an aggregate with its
//                 free strong neighbours
//   prolongation  P = T - (omega D^-1) (A T), T the aggregate indicator with
//                 unit columns, omega = 4 / (3 rho(D^-1 A)), rho from power
//                 steps started at a fixed splitmix64 vector
//   coarse matrix (P^T A) P, exact zeros dropped; the last level (n <= 256, no
//                 coarsening possible, or the level limit) is inverted densely
//   V-cycle       `sweeps` damped-Jacobi sweeps (omega) from x = 0, residual,
//                 restriction by P^T, recursion, prolongation, `sweeps` sweeps:
//                 symmetric, so it preconditions CG
// Sparse products follow the row-by-row (Gustavson) order, columns ascending.
#include <algorithm>
#include <cmath>

#include "amg.hpp"

namespace nrs {

namespace {

constexpr int kBlock = 256;

struct SplitMix64 {
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

// C = A B (CSR, columns ascending, exact zeros dropped)
CsrHost spgemm(const CsrHost &A, const CsrHost &B)
{
  NRS_CHECK(A.m == B.n, "amg: product shapes");
  CsrHost C;
  C.n = A.n;
  C.m = B.m;
  C.rp.assign(A.n + 1, 0);
  std::vector<double> acc(B.m, 0.0);
  std::vector<int> mark(B.m, -1), cols;
  for (int i = 0; i < A.n; ++i) {
    cols.clear();
    for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) {
      const int k = A.ci[p];
      const double a = A.v[p];
      for (int q = B.rp[k]; q < B.rp[k + 1]; ++q) {
        const int j = B.ci[q];
        if (mark[j] != i) { mark[j] = i; acc[j] = 0.0; cols.push_back(j); }
        acc[j] += a * B.v[q];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int j : cols)
      if (acc[j] != 0.0) { C.ci.push_back(j); C.v.push_back(acc[j]); }
    C.rp[i + 1] = (int)C.ci.size();
  }
  return C;
}

CsrHost transpose(const CsrHost &A)
{
  CsrHost T;
  T.n = A.m;
  T.m = A.n;
  T.rp.assign(A.m + 1, 0);
  for (int c : A.ci) T.rp[c + 1]++;
  for (int j = 0; j < A.m; ++j) T.rp[j + 1] += T.rp[j];
  T.ci.resize(A.ci.size());
  T.v.resize(A.v.size());
  std::vector<int> pos(T.rp.begin(), T.rp.end() - 1);
  for (int i = 0; i < A.n; ++i)
    for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) {
      const int d = pos[A.ci[p]]++;
      T.ci[d] = i;
      T.v[d] = A.v[p];
    }
  return T;
}

double diag_of(const CsrHost &A, int i)
{
  for (int p = A.rp[i]; p < A.rp[i + 1]; ++p)
    if (A.ci[p] == i) return A.v[p];
  return 0.0;
}

// y = A x (host, row order)
void host_spmv(const CsrHost &A, const std::vector<double> &x, std::vector<double> &y)
{
  y.assign(A.n, 0.0);
  for (int i = 0; i < A.n; ++i) {
    double s = 0.0;
    for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) s += A.v[p] * x[A.ci[p]];
    y[i] = s;
  }
}

double rho_dinv_a(const CsrHost &A, const std::vector<double> &dinv, const AmgParams &p)
{
  const int n = A.n;
  SplitMix64 rng{p.seed};
  std::vector<double> v(n), w;
  for (auto &x : v) x = rng.next() - 0.5;
  double nv = 0.0;
  for (double x : v) nv += x * x;
  nv = std::sqrt(nv);
  for (auto &x : v) x /= nv;
  double lam = 0.0;
  for (int it = 0; it < p.powerSteps; ++it) {
    host_spmv(A, v, w);
    double s = 0.0;
    for (int i = 0; i < n; ++i) { w[i] *= dinv[i]; s += w[i] * w[i]; }
    lam = std::sqrt(s);
    if (!(lam > 0.0)) return 1.0;
    for (int i = 0; i < n; ++i) v[i] = w[i] / lam;
  }
  return lam;
}

// dense inverse of an SPD matrix by Cholesky (row-major n x n).  B given: A is
// symmetric positive semidefinite with null space span{B}; with b = B / |B|,
// A + a b b^T (a = the largest diagonal entry) is SPD, its inverse maps b to
// b / a and agrees with A^+ on the range of A, so P (A + a b b^T)^-1 P with
// P = I - b b^T is the pseudo-inverse A^+.
std::vector<double> dense_inverse(const CsrHost &A, const std::vector<double> *B)
{
  const int n = A.n;
  std::vector<double> M((size_t)n * n, 0.0), L((size_t)n * n, 0.0), inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) M[(size_t)i * n + A.ci[p]] += A.v[p];
  std::vector<double> b;   // unit null vector
  if (B) {
    double nb = 0.0, shift = 0.0;
    for (double v : *B) nb += v * v;
    nb = std::sqrt(nb);
    b.resize(n);
    for (int i = 0; i < n; ++i) b[i] = (*B)[i] / nb;
    for (int i = 0; i < n; ++i) shift = std::max(shift, M[(size_t)i * n + i]);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) M[(size_t)i * n + j] += shift * b[i] * b[j];
  }
  for (int j = 0; j < n; ++j) {
    double d = M[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    NRS_CHECK(d > 0.0, "amg: coarsest matrix not positive definite");
    const double ljj = std::sqrt(d);
    L[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double s = M[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = s / ljj;
    }
  }
  std::vector<double> e(n), y(n);
  for (int c = 0; c < n; ++c) {
    std::fill(e.begin(), e.end(), 0.0);
    e[c] = 1.0;
    for (int i = 0; i < n; ++i) {   // L y = e
      double s = e[i];
      for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * y[k];
      y[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {   // L^T x = y
      double s = y[i];
      for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * inv[(size_t)k * n + c];
      inv[(size_t)i * n + c] = s / L[(size_t)i * n + i];
    }
  }
  if (B) {   // inv <- P inv P, P = I - b b^T (rows, then columns projected)
    for (int i = 0; i < n; ++i) {
      double m = 0.0;
      for (int j = 0; j < n; ++j) m += inv[(size_t)i * n + j] * b[j];
      for (int j = 0; j < n; ++j) inv[(size_t)i * n + j] -= m * b[j];
    }
    for (int j = 0; j < n; ++j) {
      double m = 0.0;
      for (int i = 0; i < n; ++i) m += b[i] * inv[(size_t)i * n + j];
      for (int i = 0; i < n; ++i) inv[(size_t)i * n + j] -= b[i] * m;
    }
  }
  return inv;
}

// LANES lanes per row: 8 for the sparse finest matrices, a whole wavefront
// for the dense rows the Galerkin products produce on coarse levels
template <int MODE, int LANES>
__global__ void __launch_bounds__(kBlock)
csr_kernel(int n, const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
           const double *__restrict__ x, double *__restrict__ y, const double *__restrict__ b,
           const double *__restrict__ dinv, double omega)
{
  const int lane = threadIdx.x % LANES;
  const int row = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / LANES);
  double sum = 0.0;
  if (row < n)
    for (int k = rp[row] + lane; k < rp[row + 1]; k += LANES) sum += v[k] * x[ci[k]];
#pragma unroll
  for (int o = LANES / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, LANES);
  if (row >= n || lane != 0) return;
  if (MODE == SPMV_Y) y[row] = sum;
  else if (MODE == SPMV_RES) y[row] = b[row] - sum;
  else if (MODE == SPMV_ADD) y[row] += sum;
  else y[row] = x[row] + omega * dinv[row] * (b[row] - sum);
}

__global__ void jacobi0_kernel(int n, const double *__restrict__ b, const double *__restrict__ dinv,
                               double omega, double *__restrict__ x)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = omega * dinv[i] * b[i];
}

// x = inv b, one wave per row: the lanes read the row of the symmetric inverse
// (coalesced), each sums every 64th product, then a fixed butterfly adds the
// lanes' partial sums (deterministic; rows spread over the whole GPU instead of
// one thread walking a row)
__global__ void __launch_bounds__(256) dense_kernel(int n, const double *__restrict__ inv,
                                                    const double *__restrict__ b, double *__restrict__ x)
{
  const int i = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const double *row = inv + (size_t)i * n;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += row[j] * b[j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) x[i] = s;
}

__global__ void copy_kernel(int n, const double *__restrict__ a, double *__restrict__ b)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

template <typename T>
__global__ void gather_kernel(int n, const int *__restrict__ first, const T *__restrict__ e,
                              double *__restrict__ u)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] = (double)e[first[i]];
}

template <typename T>
__global__ void scatter_kernel(int64_t L, const int *__restrict__ rowOf, const double *__restrict__ u,
                               T *__restrict__ e)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= L) return;
  const int r = rowOf[i];
  e[i] = (r >= 0) ? (T)u[r] : T(0);
}

unsigned blocks(int64_t n, int per) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

}  // namespace

template <int LANES>
static void csr_launch(int mode, int n, const int *rp, const int *ci, const double *v, const double *x,
                       double *y, const double *b, const double *dinv, double omega, hipStream_t s)
{
  const unsigned g = blocks((int64_t)n * LANES, kBlock);
  switch (mode) {
    case SPMV_Y: csr_kernel<SPMV_Y, LANES><<<g, kBlock, 0, s>>>(n, rp, ci, v, x, y, b, dinv, omega); break;
    case SPMV_RES: csr_kernel<SPMV_RES, LANES><<<g, kBlock, 0, s>>>(n, rp, ci, v, x, y, b, dinv, omega); break;
    case SPMV_ADD: csr_kernel<SPMV_ADD, LANES><<<g, kBlock, 0, s>>>(n, rp, ci, v, x, y, b, dinv, omega); break;
    default: csr_kernel<SPMV_JAC, LANES><<<g, kBlock, 0, s>>>(n, rp, ci, v, x, y, b, dinv, omega); break;
  }
  NRS_LAUNCH_CHECK();
}

void csr_spmv(int mode, int n, const int *rp, const int *ci, const double *v, const double *x, double *y,
              const double *b, const double *dinv, double omega, hipStream_t s, int lanes)
{
  if (n <= 0) return;
  if (lanes >= 64) csr_launch<64>(mode, n, rp, ci, v, x, y, b, dinv, omega, s);
  else csr_launch<8>(mode, n, rp, ci, v, x, y, b, dinv, omega, s);
}

int csr_lanes(const CsrHost &A) { return (A.n > 0 && A.ci.size() >= (size_t)48 * A.n) ? 64 : 8; }

template <typename T>
void rows_gather(int n, const int *first, const T *e, double *u, hipStream_t s)
{
  if (n <= 0) return;
  gather_kernel<T><<<blocks(n, kBlock), kBlock, 0, s>>>(n, first, e, u);
  NRS_LAUNCH_CHECK();
}
template <typename T>
void rows_scatter(int64_t L, const int *rowOf, const double *u, T *e, hipStream_t s)
{
  if (L <= 0) return;
  scatter_kernel<T><<<blocks(L, kBlock), kBlock, 0, s>>>(L, rowOf, u, e);
  NRS_LAUNCH_CHECK();
}
template void rows_gather<double>(int, const int *, const double *, double *, hipStream_t);
template void rows_gather<float>(int, const int *, const float *, double *, hipStream_t);
template void rows_scatter<double>(int64_t, const int *, const double *, double *, hipStream_t);
template void rows_scatter<float>(int64_t, const int *, const double *, float *, hipStream_t);

std::vector<AmgHostLevel> amg_build(const CsrHost &A0, const AmgParams &p)
{
  NRS_CHECK(A0.n == A0.m && A0.n > 0, "amg: square nonempty matrix required");
  std::vector<AmgHostLevel> out;
  CsrHost A = A0;
  // null vector carried down the levels (singular problems): T B_c = B, so the
  // Galerkin matrices keep B as null vector; without a null space B = 1 on
  // every level (unit aggregate indicators)
  std::vector<double> B(A0.n, 1.0);
  for (;;) {
    AmgHostLevel h;
    h.A = A;
    const int n = A.n;
    h.dinv.resize(n);
    for (int i = 0; i < n; ++i) {
      const double d = diag_of(A, i);
      NRS_CHECK(d > 0.0, "amg: nonpositive diagonal");
      h.dinv[i] = 1.0 / d;
    }
    if (n <= p.coarsest || (int)out.size() + 1 == p.maxLevels) {
      h.inv = dense_inverse(A, p.nullSpace ? &B : nullptr);
      out.push_back(std::move(h));
      break;
    }
    h.omega = 4.0 / (3.0 * rho_dinv_a(A, h.dinv, p));
    // strength: |a_ij| >= theta sqrt(|a_ii a_jj|), i != j
    std::vector<double> ad(n);
    for (int i = 0; i < n; ++i) ad[i] = std::fabs(1.0 / h.dinv[i]);
    std::vector<int> srp(n + 1, 0), sci;
    for (int i = 0; i < n; ++i) {
      for (int q = A.rp[i]; q < A.rp[i + 1]; ++q) {
        const int j = A.ci[q];
        if (j != i && std::fabs(A.v[q]) >= p.theta * std::sqrt(ad[i] * ad[j])) sci.push_back(j);
      }
      srp[i + 1] = (int)sci.size();
    }
    // aggregation
    std::vector<int> agg(n, -1);
    int na = 0;
    for (int i = 0; i < n; ++i) {
      if (agg[i] >= 0) continue;
      bool free = true;
      for (int q = srp[i]; q < srp[i + 1] && free; ++q) free = agg[sci[q]] < 0;
      if (!free) continue;
      agg[i] = na;
      for (int q = srp[i]; q < srp[i + 1]; ++q) agg[sci[q]] = na;
      ++na;
    }
    const std::vector<int> first = agg;
    for (int i = 0; i < n; ++i) {
      if (agg[i] >= 0) continue;
      for (int q = srp[i]; q < srp[i + 1]; ++q)
        if (first[sci[q]] >= 0) { agg[i] = first[sci[q]]; break; }
    }
    for (int i = 0; i < n; ++i) {
      if (agg[i] >= 0) continue;
      agg[i] = na;
      for (int q = srp[i]; q < srp[i + 1]; ++q)
        if (agg[sci[q]] < 0) agg[sci[q]] = na;
      ++na;
    }
    if (na >= n) {   // nothing to coarsen: solve this level directly
      h.omega = 0.0;
      h.inv = dense_inverse(A, p.nullSpace ? &B : nullptr);
      NRS_CHECK(n <= 8192, "amg: level cannot be coarsened and is too large to invert");
      out.push_back(std::move(h));
      break;
    }
    // tentative prolongation T (unit columns), then P = T - (omega D^-1) (A T)
    std::vector<double> norm(na, 0.0);   // |B on each aggregate| (B = 1: sqrt of its size)
    for (int i = 0; i < n; ++i) norm[agg[i]] += B[i] * B[i];
    for (auto &v : norm) v = std::sqrt(v);
    CsrHost T;
    T.n = n;
    T.m = na;
    T.rp.resize(n + 1);
    T.ci.resize(n);
    T.v.resize(n);
    for (int i = 0; i < n; ++i) {
      T.rp[i] = i;
      T.ci[i] = agg[i];
      T.v[i] = B[i] / norm[agg[i]];
    }
    T.rp[n] = n;
    const CsrHost AT = spgemm(A, T);
    CsrHost P;
    P.n = n;
    P.m = na;
    P.rp.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
      const double sc = h.omega * h.dinv[i];
      int q = AT.rp[i];
      const int qe = AT.rp[i + 1];
      bool tdone = false;
      const int tc = T.ci[i];
      // merge row i of T (one entry) with row i of sc * AT, columns ascending
      while (q < qe || !tdone) {
        int c;
        double val;
        if (!tdone && (q >= qe || tc <= AT.ci[q])) {
          c = tc;
          val = T.v[i];
          if (q < qe && AT.ci[q] == tc) { val = T.v[i] - sc * AT.v[q]; ++q; }
          tdone = true;
        } else {
          c = AT.ci[q];
          val = -(sc * AT.v[q]);
          ++q;
        }
        if (val != 0.0) { P.ci.push_back(c); P.v.push_back(val); }
      }
      P.rp[i + 1] = (int)P.ci.size();
    }
    h.P = P;
    h.R = transpose(P);
    A = spgemm(spgemm(h.R, A), P);
    out.push_back(std::move(h));
    if (p.nullSpace) B = norm;
    else B.assign(na, 1.0);
  }
  return out;
}

void Amg::setup(const CsrHost &A, const AmgParams &p)
{
  params = p;
  sweeps = p.sweeps;
  const auto host = amg_build(A, p);
  lv.clear();
  lv.resize(host.size());
  for (size_t k = 0; k < host.size(); ++k) {
    const AmgHostLevel &h = host[k];
    AmgDevLevel &d = lv[k];
    d.n = h.A.n;
    d.omega = h.omega;
    d.dense = !h.inv.empty();
    d.arp.upload(h.A.rp); d.aci.upload(h.A.ci); d.av.upload(h.A.v);
    d.laneA = csr_lanes(h.A);
    d.dinv.upload(h.dinv);
    if (d.dense) {
      d.inv.upload(h.inv);
    } else {
      d.nc = h.P.m;
      d.prp.upload(h.P.rp); d.pci.upload(h.P.ci); d.pv.upload(h.P.v);
      d.rrp.upload(h.R.rp); d.rci.upload(h.R.ci); d.rv.upload(h.R.v);
      d.laneP = csr_lanes(h.P);
      d.laneR = csr_lanes(h.R);
    }
    if (k > 0) { d.b.alloc(d.n); d.xa.alloc(d.n); d.xb.alloc(d.n); }
    d.r.alloc(d.n);
  }
}

// One V-cycle: returns the buffer holding the result (xa or xb).
static double *vcycle(std::vector<AmgDevLevel> &lv, int sweeps, size_t k, const double *b, double *xa,
                      double *xb, hipStream_t s)
{
  AmgDevLevel &l = lv[k];
  if (l.dense) {
    dense_kernel<<<(unsigned)((l.n + 3) / 4), 256, 0, s>>>(l.n, l.inv.get(), b, xa);
    NRS_LAUNCH_CHECK();
    return xa;
  }
  jacobi0_kernel<<<blocks(l.n, kBlock), kBlock, 0, s>>>(l.n, b, l.dinv.get(), l.omega, xa);
  NRS_LAUNCH_CHECK();
  double *x = xa, *t = xb;
  for (int sw = 1; sw < sweeps; ++sw) {
    csr_spmv(SPMV_JAC, l.n, l.arp.get(), l.aci.get(), l.av.get(), x, t, b, l.dinv.get(), l.omega, s, l.laneA);
    std::swap(x, t);
  }
  csr_spmv(SPMV_RES, l.n, l.arp.get(), l.aci.get(), l.av.get(), x, l.r.get(), b, nullptr, 0.0, s, l.laneA);
  AmgDevLevel &c = lv[k + 1];
  csr_spmv(SPMV_Y, l.nc, l.rrp.get(), l.rci.get(), l.rv.get(), l.r.get(), c.b.get(), nullptr, nullptr, 0.0, s,
           l.laneR);
  const double *xc = vcycle(lv, sweeps, k + 1, c.b.get(), c.xa.get(), c.xb.get(), s);
  csr_spmv(SPMV_ADD, l.n, l.prp.get(), l.pci.get(), l.pv.get(), xc, x, nullptr, nullptr, 0.0, s, l.laneP);
  for (int sw = 0; sw < sweeps; ++sw) {
    csr_spmv(SPMV_JAC, l.n, l.arp.get(), l.aci.get(), l.av.get(), x, t, b, l.dinv.get(), l.omega, s, l.laneA);
    std::swap(x, t);
  }
  return x;
}

void Amg::apply(const double *b, double *z, double *tmp, hipStream_t s)
{
  NRS_CHECK(!lv.empty(), "amg: not set up");
  const double *res = vcycle(lv, sweeps, 0, b, z, tmp, s);
  if (res != z) {
    copy_kernel<<<blocks(lv[0].n, kBlock), kBlock, 0, s>>>(lv[0].n, res, z);
    NRS_LAUNCH_CHECK();
  }
}

}  // namespace nrs
