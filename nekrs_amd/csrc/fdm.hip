// fdm.hip -- overlapping Schwarz smoother (ASM / RAS) with fast diagonalisation.
//
// Reference behaviour followed (src/solvers/elliptic/MG/ellipticMultiGridSchwarz.cpp):
//   element lengths         :58-276  (harmonic mean over face-interior lines,
//                                     neighbour lengths, fallback to the middle one)
//   boundary flags          :278-298 (faces 4,2 / 1,3 / 0,5 for r / s / t)
//   1D stiffness and mass   :308-427 on N+3 points (own element plus one node of
//                                     each neighbour)
//   eigenproblem, zeroing   :455-556 (dsygv; here a Jacobi eigensolver on
//                                     B^-1/2 A B^-1/2, B being diagonal)
//   operators               :558-641 (S per direction, invL = 1 / (lx + ly + lz))
//   extended-mesh mask      :643-743 (edges and Dirichlet faces of the N+2 mesh)
//   weights                 :798-835
//   application             :1056-1156 with kernels/elliptic/preFDM.c,
//                                     fusedFDM.c and postFDM.c
//
// The reference builds an order N+2 mesh and its gather-scatter to move one
// layer of values between face neighbours.  That exchange only ever pairs an
// element's ghost face layer with the layer next to the shared face in the
// neighbour, so here each element reads that layer straight from its
// neighbour through precomputed face links: one kernel fills the extended box,
// runs the six tensor contractions out of LDS and restricts (RAS), or writes
// its overlap results for a second small kernel to exchange (ASM).  The
// arithmetic of the exchange is kept: a ghost value is (own + neighbour) - own,
// which is what gather-scatter-then-subtract computes.
#include <algorithm>
#include <array>
#include <cmath>
#include <thread>

#include "fdm.hpp"
#include "kernels.hpp"
#include "linalg.hpp"
#include "mesh.hpp"

namespace nrs {

namespace {

// regular node index of face point (u, v) at depth d from face f
__host__ __device__ inline int face_node(int f, int u, int v, int d, int N)
{
  const int Nq = N + 1;
  int i, j, k;
  switch (f) {
    case 0: i = u; j = v; k = d; break;
    case 5: i = u; j = v; k = N - d; break;
    case 1: i = u; k = v; j = d; break;
    case 3: i = u; k = v; j = N - d; break;
    case 4: j = u; k = v; i = d; break;
    default: j = u; k = v; i = N - d; break;   // face 2
  }
  return i + Nq * (j + Nq * k);
}

__host__ __device__ inline void link_map(int code, int N, int u, int v, int &up, int &vp)
{
  const int p = (code & 8) ? v : u, q = (code & 8) ? u : v;
  up = (code & 16) ? N - p : p;
  vp = (code & 32) ? N - q : q;
}

inline int face_dir(int f) { return (f == 4 || f == 2) ? 0 : (f == 1 || f == 3) ? 1 : 2; }

// Cyclic Jacobi eigenvalue iteration for a symmetric n x n matrix (row-major,
// overwritten): eigenvalues in d, eigenvectors in the columns of v.
void jacobi_eig(int n, double *a, double *d, double *v)
{
  for (int i = 0; i < n * n; ++i) v[i] = 0.0;
  for (int i = 0; i < n; ++i) v[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < n; ++p) {
      diag += a[p * n + p] * a[p * n + p];
      for (int q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
    }
    if (off <= 1e-34 * diag || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if (apq == 0.0) continue;
        const double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {   // A <- A J
          const double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {   // A <- J^T A
          const double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {   // V <- V J
          const double vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = c * vkp - s * vkq;
          v[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) d[i] = a[i * n + i];
}

template <typename F> void parallel_for(int64_t n, F &&fn)
{
  const int nt = (int)std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  if (n < 64 || nt == 1) {
    for (int64_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (int64_t i = t; i < n; i += nt) fn(i);
    });
  for (auto &x : th) x.join();
}

}  // namespace

// ------------------------------------------------------------------ host setup

std::vector<FaceLink> face_links(int N, int E, const long long *gid)
{
  const int Np = (N + 1) * (N + 1) * (N + 1);
  struct Key {
    std::array<long long, 4> g;
    int e, f;
  };
  std::vector<Key> keys((size_t)E * 6);
  for (int e = 0; e < E; ++e)
    for (int f = 0; f < 6; ++f) {
      const long long *ge = gid + (int64_t)e * Np;
      std::array<long long, 4> c = {ge[face_node(f, 0, 0, 0, N)], ge[face_node(f, N, 0, 0, N)],
                                    ge[face_node(f, 0, N, 0, N)], ge[face_node(f, N, N, 0, N)]};
      std::sort(c.begin(), c.end());
      keys[(size_t)e * 6 + f] = {c, e, f};
    }
  std::sort(keys.begin(), keys.end(), [](const Key &x, const Key &y) {
    if (x.g != y.g) return x.g < y.g;
    return x.e * 6 + x.f < y.e * 6 + y.f;
  });
  std::vector<FaceLink> links((size_t)E * 6, FaceLink{-1, 0});
  auto find_code = [&](int ea, int fa, int eb, int fb) {
    const long long *ga = gid + (int64_t)ea * Np, *gb = gid + (int64_t)eb * Np;
    for (int t = 0; t < 8; ++t) {
      const int code = fb | (t << 3);
      bool ok = true;
      for (int v = 0; v <= N && ok; ++v)
        for (int u = 0; u <= N && ok; ++u) {
          int up, vp;
          link_map(code, N, u, v, up, vp);
          ok = ga[face_node(fa, u, v, 0, N)] == gb[face_node(fb, up, vp, 0, N)];
        }
      if (ok) return code;
    }
    throw Error(NRS_ERR_INVALID, "schwarz: faces share corners but not their nodes");
  };
  for (size_t s = 0; s < keys.size();) {
    size_t t = s + 1;
    while (t < keys.size() && keys[t].g == keys[s].g) ++t;
    NRS_CHECK(t - s <= 2, "schwarz: a face is shared by more than two elements");
    if (t - s == 2) {
      const Key &A = keys[s], &B = keys[s + 1];
      NRS_CHECK(A.e != B.e, "schwarz: an element is its own face neighbour");
      links[(size_t)A.e * 6 + A.f] = {B.e, find_code(A.e, A.f, B.e, B.f)};
      links[(size_t)B.e * 6 + B.f] = {A.e, find_code(B.e, B.f, A.e, A.f)};
    }
    s = t;
  }
  return links;
}

std::vector<double> element_lengths(int N, int E, const double *x, const double *y,
                                    const double *z, const std::vector<FaceLink> &links)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  std::vector<double> r, w;
  gll_nodes(N, r, w);
  std::vector<double> out((size_t)9 * E);
  double *left = out.data(), *mid = left + 3 * (size_t)E, *right = mid + 3 * (size_t)E;
  const int lo = (Nq == 2) ? 0 : 1, hi = (Nq == 2) ? Nq : Nq - 1;
  auto at = [&](int i, int j, int k) { return i + Nq * (j + Nq * k); };
  for (int64_t e = 0; e < E; ++e) {
    const double *X = x + e * Np, *Y = y + e * Np, *Z = z + e * Np;
    // harmonic_mean_element_length (:58-131): lines between opposite faces
    for (int d = 0; d < 3; ++d) {
      double acc = 0.0, wsum = 0.0;
      for (int b = lo; b < hi; ++b)
        for (int a = lo; a < hi; ++a) {
          int n1, n2;
          if (d == 0) { n1 = at(0, a, b); n2 = at(Nq - 1, a, b); }
          else if (d == 1) { n1 = at(a, 0, b); n2 = at(a, Nq - 1, b); }
          else { n1 = at(a, b, 0); n2 = at(a, b, Nq - 1); }
          const double wt = (Nq == 2) ? 1.0 : w[a - 1] * w[b - 1];
          const double dx = X[n2] - X[n1], dy = Y[n2] - Y[n1], dz = Z[n2] - Z[n1];
          acc += wt / (dx * dx + dy * dy + dz * dz);
          wsum += wt;
        }
      mid[d * E + e] = 1.0 / std::sqrt(acc / wsum);
    }
  }
  for (int64_t n = 0; n < 3 * (int64_t)E; ++n) {
    const double m = mid[n];
    NRS_CHECK(std::isfinite(m) && std::fabs(m) >= 1e-12 && m >= -1e-12,
              "schwarz: element with zero, negative or invalid length");
  }
  if (Nq == 2) {
    std::copy(mid, mid + 3 * (size_t)E, left);
    std::copy(mid, mid + 3 * (size_t)E, right);
    return out;
  }
  // neighbour lengths (:231-275): the gather-scatter sum on the shared face,
  // minus the element's own value, falling back to the middle length
  static const int lf[3] = {4, 1, 0}, rf[3] = {2, 3, 5};
  for (int64_t e = 0; e < E; ++e)
    for (int d = 0; d < 3; ++d)
      for (int side = 0; side < 2; ++side) {
        const int f = side ? rf[d] : lf[d];
        const FaceLink &lk = links[e * 6 + f];
        const double own = mid[d * E + e];
        double len = 0.0;
        if (lk.nbr >= 0) len = (own + mid[face_dir(lk.code & 7) * E + lk.nbr]) - own;
        if (std::fabs(len) < 1e-12 || len < -1e-12) len = own;
        (side ? right : left)[d * E + e] = len;
      }
  return out;
}

bool fdm_1d(int N, const double *D, const double *w, int lbc, int rbc, double ll, double lm,
            double lr, double *S, double *lam)
{
  const int n = N, nq = N + 1, nl = N + 3;
  // ah = D^T W D (:318-338)
  std::vector<double> tmp(nq * nq), ah(nq * nq), a(nl * nl, 0.0), b(nl, 0.0);
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < nq; ++j) tmp[i * nq + j] = D[i * nq + j] * w[i];
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < nq; ++j) {
      double s = 0.0;
      for (int k = 0; k < nq; ++k) s += D[k * nq + i] * tmp[k * nq + j];
      ah[i + j * nq] = s;
    }
  auto AH = [&](int i, int j) { return ah[i + j * nq]; };
  auto A = [&](int i, int j) -> double & { return a[i * nl + j]; };
  const int i0 = (lbc == 1) ? 1 : 0, i1 = (rbc == 1) ? n - 1 : n;
  // stiffness (:339-377)
  A(1, 1) = 1.0;
  A(n + 1, n + 1) = 1.0;
  for (int j = i0; j <= i1; ++j)
    for (int i = i0; i <= i1; ++i) A(i + 1, j + 1) = (2.0 / lm) * AH(i, j);
  if (lbc == 0) {
    const double f = 2.0 / ll;
    A(0, 0) = f * AH(n - 1, n - 1);
    A(1, 0) = f * AH(n, n - 1);
    A(0, 1) = f * AH(n - 1, n);
    A(1, 1) += f * AH(n, n);
  } else {
    A(0, 0) = 1.0;
  }
  if (rbc == 0) {
    const double f = 2.0 / lr;
    A(n + 1, n + 1) += f * AH(0, 0);
    A(n + 2, n + 1) = f * AH(1, 0);
    A(n + 1, n + 2) = f * AH(0, 1);
    A(n + 2, n + 2) = f * AH(1, 1);
  } else {
    A(n + 2, n + 2) = 1.0;
  }
  // mass, diagonal (:380-427)
  b[1] = 1.0;
  b[n + 1] = 1.0;
  for (int i = i0; i <= i1; ++i) b[i + 1] = 0.5 * lm * w[i];
  if (lbc == 0) {
    b[0] = 0.5 * ll * w[n - 1];
    b[1] += 0.5 * ll * w[n];
  } else {
    b[0] = 1.0;
  }
  if (rbc == 0) {
    b[n + 1] += 0.5 * lr * w[0];
    b[n + 2] = 0.5 * lr * w[1];
  } else {
    b[n + 2] = 1.0;
  }
  for (int i = 0; i < nl; ++i)
    if (!(b[i] > 0.0)) return false;
  // A z = lam B z  <=>  (B^-1/2 A B^-1/2) y = lam y, z = B^-1/2 y.  dsygv reads
  // the upper triangle (UPLO='U'), so the matrix is symmetrised from it.
  std::vector<double> c(nl * nl), v(nl * nl), d(nl);
  for (int i = 0; i < nl; ++i)
    for (int j = i; j < nl; ++j) {
      const double s = A(i, j) / std::sqrt(b[i] * b[j]);
      c[i * nl + j] = c[j * nl + i] = s;
    }
  jacobi_eig(nl, c.data(), d.data(), v.data());
  std::vector<int> ord(nl);
  for (int i = 0; i < nl; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int m = 0; m < nl; ++m) {
    lam[m] = d[ord[m]];
    for (int i = 0; i < nl; ++i) S[i * nl + m] = v[i * nl + ord[m]] / std::sqrt(b[i]);
  }
  // row zeroing (:547-554)
  auto zero_row = [&](int r) { for (int m = 0; m < nl; ++m) S[r * nl + m] = 0.0; };
  if (lbc > 0) zero_row(0);
  if (lbc == 1) zero_row(1);
  if (rbc > 0) zero_row(nl - 1);
  if (rbc == 1) zero_row(nl - 2);
  return true;
}

// ------------------------------------------------------------------ kernels

namespace {

template <int NQ> struct FdmCfg {
  static constexpr int NL = NQ + 2;                           // extended box edge
  static constexpr int NLP = (NL % 2 == 0) ? NL + 1 : NL;     // odd row pitch in LDS
  static constexpr int NP = NQ * NQ * NQ;
  static constexpr int WPE = (NL * NL + 63) / 64;             // whole waves per element
  static constexpr int TPE = 64 * WPE;
  static constexpr int EPB = (WPE >= 4) ? 1 : 4 / WPE;        // elements per block
  static constexpr int BLOCK = TPE * EPB;
  static constexpr int BOX = NL * NL * NLP;
};

// One element per TPE threads (whole waves, so the element's operators are
// read with wave-uniform loads).  Thread (a, b) owns one pencil of the box in
// each of the six contraction phases; the box lives in two LDS buffers.
//   RAS: out = restricted result * w (w = invDegree)
//   ASM: out = restricted result, gout = overlap (ghost face) results
template <typename T, int NQ, bool ASM>
__global__ __launch_bounds__(FdmCfg<NQ>::BLOCK) void fdm_kernel(
    int E, const T *__restrict__ u, const T *__restrict__ S, const double *__restrict__ lam,
    const int2 *__restrict__ links, const T *__restrict__ w, T *__restrict__ out, T *__restrict__ gout)
{
  using C = FdmCfg<NQ>;
  constexpr int NL = C::NL, NLP = C::NLP, NP = C::NP, NQ2 = NQ * NQ, TPE = C::TPE;
  __shared__ T lds[C::EPB][2][C::BOX];
  const int slot = threadIdx.x / TPE;
  const int t = threadIdx.x % TPE;
  const int e = __builtin_amdgcn_readfirstlane(blockIdx.x * C::EPB + slot);
  const bool live = e < E;
  T *X = lds[slot][0], *Y = lds[slot][1];
  auto at = [](int k, int j, int i) { return (k * NL + j) * NLP + i; };

  // 1. box interior from u (preFDM.c:16-37); edges and corners are zero
  if (live) {
    const T *ue = u + (int64_t)e * NP;
    for (int n = t; n < NP; n += TPE) {
      const int i = n % NQ, j = (n / NQ) % NQ, k = n / NQ2;
      X[at(k + 1, j + 1, i + 1)] = ue[n];
    }
    for (int n = t; n < 12 * NL; n += TPE) {
      const int line = n / NL, s = n % NL;
      const int c0 = (line & 1) ? NL - 1 : 0, c1 = (line & 2) ? NL - 1 : 0;
      const int dir = line >> 2;
      if (dir == 0) X[at(c1, c0, s)] = T(0);
      else if (dir == 1) X[at(c1, s, c0)] = T(0);
      else X[at(s, c1, c0)] = T(0);
    }
  }
  __syncthreads();
  // 2. ghost faces: (own second layer + neighbour's second layer) - own
  //    (preFDM.c:40-96, extended gather-scatter, fusedFDM.c:47-104)
  if (live) {
    for (int n = t; n < 6 * NQ2; n += TPE) {
      const int f = n / NQ2, uu = n % NQ, vv = (n / NQ) % NQ;
      int g, o;
      switch (f) {
        case 0: g = at(0, vv + 1, uu + 1); o = at(2, vv + 1, uu + 1); break;
        case 5: g = at(NL - 1, vv + 1, uu + 1); o = at(NL - 3, vv + 1, uu + 1); break;
        case 1: g = at(vv + 1, 0, uu + 1); o = at(vv + 1, 2, uu + 1); break;
        case 3: g = at(vv + 1, NL - 1, uu + 1); o = at(vv + 1, NL - 3, uu + 1); break;
        case 4: g = at(vv + 1, uu + 1, 0); o = at(vv + 1, uu + 1, 2); break;
        default: g = at(vv + 1, uu + 1, NL - 1); o = at(vv + 1, uu + 1, NL - 3); break;
      }
      const T own = X[o];
      T val = T(0);
      const int2 lk = links[(int64_t)e * 6 + f];
      if (lk.x >= 0) {
        int up, vp;
        link_map(lk.y, NQ - 1, uu, vv, up, vp);
        const T nb = u[(int64_t)lk.x * NP + face_node(lk.y & 7, up, vp, 1, NQ - 1)];
        val = (own + nb) - own;
      }
      X[g] = val;
    }
  }
  __syncthreads();

  // 3. u <- (Sz x Sy x Sx) invL (Sz x Sy x Sx)^T u  (fusedFDM.c:121-175)
  const int a = t / NL, b = t % NL;
  const bool act = live && t < NL * NL;
  const T *Sx = S + (int64_t)e * 3 * NL * NL, *Sy = Sx + NL * NL, *Sz = Sy + NL * NL;
  T r[NL];
  if (act) {   // x forward: Y[k][j][m] = sum_l Sx[l][m] X[k][j][l]
#pragma unroll
    for (int l = 0; l < NL; ++l) r[l] = X[at(a, b, l)];
#pragma unroll
    for (int m = 0; m < NL; ++m) {
      T s = T(0);
#pragma unroll
      for (int l = 0; l < NL; ++l) s += Sx[l * NL + m] * r[l];
      Y[at(a, b, m)] = s;
    }
  }
  __syncthreads();
  if (act) {   // y forward: X[k][n][m] = sum_l Sy[l][n] Y[k][l][m]
#pragma unroll
    for (int l = 0; l < NL; ++l) r[l] = Y[at(a, l, b)];
#pragma unroll
    for (int n = 0; n < NL; ++n) {
      T s = T(0);
#pragma unroll
      for (int l = 0; l < NL; ++l) s += Sy[l * NL + n] * r[l];
      X[at(a, n, b)] = s;
    }
  }
  __syncthreads();
  if (act) {   // z forward, eigenvalue scaling: Y[p][n][m] = invL * sum_l Sz[l][p] X[l][n][m]
    const double *le = lam + (int64_t)e * 3 * NL;
    const double lxy = le[b] + le[NL + a];
#pragma unroll
    for (int l = 0; l < NL; ++l) r[l] = X[at(l, a, b)];
#pragma unroll
    for (int p = 0; p < NL; ++p) {
      T s = T(0);
#pragma unroll
      for (int l = 0; l < NL; ++l) s += Sz[l * NL + p] * r[l];
      const double dg = lxy + le[2 * NL + p];
      const T inv = (dg > 1e-5) ? T(1.0 / dg) : T(0);
      Y[at(p, a, b)] = s * inv;
    }
  }
  __syncthreads();
  if (act) {   // x back: X[p][n][i] = sum_m Sx[i][m] Y[p][n][m]
#pragma unroll
    for (int m = 0; m < NL; ++m) r[m] = Y[at(a, b, m)];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      T s = T(0);
#pragma unroll
      for (int m = 0; m < NL; ++m) s += Sx[i * NL + m] * r[m];
      X[at(a, b, i)] = s;
    }
  }
  __syncthreads();
  if (act) {   // y back: Y[p][j][i] = sum_n Sy[j][n] X[p][n][i]
#pragma unroll
    for (int n = 0; n < NL; ++n) r[n] = X[at(a, n, b)];
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      T s = T(0);
#pragma unroll
      for (int n = 0; n < NL; ++n) s += Sy[j * NL + n] * r[n];
      Y[at(a, j, b)] = s;
    }
  }
  __syncthreads();
  if (!act) return;
  // z back into registers: o[k] = sum_p Sz[k][p] Y[p][j][i], pencil (j, i) = (a, b)
  T o[NL];
#pragma unroll
  for (int p = 0; p < NL; ++p) r[p] = Y[at(p, a, b)];
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    T s = T(0);
#pragma unroll
    for (int p = 0; p < NL; ++p) s += Sz[k * NL + p] * r[p];
    o[k] = s;
  }
  const bool ai = a >= 1 && a <= NL - 2, bi = b >= 1 && b <= NL - 2;
  const int64_t eb = (int64_t)e * NP;
  if (ai && bi) {
#pragma unroll
    for (int k = 1; k <= NL - 2; ++k) {
      const int64_t idx = eb + ((k - 1) * NQ + (a - 1)) * NQ + (b - 1);
      out[idx] = ASM ? o[k] : o[k] * w[idx];
    }
  }
  if (!ASM) return;
  T *ge = gout + (int64_t)e * 6 * NQ2;
  if (ai && bi) {   // z faces 0 / 5: (u, v) = (i, j)
    ge[0 * NQ2 + (a - 1) * NQ + (b - 1)] = o[0];
    ge[5 * NQ2 + (a - 1) * NQ + (b - 1)] = o[NL - 1];
  }
  if (ai && (b == 0 || b == NL - 1)) {   // x faces 4 / 2: (u, v) = (j, k)
    T *gf = ge + (b == 0 ? 4 : 2) * NQ2;
#pragma unroll
    for (int k = 1; k <= NL - 2; ++k) gf[(k - 1) * NQ + (a - 1)] = o[k];
  }
  if (bi && (a == 0 || a == NL - 1)) {   // y faces 1 / 3: (u, v) = (i, k)
    T *gf = ge + (a == 0 ? 1 : 3) * NQ2;
#pragma unroll
    for (int k = 1; k <= NL - 2; ++k) gf[(k - 1) * NQ + (b - 1)] = o[k];
  }
}

// ASM overlap exchange (extended gather-scatter + postFDM.c:3-155): each node
// on the second layer from a face with a neighbour gets (own + neighbour) -
// own of the overlap results there, in the order k faces, j faces, i faces;
// then the Schwarz weight.
template <typename T, int NQ>
__global__ void fdm_post_kernel(int64_t n, const T *__restrict__ box, const T *__restrict__ gout,
                                const int2 *__restrict__ links, const T *__restrict__ wts,
                                T *__restrict__ Su)
{
  constexpr int NP = NQ * NQ * NQ, NQ2 = NQ * NQ;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = idx / NP;
    const int r = (int)(idx - e * NP);
    const int i = r % NQ, j = (r / NQ) % NQ, k = r / NQ2;
    T v = box[idx];
    auto add = [&](int f, int uu, int vv) {
      const int2 lk = links[e * 6 + f];
      if (lk.x < 0) return;
      int up, vp;
      link_map(lk.y, NQ - 1, uu, vv, up, vp);
      const T own = gout[(e * 6 + f) * NQ2 + vv * NQ + uu];
      const T nb = gout[((int64_t)lk.x * 6 + (lk.y & 7)) * NQ2 + vp * NQ + up];
      v += (own + nb) - own;
    };
    if (k == 1) add(0, i, j);
    if (k == NQ - 2) add(5, i, j);
    if (j == 1) add(1, i, k);
    if (j == NQ - 2) add(3, i, k);
    if (i == 1) add(4, j, k);
    if (i == NQ - 2) add(2, j, k);
    Su[idx] = v * wts[idx];
  }
}

// overlap counts behind the ASM weights (generate_weights, :798-835 before
// its gather-scatter): 1 + number of faces with a neighbour whose second
// layer holds the node
template <typename T>
__global__ void asm_count_kernel(int N, int64_t n, const int2 *__restrict__ links, T *__restrict__ c)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = idx / Np;
    const int r = (int)(idx - e * Np);
    const int i = r % Nq, j = (r / Nq) % Nq, k = r / (Nq * Nq);
    int cnt = 1;
    const int2 *lk = links + e * 6;
    cnt += (k == 1 && lk[0].x >= 0) + (k == N - 1 && lk[5].x >= 0);
    cnt += (j == 1 && lk[1].x >= 0) + (j == N - 1 && lk[3].x >= 0);
    cnt += (i == 1 && lk[4].x >= 0) + (i == N - 1 && lk[2].x >= 0);
    c[idx] = T(cnt);
  }
}

template <typename T, int NQ>
void launch_fdm_nq(bool ras, int E, const T *u, const T *S, const double *lam, const int2 *links,
                   const T *w, T *out, T *gout, hipStream_t s)
{
  using C = FdmCfg<NQ>;
  const unsigned grid = (unsigned)((E + C::EPB - 1) / C::EPB);
  if (ras)
    fdm_kernel<T, NQ, false><<<grid, C::BLOCK, 0, s>>>(E, u, S, lam, links, w, out, gout);
  else
    fdm_kernel<T, NQ, true><<<grid, C::BLOCK, 0, s>>>(E, u, S, lam, links, w, out, gout);
  NRS_LAUNCH_CHECK();
}

template <typename T, int NQ>
void launch_post_nq(int64_t n, const T *box, const T *gout, const int2 *links, const T *wts, T *Su,
                    hipStream_t s)
{
  fdm_post_kernel<T, NQ><<<grid_for(n, 256, 16384), 256, 0, s>>>(n, box, gout, links, wts, Su);
  NRS_LAUNCH_CHECK();
}

}  // namespace

template <typename T>
void Schwarz<T>::setup(int N_, int E_, bool ras_, int prec, const std::vector<double> &lengths,
                       const int *EToB, const std::vector<FaceLink> &lk, const nrs_ogs *ogs_,
                       const T *invDegree_, int64_t Nmasked_, const int *maskIds_, hipStream_t s)
{
  N = N_; E = E_; ras = ras_; fdmPrecision = prec;
  NRS_CHECK(N >= 1 && N <= 9, "schwarz: orders 1..9 are built");
  NRS_CHECK(prec == NRS_FDM_F32, "schwarz: only FP32/FP64 FDM contractions are built");
  Nq = N + 1; Np = Nq * Nq * Nq; nl = N + 3;
  ogs = ogs_; invDegree = invDegree_; Nmasked = Nmasked_; maskIds = maskIds_;
  std::vector<double> r, wq;
  gll_nodes(N, r, wq);
  const auto D = dmatrix(r);
  const double *left = lengths.data(), *mid = left + 3 * (size_t)E, *right = mid + 3 * (size_t)E;
  std::vector<T> Sh((size_t)E * 3 * nl * nl);
  std::vector<double> lh((size_t)E * 3 * nl);
  std::vector<int> bad(E, 0);
  static const int lf[3] = {4, 1, 0}, rf[3] = {2, 3, 5};   // compute_element_boundary_conditions
  parallel_for(E, [&](int64_t e) {
    std::vector<double> Sd(nl * nl);
    for (int d = 0; d < 3; ++d) {
      const int lbc = EToB[e * 6 + lf[d]], rbc = EToB[e * 6 + rf[d]];
      if (!fdm_1d(N, D.data(), wq.data(), lbc, rbc, left[d * E + e], mid[d * E + e],
                  right[d * E + e], Sd.data(), &lh[(e * 3 + d) * nl])) {
        bad[e] = 1;
        return;
      }
      for (int q = 0; q < nl * nl; ++q) Sh[(e * 3 + d) * nl * nl + q] = (T)Sd[q];
    }
  });
  NRS_CHECK(std::find(bad.begin(), bad.end(), 1) == bad.end(),
            "schwarz: 1D mass matrix not positive definite");
  S.upload(Sh);
  lam.upload(lh);
  std::vector<int2> l2((size_t)E * 6);
  for (size_t q = 0; q < l2.size(); ++q) l2[q] = make_int2(lk[q].nbr, lk[q].code);
  links.upload(l2);
  const int64_t L = (int64_t)E * Np;
  if (!ras) {
    wts.alloc(L);
    asm_count_kernel<T><<<grid_for(L, 256, 16384), 256, 0, s>>>(N, L, links.get(), wts.get());
    NRS_LAUNCH_CHECK();
    ogs_gather_scatter<T>(ogs, wts.get(), 1, L, NRS_ADD, s);
    reciprocal<T>(L, wts.get(), s);
    boxOut.alloc(L);
    ghostOut.alloc((size_t)E * 6 * Nq * Nq);
  }
  NRS_HIP(hipStreamSynchronize(s));
}

template <typename T>
void Schwarz<T>::apply(const T *u, T *Su, hipStream_t s)
{
  NRS_CHECK((const void *)u != (const void *)Su, "schwarz: input and output must differ");
  const int64_t L = (int64_t)E * Np;
  T *out = ras ? Su : boxOut.get();
  T *gout = ras ? nullptr : ghostOut.get();
  const T *w = ras ? invDegree : nullptr;
#define NRS_FDM_CASE(NQ_)                                                                      \
  case NQ_:                                                                                    \
    launch_fdm_nq<T, NQ_>(ras, E, u, S.get(), lam.get(), links.get(), w, out, gout, s);       \
    if (!ras) launch_post_nq<T, NQ_>(L, boxOut.get(), ghostOut.get(), links.get(), wts.get(), Su, s); \
    break;
  switch (Nq) {
    NRS_FDM_CASE(2) NRS_FDM_CASE(3) NRS_FDM_CASE(4) NRS_FDM_CASE(5) NRS_FDM_CASE(6)
    NRS_FDM_CASE(7) NRS_FDM_CASE(8) NRS_FDM_CASE(9) NRS_FDM_CASE(10)
    default: throw Error(NRS_ERR_INVALID, "schwarz: unsupported order");
  }
#undef NRS_FDM_CASE
  ogs_gather_scatter<T>(ogs, Su, 1, L, NRS_ADD, s);
  launch_mask<T>(Nmasked, maskIds, Su, s);
}

template struct Schwarz<float>;
template struct Schwarz<double>;

}  // namespace nrs
