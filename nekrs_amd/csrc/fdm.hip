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
#include <cstring>
#include <thread>
#include <type_traits>

#include "fdm.hpp"
#include "tune.hpp"
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

// FP32 -> BF16 / FP16 bit patterns, round to nearest even (the device rounding)
uint16_t host_bf16(float x)
{
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
uint16_t host_f16(float x)
{
  uint32_t u;
  std::memcpy(&u, &x, 4);
  const uint32_t sign = (u >> 16) & 0x8000u;
  const uint32_t a = u & 0x7FFFFFFFu;
  if (a >= 0x7F800000u) return (uint16_t)(sign | 0x7C00u | (a > 0x7F800000u ? 0x200u : 0u));
  if (a >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);          // rounds to >= 65520: inf
  if (a < 0x38800000u) {                                            // below 2^-14: subnormal
    const int ex = (int)(a >> 23);
    const int shift = 126 - ex;                                     // value = m * 2^-shift units of 2^-24
    if (ex == 0 || shift > 24) return (uint16_t)sign;
    const uint32_t m = (a & 0x7FFFFFu) | 0x800000u;
    uint32_t q = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return (uint16_t)(sign | q);
  }
  uint32_t q = a - 0x38000000u;                                     // rebias 127 -> 15
  const uint32_t rem = q & 0x1FFFu;
  q >>= 13;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
  return (uint16_t)(sign | q);
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
//   RAS and ASM: out = restricted result; ASM also gout = overlap (ghost face)
//   results.  The RAS weight (invDegree) is applied inside the gather-scatter.
template <typename T, int NQ, bool ASM>
__global__ __launch_bounds__(FdmCfg<NQ>::BLOCK) void fdm_kernel(
    int E, const T *__restrict__ u, const T *__restrict__ S, const double *__restrict__ lam,
    double2 helm, const int4 *__restrict__ gmap, T *__restrict__ out, T *__restrict__ gout,
    const uint32_t *__restrict__ mbits)
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
      const int4 gm = gmap[(int64_t)e * 6 + f];
      const T nb = (gm.x >= 0) ? u[gm.x + uu * gm.y + vv * gm.z] : T(0);
      X[g] = (own + nb) - own;     // 0 without a neighbour, as the unshared extended node
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
      const T inv = (dg > 1e-5) ? T(1.0 / (helm.x * dg + helm.y)) : T(0);
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
    // RAS writes Su itself: its masked nodes (in no gather-scatter row) as zero
    const uint32_t *mb = (!ASM && mbits) ? mbits + (int64_t)e * ((NP + 31) / 32) : nullptr;
#pragma unroll
    for (int k = 1; k <= NL - 2; ++k) {
      const int n = ((k - 1) * NQ + (a - 1)) * NQ + (b - 1);
      out[eb + n] = (mb && ((mb[n >> 5] >> (n & 31)) & 1u)) ? T(0) : o[k];
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
// then the Schwarz weight.  One thread per (j, k) row of i values.
template <typename T, int NQ>
__global__ __launch_bounds__(256) void fdm_post_kernel(int E, const T *__restrict__ box,
                                                       const T *__restrict__ gout,
                                                       const int4 *__restrict__ smap,
                                                       const T *__restrict__ wts, T *__restrict__ Su,
                                                       const uint32_t *__restrict__ mbits)
{
  constexpr int NP = NQ * NQ * NQ, NQ2 = NQ * NQ;
  const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (row >= (int64_t)E * NQ2) return;
  const int64_t e = row / NQ2;
  const int r = (int)(row - e * NQ2), j = r % NQ, k = r / NQ;
  const int64_t base = e * NP + (int64_t)r * NQ;
  const T *ge = gout + e * 6 * NQ2;
  T v[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) v[i] = box[base + i];
  // overlap value at face point (uu, vv) of face f: (own + neighbour) - own
  auto ov = [&](int f, int uu, int vv) -> T {
    const int4 sm = smap[e * 6 + f];
    if (sm.x < 0) return T(0);
    const T own = ge[f * NQ2 + vv * NQ + uu];
    return (own + gout[sm.x + uu * sm.y + vv * sm.z]) - own;
  };
  if (k == 1)
#pragma unroll
    for (int i = 0; i < NQ; ++i) v[i] += ov(0, i, j);
  if (k == NQ - 2)
#pragma unroll
    for (int i = 0; i < NQ; ++i) v[i] += ov(5, i, j);
  if (j == 1)
#pragma unroll
    for (int i = 0; i < NQ; ++i) v[i] += ov(1, i, k);
  if (j == NQ - 2)
#pragma unroll
    for (int i = 0; i < NQ; ++i) v[i] += ov(3, i, k);
  v[1] += ov(4, j, k);
  v[NQ - 2] += ov(2, j, k);
  // masked nodes (in no gather-scatter row) stored as zero
  const uint32_t *mb = mbits ? mbits + e * ((NP + 31) / 32) : nullptr;
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    const int n = r * NQ + i;
    Su[base + i] = (mb && ((mb[n >> 5] >> (n & 31)) & 1u)) ? T(0) : v[i] * wts[base + i];
  }
}

// overlap counts behind the ASM weights (generate_weights, :798-835 before
// its gather-scatter): 1 + number of faces with a neighbour whose second
// layer holds the node
template <typename T>
__global__ void asm_count_kernel(int N, int64_t n, const int4 *__restrict__ gmap, T *__restrict__ c)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = idx / Np;
    const int r = (int)(idx - e * Np);
    const int i = r % Nq, j = (r / Nq) % Nq, k = r / (Nq * Nq);
    int cnt = 1;
    const int4 *lk = gmap + e * 6;
    cnt += (k == 1 && lk[0].x >= 0) + (k == N - 1 && lk[5].x >= 0);
    cnt += (j == 1 && lk[1].x >= 0) + (j == N - 1 && lk[3].x >= 0);
    cnt += (i == 1 && lk[4].x >= 0) + (i == N - 1 && lk[2].x >= 0);
    c[idx] = T(cnt);
  }
}

// ---- BF16 / FP16 variant: the six contractions on the matrix cores
//
// One element per wavefront.  The extended box lives in LDS as 16-bit values in
// a [plane][row][16] layout; every contraction is a (16 x 16) x (16 x 16) MFMA
// per plane (v_mfma_f32_16x16x16_{bf16,f16}, FP32 accumulation) whose result is
// written back transposed so that the next contraction index is the contiguous
// one: X[k][j][i] -> T1[m][k][j] -> T2[n][m][k] -> T3[p][n][m] (times invL) ->
// T4[i][p][n] -> T5[j][i][p] -> out[k][j][i].  Each axis is stored in the
// order (interior 0..NQ-1, left ghost, right ghost, zero padding), so interior
// rows are contiguous and aligned; the 1D operators are stored with their rows
// permuted the same way.  FP16 data are kept in range by a power-of-two scale
// per element and contraction (block floating point), which leaves the
// rounding of an 11-bit significand.

typedef _Float16 fdm_h4 __attribute__((ext_vector_type(4)));
typedef short fdm_s4 __attribute__((ext_vector_type(4)));
typedef float fdm_f4 __attribute__((ext_vector_type(4)));

typedef __bf16 fdm_bf2 __attribute__((ext_vector_type(2)));
typedef _Float16 fdm_h2 __attribute__((ext_vector_type(2)));
typedef float fdm_f2 __attribute__((ext_vector_type(2)));
// two FP32 -> two 16-bit values in one word (v_cvt_pk_bf16_f32 / v_cvt_pk_f16_f32,
// round to nearest even)
template <int FMT> __device__ __forceinline__ uint32_t pack2(float a, float b)
{
  const fdm_f2 v = {a, b};
  if constexpr (FMT == NRS_FDM_BF16) return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, fdm_bf2));
  else return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, fdm_h2));
}
template <int FMT> __device__ __forceinline__ uint2 pack4(fdm_f4 v)
{
  return make_uint2(pack2<FMT>(v[0], v[1]), pack2<FMT>(v[2], v[3]));
}
template <int FMT> __device__ __forceinline__ fdm_f4 mfma16(uint2 a, uint2 b)
{
  const fdm_f4 z = {0.f, 0.f, 0.f, 0.f};
  if constexpr (FMT == NRS_FDM_BF16)
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(fdm_s4, a),
                                                     __builtin_bit_cast(fdm_s4, b), z, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(fdm_h4, a),
                                                 __builtin_bit_cast(fdm_h4, b), z, 0, 0, 0);
}
// LDS traffic of one wavefront is ordered; the fence keeps the compiler from
// moving LDS accesses across a contraction boundary.
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

template <int NQ> struct MfmaCfg {
  static constexpr int NL = NQ + 2;
  // Row stride RW = 20 halves (10 dwords): the rows an operand read takes
  // (lane c reads row c) start at 10 c mod 32 dwords, distinct even banks.
  // Plane stride PS: NL rows (the MFMA's rows NL..15 are zero, not stored), at
  // least the staged results of an element (ASM: NP + 6 NQ^2 floats) over NL
  // planes, and 2 mod 4 dwords, so the transposed result store (lane c writes
  // plane c) starts at distinct even banks too.
  static constexpr int RW = 20;
  static constexpr int PSMIN0 = (2 * (NQ * NQ * NQ + 6 * NQ * NQ) + NL - 1) / NL;
  static constexpr int PSMIN = (NL * RW > PSMIN0) ? NL * RW : PSMIN0;
  static constexpr int PSD0 = (PSMIN + 1) / 2;
  static constexpr int PSD = PSD0 + ((2 - PSD0 % 4) + 4) % 4;
  static constexpr int PS = 2 * PSD;
  static constexpr int BUF = NL * PS;
  static constexpr int SST = (3 * NL * NL + 7) / 8 * 8;   // operator halves per element, 16-byte multiple
  static constexpr int TAIL = SST;
  static constexpr int WAVES = 4;
  static_assert(NL <= 16, "one 16-wide MFMA tile per axis");
  static_assert(PSD % 4 == 2, "plane stride 2 mod 4 dwords");
};

// Each wavefront walks its elements in a grid-stride loop.  The global loads
// of the next element (interior rows, the own and neighbour values behind each
// ghost, operators, eigenvalues, ghost maps) are issued before the contractions
// of the current one, and the results of an element are staged in LDS and
// written out (as whole rows) only after the next element's loads have
// arrived, so no wait on outstanding stores sits in front of the prefetch.
template <int NQ, int FMT, bool ASM>
__global__ __launch_bounds__(256) void fdm_mfma_kernel(
    int E, const float *__restrict__ u, const uint16_t *__restrict__ S16,
    const float *__restrict__ lam, float2 helm, const int4 *__restrict__ gmap,
    float *__restrict__ out, float *__restrict__ gout)
{
  using C = MfmaCfg<NQ>;
  constexpr int NL = C::NL, PS = C::PS, RW = C::RW, BUF = C::BUF, NP = NQ * NQ * NQ, NQ2 = NQ * NQ;
  constexpr int SST = C::SST;                  // halves of operators per element (16-byte rows)
  constexpr int SQ4 = SST / 8;                 // uint4 per element's operators
  constexpr bool SCALE = (FMT == NRS_FDM_F16);
  constexpr int RI = (NQ2 + 63) / 64;          // interior rows per lane
  constexpr int RG = (NQ2 + 63) / 64;          // positions per face per lane
  constexpr int RS = (SQ4 + 63) / 64;          // operator uint4 per lane
  constexpr int OUTF = ASM ? NP + 6 * NQ2 : NP;   // staged floats per element
  static_assert(2 * OUTF <= BUF, "staged results fit one box buffer");
  __shared__ __align__(16) uint16_t lds[C::WAVES][2 * BUF + C::TAIL];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  uint16_t *X = lds[wv], *Y = lds[wv] + BUF, *SB = lds[wv] + 2 * BUF;
  float *stage = reinterpret_cast<float *>(Y);   // results of the previous element
  {
    uint4 *p = reinterpret_cast<uint4 *>(lds[wv]);
    for (int n = lane; n < (2 * BUF + C::TAIL) * 2 / 16; n += 64) p[n] = make_uint4(0u, 0u, 0u, 0u);
  }
  wave_lds_sync();
  const int stride = gridDim.x * C::WAVES;
  int e = __builtin_amdgcn_readfirstlane(blockIdx.x * C::WAVES + wv);
  if (e >= E) return;

  // ---- prefetch state of the next element
  float pvi[RI][NQ], pown[6][RG], pnb[6][RG], plx[4], plz, ply[NL];
  uint4 ps[RS];
  int4 pgm[6];
  auto load_a = [&](int en) {
    const float *ue = u + (int64_t)en * NP;
#pragma unroll
    for (int t = 0; t < RI; ++t) {
      const int n = lane + 64 * t;
#pragma unroll
      for (int i = 0; i < NQ; ++i) pvi[t][i] = (n < NQ2) ? ue[n * NQ + i] : 0.f;
    }
#pragma unroll
    for (int f = 0; f < 6; ++f)
#pragma unroll
      for (int t = 0; t < RG; ++t) {
        const int n = lane + 64 * t;
        pown[f][t] = (n < NQ2) ? ue[face_node(f, n % NQ, n / NQ, 1, NQ - 1)] : 0.f;
      }
    const uint4 *Se = reinterpret_cast<const uint4 *>(S16 + (int64_t)en * SST);
#pragma unroll
    for (int t = 0; t < RS; ++t) {
      const int n = lane + 64 * t;
      ps[t] = (n < SQ4) ? Se[n] : make_uint4(0u, 0u, 0u, 0u);
    }
    const float *le = lam + (int64_t)en * 3 * NL;
#pragma unroll
    for (int q = 0; q < 4; ++q) plx[q] = (4 * g + q < NL) ? le[4 * g + q] : 0.f;
    plz = (c < NL) ? le[2 * NL + c] : 0.f;
#pragma unroll
    for (int a = 0; a < NL; ++a) ply[a] = le[NL + a];
#pragma unroll
    for (int f = 0; f < 6; ++f) pgm[f] = gmap[(int64_t)en * 6 + f];
  };
  auto load_b = [&]() {   // neighbour layers through the affine ghost maps
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int4 gm = pgm[f];
#pragma unroll
      for (int t = 0; t < RG; ++t) {
        const int n = lane + 64 * t;
        pnb[f][t] = (gm.x >= 0 && n < NQ2) ? u[gm.x + (n % NQ) * gm.y + (n / NQ) * gm.z] : 0.f;
      }
    }
  };
  auto flush = [&](int ep) {   // staged results of element ep -> global rows
    for (int n = lane; n < NQ2; n += 64) {
#pragma unroll
      for (int i = 0; i < NQ; ++i) out[(int64_t)ep * NP + n * NQ + i] = stage[n * NQ + i];
    }
    if constexpr (ASM) {
      for (int n = lane; n < 6 * NQ2; n += 64) gout[(int64_t)ep * 6 * NQ2 + n] = stage[NP + n];
    }
  };
  load_a(e);
  load_b();
  int prev = -1;   // element whose results are staged

  for (;;) {
    float vi[RI][NQ], vg[6][RG], lx[4], lz, ly[NL];
    uint4 sr[RS];
#pragma unroll
    for (int t = 0; t < RI; ++t)
#pragma unroll
      for (int i = 0; i < NQ; ++i) vi[t][i] = pvi[t][i];
#pragma unroll
    for (int f = 0; f < 6; ++f)
#pragma unroll
      for (int t = 0; t < RG; ++t) vg[f][t] = (pown[f][t] + pnb[f][t]) - pown[f][t];   // fusedFDM.c:47-104
#pragma unroll
    for (int t = 0; t < RS; ++t) sr[t] = ps[t];
#pragma unroll
    for (int q = 0; q < 4; ++q) lx[q] = plx[q];
    lz = plz;
#pragma unroll
    for (int a = 0; a < NL; ++a) ly[a] = ply[a];
    const int en = e + stride;
    const bool more = en < E;

    // -- 1. the extended box: interior rows, ghost faces, zero edges; operators
    float sc = 1.f;   // stored = true * sc
    if constexpr (SCALE) {
      float m = 0.f;
#pragma unroll
      for (int t = 0; t < RI; ++t)
#pragma unroll
        for (int i = 0; i < NQ; ++i) m = fmaxf(m, fabsf(vi[t][i]));
#pragma unroll
      for (int f = 0; f < 6; ++f)
#pragma unroll
        for (int t = 0; t < RG; ++t) m = fmaxf(m, fabsf(vg[f][t]));
      m = wave_max(m);
      if (m > 0.f) {
        int ex;
        frexpf(m, &ex);
        sc = ldexpf(1.f, 15 - ex);
      }
    }
#pragma unroll
    for (int t = 0; t < RI; ++t) {
      const int n = lane + 64 * t;
      if (n < NQ2) {
        uint16_t *row = X + (n / NQ) * PS + (n % NQ) * RW;
#pragma unroll
        for (int i = 0; i + 1 < NQ; i += 2)
          *reinterpret_cast<uint32_t *>(row + i) = pack2<FMT>(vi[t][i] * sc, vi[t][i + 1] * sc);
        if (NQ & 1) row[NQ - 1] = (uint16_t)(pack2<FMT>(vi[t][NQ - 1] * sc, 0.f) & 0xFFFFu);
      }
    }
#pragma unroll
    for (int f = 0; f < 6; ++f)
#pragma unroll
      for (int t = 0; t < RG; ++t) {
        const int n = lane + 64 * t;
        if (n < NQ2) {
          const int uu = n % NQ, vv = n / NQ;
          int pos;
          switch (f) {   // (plane, row, col) of the ghost in storage order
            case 0: pos = NQ * PS + vv * RW + uu; break;
            case 5: pos = (NQ + 1) * PS + vv * RW + uu; break;
            case 1: pos = vv * PS + NQ * RW + uu; break;
            case 3: pos = vv * PS + (NQ + 1) * RW + uu; break;
            case 4: pos = vv * PS + uu * RW + NQ; break;
            default: pos = vv * PS + uu * RW + NQ + 1; break;
          }
          X[pos] = (uint16_t)(pack2<FMT>(vg[f][t] * sc, 0.f) & 0xFFFFu);
        }
      }
    for (int n = lane; n < 12 * NL; n += 64) {   // box edges (two ghost indices) are zero
      const int line = n / NL, sidx = n % NL;
      const int c0 = NQ + (line & 1), c1 = NQ + ((line >> 1) & 1), dir = line >> 2;
      const int pos = (dir == 0) ? c1 * PS + c0 * RW + sidx
                                 : (dir == 1) ? c1 * PS + sidx * RW + c0 : sidx * PS + c1 * RW + c0;
      X[pos] = 0;
    }
#pragma unroll
    for (int t = 0; t < RS; ++t) {
      const int n = lane + 64 * t;
      if (n < SQ4) reinterpret_cast<uint4 *>(SB)[n] = sr[t];
    }
    // the previous element's results leave before the next element's loads are
    // issued, so waiting for those loads never waits on these stores
    if (prev >= 0) flush(prev);
    wave_lds_sync();
    if (more) load_a(en);
    // forward B[s][m] = S[s][m], backward B[m][s] = S[s][m] (rows in storage order)
    uint2 bf[3], bb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      uint32_t f4[4], b4[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = 4 * g + q;
        const bool ok = (k < NL && c < NL);
        f4[q] = ok ? SB[(d * NL + k) * NL + c] : 0u;
        b4[q] = ok ? SB[(d * NL + c) * NL + k] : 0u;
      }
      bf[d] = make_uint2(f4[0] | (f4[1] << 16), f4[2] | (f4[3] << 16));
      bb[d] = make_uint2(b4[0] | (b4[1] << 16), b4[2] | (b4[3] << 16));
    }

    // -- 2. six contractions
    fdm_f4 D[NL];
    auto contract = [&](const uint16_t *in, uint2 bfrag) {
#pragma unroll
      for (int a = 0; a < NL; ++a)
      {
        // rows NL..15 of the operand are zero (never read): their outputs land in
        // the padding columns of the next operand, which meet the zero rows of B
        uint2 av = make_uint2(0u, 0u);
        if (c < NL) av = *reinterpret_cast<const uint2 *>(in + a * PS + c * RW + 4 * g);
        D[a] = mfma16<FMT>(av, bfrag);
      }
    };
    auto rescale = [&]() {   // FP16: bring the phase result back to [2^14, 2^15)
      if constexpr (SCALE) {
        float m = 0.f;
#pragma unroll
        for (int a = 0; a < NL; ++a)
#pragma unroll
          for (int q = 0; q < 4; ++q) m = fmaxf(m, fabsf(D[a][q]));
        m = wave_max(m);
        if (m > 0.f) {
          int ex;
          frexpf(m, &ex);
          const float sf = ldexpf(1.f, 15 - ex);
#pragma unroll
          for (int a = 0; a < NL; ++a) D[a] *= sf;
          sc *= sf;
        }
      }
    };
    auto store = [&](uint16_t *outb) {
      if (c < NL) {
#pragma unroll
        for (int a = 0; a < NL; ++a)
          *reinterpret_cast<uint2 *>(outb + c * PS + a * RW + 4 * g) = pack4<FMT>(D[a]);
      }
      wave_lds_sync();
    };
    contract(X, bf[0]); rescale(); store(Y);          // T1[m][k][j]
    contract(Y, bf[1]); rescale(); store(X);          // T2[n][m][k]
    contract(X, bf[2]);                               // T3[p][n][m] = invL * (...)
#pragma unroll
    for (int a = 0; a < NL; ++a)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float dg = (lx[q] + ly[a]) + lz;
        D[a][q] *= (dg > 1e-5f) ? __builtin_amdgcn_rcpf(helm.x * dg + helm.y) : 0.f;   // 1 ulp: below the 16-bit rounding
      }
    rescale(); store(Y);
    if (more) load_b();
    contract(Y, bb[0]); rescale(); store(X);          // T4[i][p][n]
    contract(X, bb[1]); rescale(); store(Y);          // T5[j][i][p]
    contract(Y, bb[2]);                               // out[k][j][i]: k = c, j = a, i = 4g+q
    wave_lds_sync();                                  // T5 fully read: Y becomes the staging area
    const float inv = 1.f / sc;
    if (c < NQ) {
#pragma unroll
      for (int a = 0; a < NQ; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = 4 * g + q;
          if (i < NQ) stage[(c * NQ + a) * NQ + i] = D[a][q] * inv;
        }
    }
    if constexpr (ASM) {
      float *ge = stage + NP;
      if (c < NQ) {
#pragma unroll
        for (int a = 0; a < NQ; ++a)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int i = 4 * g + q;
            if (i == NQ) ge[4 * NQ2 + c * NQ + a] = D[a][q] * inv;          // x faces: (u,v) = (j,k)
            if (i == NQ + 1) ge[2 * NQ2 + c * NQ + a] = D[a][q] * inv;
          }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = 4 * g + q;
          if (i < NQ) {                                                       // y faces: (u,v) = (i,k)
            ge[1 * NQ2 + c * NQ + i] = D[NQ][q] * inv;
            ge[3 * NQ2 + c * NQ + i] = D[NQ + 1][q] * inv;
          }
        }
      } else if (c == NQ || c == NQ + 1) {                                    // z faces: (u,v) = (i,j)
        float *gf = ge + (c == NQ ? 0 : 5) * NQ2;
#pragma unroll
        for (int a = 0; a < NQ; ++a)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int i = 4 * g + q;
            if (i < NQ) gf[a * NQ + i] = D[a][q] * inv;
          }
      }
    }
    wave_lds_sync();
    prev = e;
    if (!more) break;
    e = en;
  }
  flush(prev);
}

// epw: elements per wavefront (the tuned variant; about four by default, so
// the prefetch has something to overlap)
template <int NQ, int FMT>
void launch_fdm_mfma_nq(bool ras, int E, const float *u, const uint16_t *S16, const float *lam,
                        float2 helm, const int4 *gmap, float *out, float *gout, int epw, hipStream_t s)
{
  using C = MfmaCfg<NQ>;
  const int64_t perBlock = (int64_t)C::WAVES * std::max(1, epw);
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((E + perBlock - 1) / perBlock, 65536));
  if (ras)
    fdm_mfma_kernel<NQ, FMT, false><<<grid, 64 * C::WAVES, 0, s>>>(E, u, S16, lam, helm, gmap, out, gout);
  else
    fdm_mfma_kernel<NQ, FMT, true><<<grid, 64 * C::WAVES, 0, s>>>(E, u, S16, lam, helm, gmap, out, gout);
  NRS_LAUNCH_CHECK();
}

template <typename T, int NQ>
void launch_fdm_nq(bool ras, int E, const T *u, const T *S, const double *lam, double2 helm, const int4 *gmap,
                   T *out, T *gout, const uint32_t *mbits, hipStream_t s)
{
  using C = FdmCfg<NQ>;
  const unsigned grid = (unsigned)((E + C::EPB - 1) / C::EPB);
  if (ras)
    fdm_kernel<T, NQ, false><<<grid, C::BLOCK, 0, s>>>(E, u, S, lam, helm, gmap, out, gout, mbits);
  else
    fdm_kernel<T, NQ, true><<<grid, C::BLOCK, 0, s>>>(E, u, S, lam, helm, gmap, out, gout, mbits);
  NRS_LAUNCH_CHECK();
}

template <typename T, int NQ>
void launch_post_nq(int E, const T *box, const T *gout, const int4 *smap, const T *wts, T *Su,
                    const uint32_t *mbits, hipStream_t s)
{
  const int64_t rows = (int64_t)E * NQ * NQ;
  fdm_post_kernel<T, NQ><<<(unsigned)((rows + 255) / 256), 256, 0, s>>>(E, box, gout, smap, wts, Su, mbits);
  NRS_LAUNCH_CHECK();
}

}  // namespace

template <typename T>
void Schwarz<T>::setup(int N_, int E_, bool ras_, int prec, const std::vector<double> &lengths,
                       const int *EToB, const std::vector<FaceLink> &lk, const nrs_ogs *ogs_,
                       const T *invDegree_, int64_t Nmasked_, const int *maskIds_,
                       const uint32_t *maskBits_, bool poisson, double lambda0, double lambda1, hipStream_t s)
{
  N = N_; E = E_; ras = ras_; fdmPrecision = prec;
  NRS_CHECK(N >= 1 && N <= 9, "schwarz: orders 1..9 are built");
  NRS_CHECK((prec == NRS_FDM_F32 || std::is_same<T, float>::value),
            "schwarz: BF16/FP16 contractions need the FP32 multigrid");
  Nq = N + 1; Np = Nq * Nq * Nq; nl = N + 3;
  ogs = ogs_; invDegree = invDegree_; Nmasked = Nmasked_; maskIds = maskIds_;
  maskBits = Nmasked > 0 ? maskBits_ : nullptr;
  if (maskBits) {
    // the kernels that write Su, and the weighted gather-scatter, store the
    // masked nodes as zero in place of a mask pass: none may be in a row of ogs
    std::vector<int> ids((size_t)Nmasked);
    NRS_HIP(hipMemcpy(ids.data(), maskIds, (size_t)Nmasked * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<uint8_t> isMasked((size_t)E * Np, 0);
    for (int id : ids) isMasked[id] = 1;
    for (const std::vector<int> *h : {&ogs->h2, &ogs->h4, &ogs->h8, &ogs->hxi})
      for (int id : *h) NRS_CHECK(!isMasked[id], "schwarz: a masked node is in a row of the gather-scatter");
  }
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
  // Helmholtz: the box solve inverts lambda0 (lx + ly + lz) + lambda1 in the
  // eigenbasis (S^T B S = I per direction), except the modes with
  // lx + ly + lz <= 1e-5 (constant on a box with free ends), dropped as in the
  // Poisson case; inverting those by 1/lambda1 would make the smoother's
  // largest eigenvalue grow like 1/(lambda1 h^2) under refinement.
  helm0 = poisson ? 1.0 : lambda0;
  helm1 = poisson ? 0.0 : lambda1;
  S.upload(Sh);
  lam.upload(lh);
  if (prec != NRS_FDM_F32) {
    // 16-bit operators with rows in the MFMA kernel's storage order (interior,
    // left ghost, right ghost), rounded from the FP32 values; FP32 eigenvalues
    const int sst = (3 * nl * nl + 7) / 8 * 8;      // 16-byte aligned per element
    std::vector<uint16_t> s16((size_t)E * sst, 0);
    std::vector<float> l32(lh.size());
    for (size_t q = 0; q < lh.size(); ++q) l32[q] = (float)lh[q];
    for (int64_t e = 0; e < E; ++e)
      for (int d = 0; d < 3; ++d)
        for (int sp = 0; sp < nl; ++sp) {
          const int phys = (sp < Nq) ? sp + 1 : (sp == Nq ? 0 : nl - 1);
          for (int m = 0; m < nl; ++m) {
            const float v = (float)Sh[((e * 3 + d) * nl + phys) * nl + m];
            s16[e * sst + (d * nl + sp) * nl + m] = (prec == NRS_FDM_BF16) ? host_bf16(v) : host_f16(v);
          }
        }
    S16.upload(s16);
    lam32.upload(l32);
  }
  // affine maps of the face links at this order: the neighbour's second-layer
  // node behind ghost point (u, v) of face f is gm.x + u gm.y + v gm.z, its
  // overlap-result slot sm.x + u sm.y + v sm.z (x < 0: no neighbour)
  NRS_CHECK((int64_t)E * Np < INT32_MAX && (int64_t)E * 6 * Nq * Nq < INT32_MAX,
            "schwarz: E-vector too long for 32-bit indices");
  std::vector<int4> gm((size_t)E * 6), sm((size_t)E * 6);
  for (int64_t q = 0; q < (int64_t)E * 6; ++q) {
    const FaceLink &l = lk[q];
    if (l.nbr < 0) {
      gm[q] = make_int4(-1, 0, 0, 0);
      sm[q] = make_int4(-1, 0, 0, 0);
      continue;
    }
    const int fb = l.code & 7;
    auto node = [&](int uu, int vv) {
      int up, vp;
      link_map(l.code, N, uu, vv, up, vp);
      return (int)((int64_t)l.nbr * Np + face_node(fb, up, vp, 1, N));
    };
    auto slot = [&](int uu, int vv) {
      int up, vp;
      link_map(l.code, N, uu, vv, up, vp);
      return (int)(((int64_t)l.nbr * 6 + fb) * Nq * Nq + vp * Nq + up);
    };
    const int n0 = node(0, 0), s0 = slot(0, 0);
    gm[q] = make_int4(n0, node(1, 0) - n0, node(0, 1) - n0, 0);
    sm[q] = make_int4(s0, slot(1, 0) - s0, slot(0, 1) - s0, 0);
  }
  gmap.upload(gm);
  if (!ras) smap.upload(sm);
  const int64_t L = (int64_t)E * Np;
  if (!ras) {
    wts.alloc(L);
    asm_count_kernel<T><<<grid_for(L, 256, 16384), 256, 0, s>>>(N, L, gmap.get(), wts.get());
    NRS_LAUNCH_CHECK();
    ogs_gather_scatter<T>(ogs, wts.get(), 1, L, NRS_ADD, s);
    reciprocal<T>(L, wts.get(), s);
    boxOut.alloc(L);
    ghostOut.alloc((size_t)E * 6 * Nq * Nq);
  }
  NRS_HIP(hipStreamSynchronize(s));
}

// the matrix-core box solves of one Schwarz level (FP32 vectors)
static void fdm_mfma_boxes(int Nq, int fmt, bool ras, int E, const float *u, const uint16_t *S16,
                           const float *lam, float2 helm, const int4 *gmap, float *out, float *gout, int epw,
                           hipStream_t s)
{
#define NRS_FDM_MFMA_CASE(NQ_)                                                                 \
  case NQ_:                                                                                    \
    if (fmt == NRS_FDM_BF16)                                                                   \
      launch_fdm_mfma_nq<NQ_, NRS_FDM_BF16>(ras, E, u, S16, lam, helm, gmap, out, gout, epw, s);     \
    else                                                                                       \
      launch_fdm_mfma_nq<NQ_, NRS_FDM_F16>(ras, E, u, S16, lam, helm, gmap, out, gout, epw, s);      \
    break;
  switch (Nq) {
    NRS_FDM_MFMA_CASE(2) NRS_FDM_MFMA_CASE(3) NRS_FDM_MFMA_CASE(4) NRS_FDM_MFMA_CASE(5)
    NRS_FDM_MFMA_CASE(6) NRS_FDM_MFMA_CASE(7) NRS_FDM_MFMA_CASE(8) NRS_FDM_MFMA_CASE(9)
    NRS_FDM_MFMA_CASE(10)
    default: throw Error(NRS_ERR_INVALID, "schwarz: unsupported order");
  }
#undef NRS_FDM_MFMA_CASE
}

template <typename T>
void Schwarz<T>::apply(const T *u, T *Su, hipStream_t s)
{
  NRS_CHECK((const void *)u != (const void *)Su, "schwarz: input and output must differ");
  T *out = ras ? Su : boxOut.get();
  T *gout = ras ? nullptr : ghostOut.get();
  if constexpr (std::is_same<T, float>::value) {
    if (fdmPrecision != NRS_FDM_F32) {
      const int epw = tuned(TUNE_FDM, N, fdmPrecision, E, 4);
      fdm_mfma_boxes(Nq, fdmPrecision, ras, E, u, S16.get(), lam32.get(), helm32(), gmap.get(), out, gout, epw, s);
      if (!ras) {
#define NRS_FDM_POST_CASE(NQ_)                                                                 \
  case NQ_: launch_post_nq<T, NQ_>(E, boxOut.get(), ghostOut.get(), smap.get(), wts.get(), Su, maskBits, s); break;
        switch (Nq) {
          NRS_FDM_POST_CASE(2) NRS_FDM_POST_CASE(3) NRS_FDM_POST_CASE(4) NRS_FDM_POST_CASE(5)
          NRS_FDM_POST_CASE(6) NRS_FDM_POST_CASE(7) NRS_FDM_POST_CASE(8) NRS_FDM_POST_CASE(9)
          NRS_FDM_POST_CASE(10)
          default: throw Error(NRS_ERR_INVALID, "schwarz: unsupported order");
        }
#undef NRS_FDM_POST_CASE
      }
      finish(Su, !ras && maskBits, s);   // the matrix-core RAS kernel stores no mask
      return;
    }
  }
#define NRS_FDM_CASE(NQ_)                                                                      \
  case NQ_:                                                                                    \
    launch_fdm_nq<T, NQ_>(ras, E, u, S.get(), lam.get(), make_double2(helm0, helm1), gmap.get(), out, gout,       \
                          maskBits, s);                                                                        \
    if (!ras) launch_post_nq<T, NQ_>(E, boxOut.get(), ghostOut.get(), smap.get(), wts.get(), Su, maskBits, s);   \
    break;
  switch (Nq) {
    NRS_FDM_CASE(2) NRS_FDM_CASE(3) NRS_FDM_CASE(4) NRS_FDM_CASE(5) NRS_FDM_CASE(6)
    NRS_FDM_CASE(7) NRS_FDM_CASE(8) NRS_FDM_CASE(9) NRS_FDM_CASE(10)
    default: throw Error(NRS_ERR_INVALID, "schwarz: unsupported order");
  }
#undef NRS_FDM_CASE
  finish(Su, maskBits != nullptr, s);
}

// RAS: Su = gs(invDegree .* restricted) (fusedFDM.c p_restrict, then oogs); the
// weighted gather-scatter multiplies each member by its weight before the row
// sum, as the reference's separate multiply does.  ASM: Su = gs(postFDM(...)).
// Then the mask (smoothSchwarz, :1146), unless the kernel that wrote Su stored
// its masked nodes as zero already (masked); RAS stores them as zero inside the
// weighted gather-scatter's launch.  Either is the same result as a pass after
// it: the masked nodes are in no row of the gather-scatter (checked at setup).
template <typename T>
void Schwarz<T>::finish(T *Su, bool masked, hipStream_t s)
{
  const int64_t L = (int64_t)E * Np;
  if (ras) {
    const bool inGs = !masked && maskBits && ogs->Nrows > 0;
    ogs_gather_scatter_weighted<T>(ogs, Su, invDegree, s, inGs ? maskIds : nullptr, inGs ? Nmasked : 0);
    if (!masked && !inGs) launch_mask<T>(Nmasked, maskIds, Su, s);
    return;
  }
  ogs_gather_scatter<T>(ogs, Su, 1, L, NRS_ADD, s);
  if (!masked) launch_mask<T>(Nmasked, maskIds, Su, s);
}

// Time the matrix-core box solve with each elements-per-wavefront variant on
// a fixed pattern, keep the variants whose result matches the default's within
// 100 eps relative (src/bench/fdm/benchmarkFDM.cpp:180-210), record the fastest.
template <typename T>
int Schwarz<T>::autotune(hipStream_t s)
{
  if (!std::is_same<T, float>::value || fdmPrecision == NRS_FDM_F32) return 0;
  const int64_t L = (int64_t)E * Np;
  const int64_t G = ras ? 0 : (int64_t)E * 6 * Nq * Nq;
  DevBuf<float> u(L), ref(L + G), out(L + G);
  fill_pattern(L, u.get(), s);
  int vals[8];
  const int nv = tune_variants(TUNE_FDM, vals, 8);
  auto run = [&](int epw, float *o) {
    fdm_mfma_boxes(Nq, fdmPrecision, ras, E, u.get(), S16.get(), lam32.get(), helm32(), gmap.get(), o,
                   ras ? nullptr : o + L, epw, s);
  };
  run(4, ref.get());
  std::vector<float> hr(L + G), ho(L + G);
  NRS_HIP(hipMemcpyAsync(hr.data(), ref.get(), (L + G) * sizeof(float), hipMemcpyDeviceToHost, s));
  NRS_HIP(hipStreamSynchronize(s));
  int best = 4;
  double bestMs = 1e300;
  for (int v = 0; v < nv; ++v) {
    run(vals[v], out.get());
    NRS_HIP(hipMemcpyAsync(ho.data(), out.get(), (L + G) * sizeof(float), hipMemcpyDeviceToHost, s));
    NRS_HIP(hipStreamSynchronize(s));
    if (max_rel_diff(ho.data(), hr.data(), L + G) > 100.0 * 1.1920928955078125e-07) continue;
    const double ms = time_launches([&] { run(vals[v], out.get()); }, s);
    if (ms < bestMs) { bestMs = ms; best = vals[v]; }
  }
  set_tuned(TUNE_FDM, N, fdmPrecision, E, best);
  return best;
}

template struct Schwarz<float>;
template struct Schwarz<double>;

}  // namespace nrs
