// mesh.hip -- setup geometry: GLL basis, structured box meshes, geometric factors.
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "kernels.hpp"
#include "mesh.hpp"

namespace nrs {

// Legendre P_n and P_n' by the three-term recurrence.
static void legendre(int n, double x, double &p, double &dp)
{
  double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
  if (n == 0) { p = 1.0; dp = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    const double d2 = d0 + (2.0 * k - 1.0) * p1;
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
  p = p1; dp = d1;
}

// GLL nodes (roots of (1-x^2) P_N'(x), Newton from Chebyshev-Lobatto points) and
// weights 2/(N(N+1) P_N(z)^2).  Same quantities as JacobiGLL
// (src/mesh/meshBasis1D.cpp:249-281).
void gll_nodes(int N, std::vector<double> &z, std::vector<double> &w)
{
  z.assign(N + 1, 0.0);
  w.assign(N + 1, 0.0);
  if (N == 0) { w[0] = 2.0; return; }
  for (int i = 0; i <= N; ++i) {
    double x = -std::cos(M_PI * i / N);
    if (i > 0 && i < N) {
      for (int it = 0; it < 100; ++it) {
        double p, dp;
        legendre(N, x, p, dp);
        const double dx = ((1.0 - x * x) * dp) / (-N * (N + 1.0) * p);
        x -= dx;
        if (std::fabs(dx) < 1e-16) break;
      }
    }
    z[i] = x;
  }
  z[0] = -1.0; z[N] = 1.0;
  for (int i = 0; i <= N / 2; ++i) {
    const double a = 0.5 * (z[N - i] - z[i]);
    z[i] = -a; z[N - i] = a;
  }
  if (N % 2 == 0) z[N / 2] = 0.0;
  for (int i = 0; i <= N; ++i) {
    double p, dp;
    legendre(N, z[i], p, dp);
    w[i] = 2.0 / (N * (N + 1.0) * p * p);
  }
}

static std::vector<double> bary(const std::vector<double> &z)
{
  const int n = (int)z.size();
  std::vector<double> lam(n);
  for (int j = 0; j < n; ++j) {
    double prod = 1.0;
    for (int k = 0; k < n; ++k)
      if (k != j) prod *= (z[j] - z[k]);
    lam[j] = 1.0 / prod;
  }
  return lam;
}

// D[i*Nq+j] = l_j'(z_i): barycentric form with negative-sum diagonal
// (replaces Dmatrix1D, src/mesh/meshBasis1D.cpp:99-119).
std::vector<double> dmatrix(const std::vector<double> &z)
{
  const int Nq = (int)z.size();
  const auto lam = bary(z);
  std::vector<double> D(Nq * Nq, 0.0);
  for (int i = 0; i < Nq; ++i) {
    double diag = 0.0;
    for (int j = 0; j < Nq; ++j) {
      if (j == i) continue;
      const double v = (lam[j] / lam[i]) / (z[i] - z[j]);
      D[i * Nq + j] = v;
      diag -= v;
    }
    D[i * Nq + i] = diag;
  }
  return D;
}

// I[i*Nin+j] = l_j(zout_i) (replaces InterpolationMatrix1D, meshBasis1D.cpp:132-151).
std::vector<double> interp_matrix(const std::vector<double> &zin, const std::vector<double> &zout)
{
  const int Nin = (int)zin.size(), Nout = (int)zout.size();
  const auto lam = bary(zin);
  std::vector<double> I(Nout * Nin, 0.0);
  for (int i = 0; i < Nout; ++i) {
    int hit = -1;
    for (int j = 0; j < Nin; ++j)
      if (zout[i] == zin[j]) hit = j;
    if (hit >= 0) { I[i * Nin + hit] = 1.0; continue; }
    double den = 0.0;
    for (int j = 0; j < Nin; ++j) den += lam[j] / (zout[i] - zin[j]);
    for (int j = 0; j < Nin; ++j) I[i * Nin + j] = (lam[j] / (zout[i] - zin[j])) / den;
  }
  return I;
}

// ---- Kershaw map: examples/kershaw/kershaw.usr:107-189 ----
static double k_right(double eps, double x) { return (x <= 0.5) ? (2.0 - eps) * x : 1.0 + eps * (x - 1.0); }
static double k_left(double eps, double x) { return 1.0 - k_right(eps, 1.0 - x); }
static double k_step(double a, double b, double x)
{
  if (x <= 0.0) return a;
  if (x >= 1.0) return b;
  return a + (b - a) * x;
}
static void kershaw(double ey, double ez, double x, double y, double z, double &xx, double &yy, double &zz)
{
  xx = x;
  const int layer = (int)(x * 6.0);
  const double lambda = (x - layer / 6.0) * 6.0;
  if (layer == 0) {
    yy = k_left(ey, y); zz = k_left(ez, z);
  } else if (layer == 1 || layer == 4) {
    yy = k_step(k_left(ey, y), k_right(ey, y), lambda);
    zz = k_step(k_left(ez, z), k_right(ez, z), lambda);
  } else if (layer == 2) {
    yy = k_step(k_right(ey, y), k_left(ey, y), lambda / 2.0);
    zz = k_step(k_right(ez, z), k_left(ez, z), lambda / 2.0);
  } else if (layer == 3) {
    yy = k_step(k_right(ey, y), k_left(ey, y), (1.0 + lambda) / 2.0);
    zz = k_step(k_right(ez, z), k_left(ez, z), (1.0 + lambda) / 2.0);
  } else {
    yy = k_right(ey, y); zz = k_right(ez, z);
  }
}

void box_mesh(int nx, int ny, int nz, int N, int map, double eps, double *x, double *y, double *z,
              long long *gid, int *EToB)
{
  NRS_CHECK(nx > 0 && ny > 0 && nz > 0 && N >= 1, "box mesh: bad sizes");
  const int Nq = N + 1;
  const int64_t Np = (int64_t)Nq * Nq * Nq;
  std::vector<double> r, w;
  gll_nodes(N, r, w);
  const long long gx = (long long)nx * N + 1, gy = (long long)ny * N + 1;
  for (int ez = 0; ez < nz; ++ez)
    for (int ey = 0; ey < ny; ++ey)
      for (int ex = 0; ex < nx; ++ex) {
        const int64_t e = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
        double vx[8], vy[8], vz[8];
        for (int c = 0; c < 8; ++c) {
          const double X = (double)(ex + (c & 1)) / nx;
          const double Y = (double)(ey + ((c >> 1) & 1)) / ny;
          const double Z = (double)(ez + ((c >> 2) & 1)) / nz;
          if (map == 2) kershaw(eps, eps, X, Y, Z, vx[c], vy[c], vz[c]);
          else { vx[c] = X; vy[c] = Y; vz[c] = Z; }
        }
        for (int k = 0; k < Nq; ++k)
          for (int j = 0; j < Nq; ++j)
            for (int i = 0; i < Nq; ++i) {
              const int64_t n = e * Np + i + Nq * (j + Nq * k);
              const double a = 0.5 * (1.0 + r[i]), b = 0.5 * (1.0 + r[j]), c = 0.5 * (1.0 + r[k]);
              double px = 0, py = 0, pz = 0;
              for (int v = 0; v < 8; ++v) {
                const double wv = ((v & 1) ? a : 1.0 - a) * (((v >> 1) & 1) ? b : 1.0 - b) *
                                  (((v >> 2) & 1) ? c : 1.0 - c);
                px += wv * vx[v]; py += wv * vy[v]; pz += wv * vz[v];
              }
              if (map == 1) {
                const double sb = 0.05 * std::sin(M_PI * px) * std::sin(M_PI * py) * std::sin(M_PI * pz);
                px += sb; py += sb; pz += sb;
              }
              x[n] = px; y[n] = py; z[n] = pz;
              const long long Gx = (long long)ex * N + i, Gy = (long long)ey * N + j, Gz = (long long)ez * N + k;
              if (gid) gid[n] = 1 + Gx + gx * (Gy + gy * Gz);
            }
        if (EToB) {
          EToB[e * 6 + 0] = (ez == 0) ? 1 : 0;
          EToB[e * 6 + 1] = (ey == 0) ? 1 : 0;
          EToB[e * 6 + 2] = (ex == nx - 1) ? 1 : 0;
          EToB[e * 6 + 3] = (ey == ny - 1) ? 1 : 0;
          EToB[e * 6 + 4] = (ex == 0) ? 1 : 0;
          EToB[e * 6 + 5] = (ez == nz - 1) ? 1 : 0;
        }
      }
}

// ---- geometric factors: restates kernels/mesh/geometricFactorsHex3D.okl:27-143 ----
// One block per element, one thread per (i,j), loop over k planes; writes the
// seven ggeo entries [e][7][Np].
template <int NQ>
__global__ void __launch_bounds__(NQ * NQ)
geofacs_kernel(int E, const double *__restrict__ Dg, const double *__restrict__ wg,
               const double *__restrict__ x, const double *__restrict__ y,
               const double *__restrict__ z, double *__restrict__ ggeo)
{
  constexpr int NP = NQ * NQ * NQ;
  __shared__ double s_D[NQ][NQ], s_x[NQ][NQ], s_y[NQ][NQ], s_z[NQ][NQ], s_w[NQ];
  const int e = blockIdx.x;
  const int i = threadIdx.x % NQ, j = threadIdx.x / NQ;
  s_D[j][i] = Dg[j * NQ + i];
  if (j == 0) s_w[i] = wg[i];
  double rx_[NQ], ry_[NQ], rz_[NQ];
  const int64_t base = (int64_t)e * NP + j * NQ + i;
#pragma unroll
  for (int l = 0; l < NQ; ++l) {
    rx_[l] = x[base + l * NQ * NQ];
    ry_[l] = y[base + l * NQ * NQ];
    rz_[l] = z[base + l * NQ * NQ];
  }
  for (int k = 0; k < NQ; ++k) {
    __syncthreads();
    s_x[j][i] = rx_[k]; s_y[j][i] = ry_[k]; s_z[j][i] = rz_[k];
    __syncthreads();
    double xr = 0, yr = 0, zr = 0, xs = 0, ys = 0, zs = 0, xt = 0, yt = 0, zt = 0;
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      const double Dim = s_D[i][m], Djm = s_D[j][m], Dkm = s_D[k][m];
      xr += Dim * s_x[j][m]; xs += Djm * s_x[m][i]; xt += Dkm * rx_[m];
      yr += Dim * s_y[j][m]; ys += Djm * s_y[m][i]; yt += Dkm * ry_[m];
      zr += Dim * s_z[j][m]; zs += Djm * s_z[m][i]; zt += Dkm * rz_[m];
    }
    const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
    const double Jinv = 1.0 / J;
    const double JW = J * s_w[i] * s_w[j] * s_w[k];
    const double rx = (ys * zt - zs * yt) * Jinv, ry = -(xs * zt - zs * xt) * Jinv, rz = (xs * yt - ys * xt) * Jinv;
    const double sx = -(yr * zt - zr * yt) * Jinv, sy = (xr * zt - zr * xt) * Jinv, sz = -(xr * yt - yr * xt) * Jinv;
    const double tx = (yr * zs - zr * ys) * Jinv, ty = -(xr * zs - zr * xs) * Jinv, tz = (xr * ys - yr * xs) * Jinv;
    const int n = k * NQ * NQ + j * NQ + i;
    double *G = ggeo + (int64_t)e * 7 * NP + n;
    G[0 * NP] = JW * (rx * rx + ry * ry + rz * rz);
    G[1 * NP] = JW * (rx * sx + ry * sy + rz * sz);
    G[4 * NP] = JW * (rx * tx + ry * ty + rz * tz);
    G[2 * NP] = JW * (sx * sx + sy * sy + sz * sz);
    G[3 * NP] = JW * (sx * tx + sy * ty + sz * tz);
    G[5 * NP] = JW * (tx * tx + ty * ty + tz * tz);
    G[6 * NP] = JW;
  }
}

void launch_geofacs(int N, int E, const double *D, const double *w, const double *x,
                    const double *y, const double *z, double *ggeo, hipStream_t s)
{
  if (E <= 0) return;
#define GF(NQ) case NQ: geofacs_kernel<NQ><<<E, NQ * NQ, 0, s>>>(E, D, w, x, y, z, ggeo); break;
  switch (N + 1) {
    GF(2) GF(3) GF(4) GF(5) GF(6) GF(7) GF(8) GF(9) GF(10)
    default: throw Error(NRS_ERR_UNSUPPORTED, "geometric factors: order N must be 1..9");
  }
#undef GF
  NRS_LAUNCH_CHECK();
}

bool trilinear_corners(int N, int E, const double *x, const double *y, const double *z,
                       std::vector<double> &corners)
{
  if (N < 1 || E <= 0) return false;
  std::vector<double> r, w;
  gll_nodes(N, r, w);
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  corners.assign((size_t)E * 24, 0.0);
  const double *X[3] = {x, y, z};
  for (int64_t e = 0; e < E; ++e) {
    double *C = corners.data() + e * 24;
    for (int d = 0; d < 3; ++d)
      for (int c = 0; c < 8; ++c) {
        const int i = (c & 1) * N, j = ((c >> 1) & 1) * N, k = ((c >> 2) & 1) * N;
        C[8 * d + c] = X[d][e * Np + i + Nq * (j + Nq * k)];
      }
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = hi[d] = C[8 * d];
      for (int c = 1; c < 8; ++c) { lo[d] = std::min(lo[d], C[8 * d + c]); hi[d] = std::max(hi[d], C[8 * d + c]); }
    }
    const double size = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (!(size > 0.0)) return false;
    const double tol = 1e-12 * size;
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const double a[2] = {0.5 * (1 - r[i]), 0.5 * (1 + r[i])};
          const double b[2] = {0.5 * (1 - r[j]), 0.5 * (1 + r[j])};
          const double c2[2] = {0.5 * (1 - r[k]), 0.5 * (1 + r[k])};
          for (int d = 0; d < 3; ++d) {
            double v = 0.0;
            for (int c = 0; c < 8; ++c) v += a[c & 1] * b[(c >> 1) & 1] * c2[(c >> 2) & 1] * C[8 * d + c];
            if (std::fabs(v - X[d][e * Np + i + Nq * (j + Nq * k)]) > tol) return false;
          }
        }
  }
  return true;
}

}  // namespace nrs
