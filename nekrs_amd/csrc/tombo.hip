// tombo.hip -- pressure right-hand side of the TOMBO splitting on the device.
//
// tombo::pressureSolve (src/navierStokes/tombo.cpp:11-121) runs, on
// [3][fieldOffset] velocity-like fields:
//   W    = JW curl(Ue);  W = gs(W) / gs(JW)                     (curlHex3D, oogs, axmyVector)
//   rhsU = BF - nu (JW curl(W) - 4/3 JW grad(div))             (curlHex3D, gradientVolumeHex3D,
//          (BF alone when the viscous term is off)              pressureRhsHex3D)
//   rhsU = gs(rhsU) / gs(JW)
//   rhs  = wdiv(rhsU) + g0/dt JW div                           (wDivergenceVolumeHex3D, pressureAddQtl)
//   rhs += -WSJ (n . flux) on inflow / interface / symmetry faces (divergenceSurfaceHex3D)
// Here the scalings by 1/gs(JW) are applied where the next kernel loads the
// field (same products, no extra pass), the second curl, the gradient and the
// combination are one kernel, the weak divergence and the mass term are one
// kernel, and the surface kernel visits only elements with a flux face.
#include <cmath>
#include <memory>

#include "elliptic.hpp"
#include "mesh.hpp"
#include "tombo.hpp"
#include "trilin.hpp"
#include "tune.hpp"

nrs_tombo::~nrs_tombo()
{
  if (ogs) delete ogs;
  if (ogsOp) delete ogsOp;
}

namespace nrs {

namespace {

constexpr int NV = 10;   // vgeo entries per node

// Volume factors (geometricFactorsHex3D.okl:27-143: rx..tz, JW) and, for
// elements with a flux face, the face normals and WSJ of those faces
// (surfaceGeometricFactorsHex3D.okl: n = +-grad(r|s|t) / |.|, sJ = |.| J,
// WSJ = sJ w_a w_b).  Faces: 0 t=-1, 1 s=-1, 2 r=+1, 3 s=+1, 4 r=-1, 5 t=+1.
template <int NQ>
__global__ void __launch_bounds__(NQ * NQ)
vgeo_kernel(int E, const double *__restrict__ Dg, const double *__restrict__ wg,
            const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
            const int *__restrict__ bslot, double *__restrict__ vgeo, double *__restrict__ bgeo)
{
  constexpr int NP = NQ * NQ * NQ, NFP = NQ * NQ;
  __shared__ double s_D[NQ][NQ], s_x[NQ][NQ], s_y[NQ][NQ], s_z[NQ][NQ], s_w[NQ];
  const int e = blockIdx.x;
  const int i = threadIdx.x % NQ, j = threadIdx.x / NQ;
  s_D[j][i] = Dg[j * NQ + i];
  if (j == 0) s_w[i] = wg[i];
  double px[NQ], py[NQ], pz[NQ];
  const int64_t base = (int64_t)e * NP + j * NQ + i;
#pragma unroll
  for (int l = 0; l < NQ; ++l) {
    px[l] = x[base + l * NFP];
    py[l] = y[base + l * NFP];
    pz[l] = z[base + l * NFP];
  }
  const int b = bslot[e];
  for (int k = 0; k < NQ; ++k) {
    __syncthreads();
    s_x[j][i] = px[k]; s_y[j][i] = py[k]; s_z[j][i] = pz[k];
    __syncthreads();
    double xr = 0, yr = 0, zr = 0, xs = 0, ys = 0, zs = 0, xt = 0, yt = 0, zt = 0;
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      const double Dim = s_D[i][m], Djm = s_D[j][m], Dkm = s_D[k][m];
      xr += Dim * s_x[j][m]; xs += Djm * s_x[m][i]; xt += Dkm * px[m];
      yr += Dim * s_y[j][m]; ys += Djm * s_y[m][i]; yt += Dkm * py[m];
      zr += Dim * s_z[j][m]; zs += Djm * s_z[m][i]; zt += Dkm * pz[m];
    }
    const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
    const double Jinv = 1.0 / J;
    const double JW = J * s_w[i] * s_w[j] * s_w[k];
    const double rx = (ys * zt - zs * yt) * Jinv, ry = -(xs * zt - zs * xt) * Jinv, rz = (xs * yt - ys * xt) * Jinv;
    const double sx = -(yr * zt - zr * yt) * Jinv, sy = (xr * zt - zr * xt) * Jinv, sz = -(xr * yt - yr * xt) * Jinv;
    const double tx = (yr * zs - zr * ys) * Jinv, ty = -(xr * zs - zr * xs) * Jinv, tz = (xr * ys - yr * xs) * Jinv;
    const int n = k * NFP + j * NQ + i;
    double *G = vgeo + (int64_t)e * NV * NP + n;
    G[0 * NP] = rx; G[1 * NP] = ry; G[2 * NP] = rz;
    G[3 * NP] = sx; G[4 * NP] = sy; G[5 * NP] = sz;
    G[6 * NP] = tx; G[7 * NP] = ty; G[8 * NP] = tz;
    G[9 * NP] = JW;
    if (b >= 0) {
      auto put = [&](int f, int nf, double nx, double ny, double nz) {
        double sJ = sqrt(nx * nx + ny * ny + nz * nz);
        const double inv = 1.0 / sJ;
        nx *= inv; ny *= inv; nz *= inv;
        sJ *= J;
        double *g = bgeo + (((int64_t)b * 6 + f) * NFP + nf) * 4;
        g[0] = nx; g[1] = ny; g[2] = nz;
        g[3] = sJ * s_w[nf % NQ] * s_w[nf / NQ];
      };
      if (k == 0) put(0, i + j * NQ, -tx, -ty, -tz);
      if (j == 0) put(1, i + k * NQ, -sx, -sy, -sz);
      if (i == NQ - 1) put(2, j + k * NQ, rx, ry, rz);
      if (j == NQ - 1) put(3, i + k * NQ, sx, sy, sz);
      if (i == 0) put(4, j + k * NQ, -rx, -ry, -rz);
      if (k == NQ - 1) put(5, i + j * NQ, tx, ty, tz);
    }
  }
}

// JW curl(s U) (curlHex3D.okl), s = invLMM when PRESCALE; with RHS also
// JW grad(div) (gradientVolumeHex3D.okl) and the combination of
// pressureRhsHex3D.okl: out = BF - mue irho (JW curl - 4/3 JW grad(div)).
// Thread (i,j) per element, k-pencils of the fields in registers.  FX (N = 7):
// four elements per block, and the interior nodes of a face i = N / i = 0 shared
// by consecutive elements with fx[e] = 1 get the pair sum here (the lower
// element's value first, as the gather-scatter adds them), so the following
// gather-scatter leaves them out.
template <int NQ, bool PRESCALE, bool RHS, bool PF, bool FX = false, bool TRI = false>
__global__ void __launch_bounds__(NQ * NQ * (FX ? 4 : 1))
curl_kernel(int E, const double *__restrict__ vgeo, const double *__restrict__ Dg, int64_t offIn,
            int64_t offOut, const double *__restrict__ invLMM, const double *__restrict__ U,
            const double *__restrict__ div, const double *__restrict__ BF, int64_t offBF,
            const double *__restrict__ mue, const double *__restrict__ irho, double *__restrict__ out,
            const uint8_t *__restrict__ fx = nullptr, const double *__restrict__ S = nullptr,
            const double *__restrict__ corners = nullptr, const double *__restrict__ zw = nullptr)
{
  constexpr int NP = NQ * NQ * NQ, NFP = NQ * NQ;
  constexpr int NF = RHS ? 4 : 3;                 // u, v, w (+ div)
  constexpr int NE = FX ? 4 : 1;
  static_assert(!FX || NQ * NQ == 64, "face fusion: one element per wave64");
  __shared__ double s_D[NQ][NQ], s_f[NE][NF][NQ][NQ];
  __shared__ double s_x[FX ? NE : 1][2][3][NQ][NQ];   // face pencils [element][i = 0, N][field][j][k]
  const int t = threadIdx.x;
  const int le = FX ? t / NFP : 0;
  const int i = (t % NFP) % NQ, j = (t % NFP) / NQ;
  const int64_t e = (int64_t)blockIdx.x * NE + le;
  const bool active = e < E;
  if (le == 0) s_D[j][i] = Dg[j * NQ + i];
  __syncthreads();
  auto sync_plane = [&]() {
    if constexpr (FX) {            // each element is one wave with its own slabs
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      __syncthreads();
    }
  };
  const bool faceLane = FX && (i == 0 || i == NQ - 1) && j > 0 && j < NQ - 1;
  const bool withNext = faceLane && active && i == NQ - 1 && le + 1 < NE && e + 1 < E && fx[e];
  const bool withPrev = faceLane && active && i == 0 && le >= 1 && fx[e - 1];
  const int64_t base = e * NP + j * NQ + i;
  if (active) {
    double pf[NF][NQ];
#pragma unroll
    for (int l = 0; l < NQ; ++l) {
      const int64_t id = base + l * NFP;
      const double sc = PRESCALE ? invLMM[id] : 1.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) pf[c][l] = PRESCALE ? sc * U[id + c * offIn] : U[id + c * offIn];
      if (RHS) pf[NF - 1][l] = div[id];
    }
    // per-plane inputs: geometry (and BF, mue, irho), at use or one plane ahead;
    // TRI: the geometry formed from the element's corners
    constexpr int NX = RHS ? NV + 5 : NV;
    const double *Ge = TRI ? nullptr : vgeo + e * NV * NP + j * NQ + i;
    TriGeo<double> tri;
    if constexpr (TRI) tri.init(corners + e * 24, zw[i], zw[j], zw[NQ + i], zw[NQ + j]);
    auto load_plane = [&](int k, double (&g)[NX]) {
      if constexpr (TRI) {
        double v[NV];
        tri.at(zw[k], zw[NQ + k], v);
#pragma unroll
        for (int f = 0; f < NV; ++f) g[f] = v[f];
      } else {
#pragma unroll
        for (int f = 0; f < NV; ++f) g[f] = __builtin_nontemporal_load(Ge + f * NP + k * NFP);
      }
      if (RHS) {
        const int64_t id = base + k * NFP;
        g[NV + 0] = BF[id];
        g[NV + 1] = BF[id + offBF];
        g[NV + 2] = BF[id + 2 * offBF];
        g[NV + 3] = mue[id];
        g[NV + 4] = irho[id];
      }
    };
    double gc[NX], gn[NX];
    if (PF) load_plane(0, gc);
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      if (PF && k + 1 < NQ) load_plane(k + 1, gn);
      sync_plane();
#pragma unroll
      for (int c = 0; c < NF; ++c) s_f[le][c][j][i] = pf[c][k];
      sync_plane();
      double dr[NF], ds[NF], dt[NF];
#pragma unroll
      for (int c = 0; c < NF; ++c) { dr[c] = 0.0; ds[c] = 0.0; dt[c] = 0.0; }
#pragma unroll
      for (int n = 0; n < NQ; ++n) {
        const double Dr = s_D[i][n], Ds = s_D[j][n], Dt = s_D[k][n];
#pragma unroll
        for (int c = 0; c < NF; ++c) {
          dr[c] += Dr * s_f[le][c][j][n];
          ds[c] += Ds * s_f[le][c][n][i];
          dt[c] += Dt * pf[c][n];
        }
      }
      const int64_t id = base + k * NFP;
      if (!PF) load_plane(k, gc);
      const double rx = gc[0], ry = gc[1], rz = gc[2];
      const double sx = gc[3], sy = gc[4], sz = gc[5];
      const double tx = gc[6], ty = gc[7], tz = gc[8];
      const double JW = gc[9];
      const double uy = ry * dr[0] + sy * ds[0] + ty * dt[0];
      const double uz = rz * dr[0] + sz * ds[0] + tz * dt[0];
      const double vx = rx * dr[1] + sx * ds[1] + tx * dt[1];
      const double vz = rz * dr[1] + sz * ds[1] + tz * dt[1];
      const double wx = rx * dr[2] + sx * ds[2] + tx * dt[2];
      const double wy = ry * dr[2] + sy * ds[2] + ty * dt[2];
      double c0 = (wy - vz) * JW, c1 = (uz - wx) * JW, c2 = (vx - uy) * JW;
      if (RHS && S) {           // VELOCITY STRESSFORMULATION: pressureStressHex3D's OUT -= JW (...)
        c0 -= S[id];
        c1 -= S[id + offOut];
        c2 -= S[id + 2 * offOut];
      }
      if (RHS) {
        constexpr int P = NF - 1;
        const double gx = JW * (rx * dr[P] + sx * ds[P] + tx * dt[P]);
        const double gy = JW * (ry * dr[P] + sy * ds[P] + ty * dt[P]);
        const double gz = JW * (rz * dr[P] + sz * ds[P] + tz * dt[P]);
        const double factor = 4. / 3;
        const double nu = gc[NV + 3] * gc[NV + 4];
        c0 = gc[NV + 0] - nu * (c0 - factor * gx);
        c1 = gc[NV + 1] - nu * (c1 - factor * gy);
        c2 = gc[NV + 2] - nu * (c2 - factor * gz);
      }
      const bool inner = k > 0 && k < NQ - 1;
      if (faceLane && inner) {
        const int side = (i == 0) ? 0 : 1;
        s_x[le][side][0][j][k] = c0;
        s_x[le][side][1][j][k] = c1;
        s_x[le][side][2][j][k] = c2;
      }
      if (!((withNext || withPrev) && inner)) {
        __builtin_nontemporal_store(c0, out + id);
        __builtin_nontemporal_store(c1, out + id + offOut);
        __builtin_nontemporal_store(c2, out + id + 2 * offOut);
      }
      if (PF && k + 1 < NQ) {
#pragma unroll
        for (int f = 0; f < NX; ++f) gc[f] = gn[f];
      }
    }
  }
  if constexpr (FX) {
    __syncthreads();
    if (withNext || withPrev) {
      const int lo = withNext ? le : le - 1;      // the lower element of the pair
#pragma unroll
      for (int k = 1; k < NQ - 1; ++k) {
        const int64_t id = base + k * NFP;
#pragma unroll
        for (int c = 0; c < 3; ++c)
          out[id + c * offOut] = s_x[lo][1][c][j][k] + s_x[lo + 1][0][c][j][k];
      }
    }
  }
}

// The stress-formulation term (pressureStressHex3D.okl): with g = grad(mue) / mue,
//   S = JW (2 sym(grad Ue) g - 2/3 div g),
// written to S [3][Nlocal]; the right-hand-side kernel subtracts it from the
// second curl.  Thread (i,j) per element with k-pencils of u, v, w, mue in
// registers; the r/s derivatives through one LDS plane per field.
template <int NQ, bool TRI>
__global__ void __launch_bounds__(NQ * NQ)
stress_kernel(int E, const double *__restrict__ vgeo, const double *__restrict__ Dg, int64_t off,
              const double *__restrict__ U, const double *__restrict__ mue, const double *__restrict__ div,
              double *__restrict__ S, const double *__restrict__ corners, const double *__restrict__ zw)
{
  constexpr int NP = NQ * NQ * NQ, NFP = NQ * NQ;
  __shared__ double s_D[NQ][NQ], s_f[4][NQ][NQ];
  const int e = blockIdx.x;
  const int i = threadIdx.x % NQ, j = threadIdx.x / NQ;
  s_D[j][i] = Dg[j * NQ + i];
  const int64_t base = (int64_t)e * NP + j * NQ + i;
  TriGeo<double> tri;
  if constexpr (TRI) tri.init(corners + (int64_t)e * 24, zw[i], zw[j], zw[NQ + i], zw[NQ + j]);
  double pf[4][NQ];
#pragma unroll
  for (int l = 0; l < NQ; ++l) {
    const int64_t id = base + l * NFP;
#pragma unroll
    for (int c = 0; c < 3; ++c) pf[c][l] = U[id + c * off];
    pf[3][l] = mue[id];
  }
  for (int k = 0; k < NQ; ++k) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) s_f[c][j][i] = pf[c][k];
    __syncthreads();
    double dr[4], ds[4], dt[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) { dr[c] = 0.0; ds[c] = 0.0; dt[c] = 0.0; }
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
      const double Dr = s_D[i][n], Ds = s_D[j][n], Dt = s_D[k][n];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        dr[c] += Dr * s_f[c][j][n];
        ds[c] += Ds * s_f[c][n][i];
        dt[c] += Dt * pf[c][n];
      }
    }
    const int64_t id = base + k * NFP;
    double v[NV];
    if constexpr (TRI) {
      tri.at(zw[k], zw[NQ + k], v);
    } else {
      const double *G = vgeo + (int64_t)e * NV * NP + k * NFP + j * NQ + i;
#pragma unroll
      for (int f = 0; f < NV; ++f) v[f] = G[f * NP];
    }
    const double rx = v[0], ry = v[1], rz = v[2];
    const double sx = v[3], sy = v[4], sz = v[5];
    const double tx = v[6], ty = v[7], tz = v[8];
    const double JW = v[9];
    double gx[4], gy[4], gz[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      gx[c] = rx * dr[c] + sx * ds[c] + tx * dt[c];
      gy[c] = ry * dr[c] + sy * ds[c] + ty * dt[c];
      gz[c] = rz * dr[c] + sz * ds[c] + tz * dt[c];
    }
    const double im = 1.0 / pf[3][k];
    const double mx = gx[3] * im, my = gy[3] * im, mz = gz[3] * im;
    const double w1 = mx * 2 * gx[0] + my * (gy[0] + gx[1]) + mz * (gz[0] + gx[2]);
    const double w2 = mx * (gy[0] + gx[1]) + my * 2 * gy[1] + mz * (gz[1] + gy[2]);
    const double w3 = mx * (gz[0] + gx[2]) + my * (gz[1] + gy[2]) + mz * 2 * gz[2];
    const double d = div[id];
    const double scale = -2. / 3;
    S[id] = JW * (w1 + scale * d * mx);
    S[id + NP * (int64_t)E] = JW * (w2 + scale * d * my);
    S[id + 2 * NP * (int64_t)E] = JW * (w3 + scale * d * mz);
  }
}

// rhs = wdiv(invLMM F) + lambda JW div   (wDivergenceVolumeHex3D.okl + pressureAddQtl.okl)
template <int NQ, bool TRI>
__global__ void __launch_bounds__(NQ * NQ)
wdiv_kernel(int E, const double *__restrict__ vgeo, const double *__restrict__ Dg, int64_t off,
            const double *__restrict__ invLMM, const double *__restrict__ F,
            const double *__restrict__ div, double lambda, double *__restrict__ rhs,
            const double *__restrict__ corners, const double *__restrict__ zw)
{
  constexpr int NP = NQ * NQ * NQ, NFP = NQ * NQ;
  __shared__ double s_D[NQ][NQ], s_u[NQ][NQ], s_v[NQ][NQ], s_w[NQ][NQ];
  const int e = blockIdx.x;
  const int i = threadIdx.x % NQ, j = threadIdx.x / NQ;
  s_D[j][i] = Dg[j * NQ + i];
  double r_div[NQ], jw[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) r_div[k] = 0.0;
  const int64_t base = (int64_t)e * NP + j * NQ + i;
  TriGeo<double> tri;
  if constexpr (TRI) tri.init(corners + (int64_t)e * 24, zw[i], zw[j], zw[NQ + i], zw[NQ + j]);
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int64_t id = base + k * NFP;
    double v[NV];
    if constexpr (TRI) {
      tri.at(zw[k], zw[NQ + k], v);
    } else {
      const double *G = vgeo + (int64_t)e * NV * NP + k * NFP + j * NQ + i;
#pragma unroll
      for (int f = 0; f < NV; ++f) v[f] = G[f * NP];
    }
    const double sc = invLMM[id];
    const double Un = sc * F[id], Vn = sc * F[id + off], Wn = sc * F[id + 2 * off];
    const double JW = v[9];
    jw[k] = JW;
    __syncthreads();
    s_u[j][i] = JW * (v[0] * Un + v[1] * Vn + v[2] * Wn);
    s_v[j][i] = JW * (v[3] * Un + v[4] * Vn + v[5] * Wn);
    s_w[j][i] = JW * (v[6] * Un + v[7] * Vn + v[8] * Wn);
    __syncthreads();
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
      r_div[k] += s_D[n][i] * s_u[j][n];
      r_div[k] += s_D[n][j] * s_v[n][i];
      r_div[n] += s_D[k][n] * s_w[j][i];
    }
  }
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int64_t id = base + k * NFP;
    rhs[id] = r_div[k] + lambda * jw[k] * div[id];
  }
}

// divergenceSurfaceHex3D.okl over the elements with a flux face: per face
// node -WSJ (coef n.U) on inflow/interface faces, -WSJ (n.F) on symmetry
// faces (F = invLMM rhsU), summed per node in the face order (0,5), (1,3),
// (2,4) and added to rhs.
template <int NQ>
__global__ void __launch_bounds__(NQ * NQ)
surface_kernel(const int *__restrict__ bElem, const int *__restrict__ bType,
               const double *__restrict__ bGeo, double coef, int64_t offU, int64_t offF,
               const double *__restrict__ invLMM, const double *__restrict__ F,
               const double *__restrict__ U, double *__restrict__ rhs)
{
  constexpr int NP = NQ * NQ * NQ, NFP = NQ * NQ;
  __shared__ double s_flux[2][NQ][NQ];
  const int b = blockIdx.x;
  const int e = bElem[b];
  const int a = threadIdx.x % NQ, c = threadIdx.x / NQ;   // face-local (a, c)
  double r_div[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) r_div[k] = 0.0;
  auto term = [&](int f, int m) -> double {
    const int nf = a + c * NQ;
    const double *g = bGeo + (((int64_t)b * 6 + f) * NFP + nf) * 4;
    const int64_t idM = (int64_t)e * NP + m;
    const int type = bType[b * 6 + f];
    double flux = 0.0;
    if (type == BC_V || type == BC_INT) {
      flux = coef * (g[0] * U[idM] + g[1] * U[idM + offU] + g[2] * U[idM + 2 * offU]);
    } else if (type == BC_SYMX || type == BC_SYMY || type == BC_SYMZ || type == BC_SYM) {
      const double sc = invLMM[idM];
      flux = g[0] * (sc * F[idM]) + g[1] * (sc * F[idM + offF]) + g[2] * (sc * F[idM + 2 * offF]);
    }
    return -g[3] * flux;
  };
  // faces 0 (k=0) and 5 (k=N): node (i=a, j=c)
  s_flux[0][c][a] = term(0, a + c * NQ);
  s_flux[1][c][a] = term(5, a + c * NQ + (NQ - 1) * NFP);
  __syncthreads();
  // thread (i, j) = (a, c) from here on when accumulating
  r_div[0] += s_flux[0][c][a];
  r_div[NQ - 1] += s_flux[1][c][a];
  __syncthreads();
  // faces 1 (j=0) and 3 (j=N): node (i=a, k=c)
  s_flux[0][c][a] = term(1, a + c * NFP);
  s_flux[1][c][a] = term(3, a + (NQ - 1) * NQ + c * NFP);
  __syncthreads();
  if (c == 0) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) r_div[k] += s_flux[0][k][a];
  }
  if (c == NQ - 1) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) r_div[k] += s_flux[1][k][a];
  }
  __syncthreads();
  // faces 2 (i=N) and 4 (i=0): node (j=a, k=c)
  s_flux[0][c][a] = term(2, NQ - 1 + a * NQ + c * NFP);
  s_flux[1][c][a] = term(4, a * NQ + c * NFP);
  __syncthreads();
  // thread (i=a, j=c): face 2 for i = N, face 4 for i = 0
  if (a == NQ - 1) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) r_div[k] += s_flux[0][k][c];
  }
  if (a == 0) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) r_div[k] += s_flux[1][k][c];
  }
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int64_t id = (int64_t)e * NP + k * NFP + c * NQ + a;
    rhs[id] += r_div[k];
  }
}

template <template <int> class L, typename... A>
void by_order(int N, A... a)
{
  switch (N + 1) {
    case 2: L<2>::run(a...); break;
    case 3: L<3>::run(a...); break;
    case 4: L<4>::run(a...); break;
    case 5: L<5>::run(a...); break;
    case 6: L<6>::run(a...); break;
    case 7: L<7>::run(a...); break;
    case 8: L<8>::run(a...); break;
    case 9: L<9>::run(a...); break;
    case 10: L<10>::run(a...); break;
    default: throw Error(NRS_ERR_UNSUPPORTED, "tombo: order N must be 1..9");
  }
}

template <int NQ> struct VgeoL {
  static void run(int E, const double *D, const double *w, const double *x, const double *y,
                  const double *z, const int *bslot, double *vgeo, double *bgeo, hipStream_t s)
  {
    vgeo_kernel<NQ><<<E, NQ * NQ, 0, s>>>(E, D, w, x, y, z, bslot, vgeo, bgeo);
  }
};
template <int NQ> struct CurlL {
  static void run(const nrs_tombo *t, int64_t offIn, int64_t offOut, bool prescale, bool pf,
                  const double *U, double *W, hipStream_t s, const uint8_t *fx = nullptr)
  {
    if (t->tri) go<true>(t, offIn, offOut, prescale, pf, U, W, s, fx);
    else go<false>(t, offIn, offOut, prescale, pf, U, W, s, fx);
  }
  template <bool TRI>
  static void go(const nrs_tombo *t, int64_t offIn, int64_t offOut, bool prescale, bool pf,
                 const double *U, double *W, hipStream_t s, const uint8_t *fx)
  {
    const double *G = t->vgeo.get(), *D = t->D.get(), *il = t->invLMM.get();
    const double *C = t->corners.get(), *zw = t->zw.get();
    if constexpr (NQ == 8) {
      if (fx) {
        const dim3 g((t->E + 3) / 4), b(256);
#define CURLF(P, F) curl_kernel<NQ, P, false, F, true, TRI><<<g, b, 0, s>>>(t->E, G, D, offIn, offOut, \
    P ? il : nullptr, U, nullptr, nullptr, 0, nullptr, nullptr, W, fx, nullptr, C, zw)
        if (prescale) { if (pf) CURLF(true, true); else CURLF(true, false); }
        else { if (pf) CURLF(false, true); else CURLF(false, false); }
#undef CURLF
        return;
      }
    }
    const dim3 g(t->E), b(NQ * NQ);
#define CURL(P, F) curl_kernel<NQ, P, false, F, false, TRI><<<g, b, 0, s>>>(t->E, G, D, offIn, offOut, \
    P ? il : nullptr, U, nullptr, nullptr, 0, nullptr, nullptr, W, nullptr, nullptr, C, zw)
    if (prescale) { if (pf) CURL(true, true); else CURL(true, false); }
    else { if (pf) CURL(false, true); else CURL(false, false); }
#undef CURL
  }
};
template <int NQ> struct RhsL {
  static void run(const nrs_tombo *t, bool pf, const double *W, const double *div, const double *BF,
                  int64_t offBF, const double *mue, const double *irho, double *out, hipStream_t s,
                  const uint8_t *fx = nullptr, const double *S = nullptr)
  {
    if (t->tri) go<true>(t, pf, W, div, BF, offBF, mue, irho, out, s, fx, S);
    else go<false>(t, pf, W, div, BF, offBF, mue, irho, out, s, fx, S);
  }
  template <bool TRI>
  static void go(const nrs_tombo *t, bool pf, const double *W, const double *div, const double *BF,
                 int64_t offBF, const double *mue, const double *irho, double *out, hipStream_t s,
                 const uint8_t *fx, const double *S)
  {
    const double *G = t->vgeo.get(), *D = t->D.get(), *il = t->invLMM.get();
    const double *C = t->corners.get(), *zw = t->zw.get();
    const int64_t L = t->Nlocal;
    if constexpr (NQ == 8) {
      if (fx) {
        const dim3 g((t->E + 3) / 4), b(256);
        if (pf)
          curl_kernel<NQ, true, true, true, true, TRI><<<g, b, 0, s>>>(t->E, G, D, L, L, il, W, div, BF, offBF,
                                                                       mue, irho, out, fx, S, C, zw);
        else
          curl_kernel<NQ, true, true, false, true, TRI><<<g, b, 0, s>>>(t->E, G, D, L, L, il, W, div, BF, offBF,
                                                                        mue, irho, out, fx, S, C, zw);
        return;
      }
    }
    const dim3 g(t->E), b(NQ * NQ);
    if (pf)
      curl_kernel<NQ, true, true, true, false, TRI><<<g, b, 0, s>>>(t->E, G, D, L, L, il, W, div, BF, offBF, mue,
                                                                    irho, out, nullptr, S, C, zw);
    else
      curl_kernel<NQ, true, true, false, false, TRI><<<g, b, 0, s>>>(t->E, G, D, L, L, il, W, div, BF, offBF,
                                                                     mue, irho, out, nullptr, S, C, zw);
  }
};
template <int NQ> struct StressL {
  static void run(const nrs_tombo *t, int64_t off, const double *U, const double *mue, const double *div,
                  double *S, hipStream_t s)
  {
    if (t->tri)
      stress_kernel<NQ, true><<<t->E, NQ * NQ, 0, s>>>(t->E, t->vgeo.get(), t->D.get(), off, U, mue, div, S,
                                                       t->corners.get(), t->zw.get());
    else
      stress_kernel<NQ, false><<<t->E, NQ * NQ, 0, s>>>(t->E, t->vgeo.get(), t->D.get(), off, U, mue, div, S,
                                                        nullptr, nullptr);
  }
};
template <int NQ> struct WdivL {
  static void run(const nrs_tombo *t, const double *F, const double *div, double lambda,
                  double *rhs, hipStream_t s)
  {
    if (t->tri)
      wdiv_kernel<NQ, true><<<t->E, NQ * NQ, 0, s>>>(t->E, t->vgeo.get(), t->D.get(), t->Nlocal,
                                                     t->invLMM.get(), F, div, lambda, rhs, t->corners.get(),
                                                     t->zw.get());
    else
      wdiv_kernel<NQ, false><<<t->E, NQ * NQ, 0, s>>>(t->E, t->vgeo.get(), t->D.get(), t->Nlocal,
                                                      t->invLMM.get(), F, div, lambda, rhs, nullptr, nullptr);
  }
};
template <int NQ> struct SurfL {
  static void run(const nrs_tombo *t, double coef, int64_t offU, const double *U, double *rhs,
                  hipStream_t s)
  {
    if (t->Nb == 0) return;
    surface_kernel<NQ><<<t->Nb, NQ * NQ, 0, s>>>(t->bElem.get(), t->bType.get(), t->bGeo.get(), coef,
                                                  offU, t->Nlocal, t->invLMM.get(), t->F.get(), U, rhs);
  }
};

__global__ void reciprocal_kernel(int64_t n, double *v)
{
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = 1.0 / v[i];
}

// invLMM <- JW (the vgeo entry 9 of every element)
__global__ void copy_jw_kernel(int E, int Np, const double *__restrict__ vgeo, double *__restrict__ out)
{
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= (int64_t)E * Np) return;
  const int64_t e = idx / Np, n = idx % Np;
  out[idx] = vgeo[e * NV * Np + 9 * Np + n];
}

bool flux_face(int t)
{
  return t == BC_V || t == BC_INT || t == BC_SYMX || t == BC_SYMY || t == BC_SYMZ || t == BC_SYM;
}

// Time the two load schedules of each curl form on this mesh with patterned
// inputs, require results within 400 eps of each other (the same arithmetic;
// the compiler may contract multiply-adds differently), record the faster
// (type 0: JW curl, type 1: the fused right-hand-side kernel).
void autotune(nrs_tombo *t, hipStream_t s)
{
  const int64_t L = t->Nlocal;
  DevBuf<double> U(3 * L), div(L), BF(3 * L), mue(L), irho(L), out2(3 * L);
  fill_pattern<double>(3 * L, U.get(), s);
  fill_pattern<double>(L, div.get(), s);
  fill_pattern<double>(3 * L, BF.get(), s);
  fill_pattern<double>(L, mue.get(), s);
  fill_pattern<double>(L, irho.get(), s);
  for (int type = 0; type < 2; ++type) {
    double ms[2];
    double *outs[2] = {t->W.get(), out2.get()};
    for (int v = 0; v < 2; ++v) {
      auto f = [&]() {
        if (type == 0) by_order<CurlL>(t->N, t, L, L, false, v == 1, U.get(), outs[v], s, t->fx.get());
        else by_order<RhsL>(t->N, t, v == 1, U.get(), div.get(), BF.get(), L, mue.get(), irho.get(), outs[v], s,
                            t->fx.get());
      };
      ms[v] = time_launches(f, s);
    }
    NRS_LAUNCH_CHECK();
    std::vector<double> a(3 * L), b(3 * L);
    NRS_HIP(hipMemcpyAsync(a.data(), outs[0], 3 * L * sizeof(double), hipMemcpyDeviceToHost, s));
    NRS_HIP(hipMemcpyAsync(b.data(), outs[1], 3 * L * sizeof(double), hipMemcpyDeviceToHost, s));
    NRS_HIP(hipStreamSynchronize(s));
    double amax = 0.0, dmax = 0.0;
    for (int64_t n = 0; n < 3 * L; ++n) {
      amax = std::max(amax, std::fabs(a[n]));
      dmax = std::max(dmax, std::fabs(a[n] - b[n]));
    }
    const bool same = dmax <= 400 * 2.220446049250313e-16 * amax;   // NaN fails
    set_tuned(TUNE_TOMBO, t->N, type, t->E, (same && ms[1] < ms[0]) ? 2 : 1);
  }
  // the face fusion stays only where the whole right-hand side times faster with it
  if (t->ogsOp) {
    auto run = [&]() {
      tombo_pressure_rhs(t, L, U.get(), U.get(), BF.get(), mue.get(), irho.get(), div.get(), 1.0, true,
                         out2.get(), s);
    };
    const double withFusion = time_launches(run, s);
    nrs_ogs *h = t->ogsOp;
    DevBuf<uint8_t> f = std::move(t->fx);
    t->ogsOp = nullptr;
    const double without = time_launches(run, s);
    if (withFusion < without) {
      t->ogsOp = h;
      t->fx = std::move(f);
    } else {
      delete h;
      t->fusedPairs = 0;
    }
  }
}

}  // namespace

nrs_tombo *tombo_setup(int N, int E, const double *x, const double *y, const double *z,
                       const long long *gids, const int *bcType, int flags)
{
  NRS_CHECK(N >= 1 && N <= 9, "tombo setup: N must be 1..9");
  NRS_CHECK(E > 0, "tombo setup: no elements");
  NRS_CHECK(x && y && z && gids && bcType, "tombo setup: null mesh arrays");
  auto t = std::make_unique<nrs_tombo>();
  t->N = N; t->Nq = N + 1; t->Np = t->Nq * t->Nq * t->Nq; t->E = E;
  t->Nlocal = (int64_t)E * t->Np;
  NRS_CHECK(t->Nlocal < (int64_t)INT32_MAX, "E*Np must fit in int32 (dlong)");
  hipStream_t s = nullptr;
  std::vector<double> z1, w1;
  gll_nodes(N, z1, w1);
  t->D.upload(dmatrix(z1));
  t->w.upload(w1);

  // elements with a flux face, and their face types
  std::vector<int> bslot(E, -1), bElem, bType;
  for (int e = 0; e < E; ++e) {
    bool any = false;
    for (int f = 0; f < 6; ++f) any = any || flux_face(bcType[e * 6 + f]);
    if (!any) continue;
    bslot[e] = (int)bElem.size();
    bElem.push_back(e);
    for (int f = 0; f < 6; ++f) bType.push_back(bcType[e * 6 + f]);
  }
  t->Nb = (int)bElem.size();
  DevBuf<int> dslot;
  dslot.upload(bslot);
  if (t->Nb) {
    t->bElem.upload(bElem);
    t->bType.upload(bType);
    t->bGeo.alloc((size_t)t->Nb * 6 * t->Nq * t->Nq * 4);
  }
  {
    DevBuf<double> dx, dy, dz;
    dx.upload(x, t->Nlocal); dy.upload(y, t->Nlocal); dz.upload(z, t->Nlocal);
    t->vgeo.alloc((size_t)t->Nlocal * NV);
    by_order<VgeoL>(N, E, t->D.get(), t->w.get(), dx.get(), dy.get(), dz.get(), dslot.get(),
                    t->vgeo.get(), t->bGeo.get(), s);
    NRS_LAUNCH_CHECK();
    NRS_HIP(hipStreamSynchronize(s));
  }
  // trilinear elements: the volume kernels form their factors from the corners
  {
    std::vector<double> cc;
    if (!(flags & NRS_TOMBO_STORED_GEOMETRY) && trilinear_corners(N, E, x, y, z, cc)) {
      std::vector<double> zwh(2 * (size_t)t->Nq);
      for (int q = 0; q < t->Nq; ++q) { zwh[q] = z1[q]; zwh[t->Nq + q] = w1[q]; }
      t->corners.upload(cc);
      t->zw.upload(zwh);
      t->tri = true;
    }
  }
  // invLMM = 1 / gs(JW)   (meshComputeInvLMM.cpp)
  t->ogs = ogs_setup(t->Nlocal, gids);
  if (flags & NRS_TOMBO_FUSION) {
    std::vector<long long> ids(gids, gids + t->Nlocal);
    const std::vector<uint8_t> fxv = face_fusion(N, E, ids, t->ogs->invDegreeHost);
    if (!fxv.empty()) {
      t->ogsOp = ogs_setup(t->Nlocal, ids.data());
      t->fx.upload(fxv);
      for (uint8_t f : fxv) t->fusedPairs += f;
    }
  }
  t->invLMM.alloc(t->Nlocal);
  const unsigned g = (unsigned)((t->Nlocal + 255) / 256);
  copy_jw_kernel<<<g, 256, 0, s>>>(E, t->Np, t->vgeo.get(), t->invLMM.get());
  ogs_gather_scatter<double>(t->ogs, t->invLMM.get(), 1, t->Nlocal, NRS_ADD, s);
  reciprocal_kernel<<<g, 256, 0, s>>>(t->Nlocal, t->invLMM.get());
  NRS_LAUNCH_CHECK();
  t->W.alloc(3 * (size_t)t->Nlocal);
  t->F.alloc(3 * (size_t)t->Nlocal);
  if (flags & NRS_TOMBO_STRESS) t->S.alloc(3 * (size_t)t->Nlocal);
  NRS_HIP(hipStreamSynchronize(s));
  if (flags & NRS_TOMBO_AUTOTUNE) autotune(t.get(), s);
  return t.release();
}

void tombo_curl(const nrs_tombo *t, int64_t off, const double *U, double *W, hipStream_t s)
{
  by_order<CurlL>(t->N, t, off, off, false, tuned(TUNE_TOMBO, t->N, 0, t->E, 1) == 2, U, W, s);
  NRS_LAUNCH_CHECK();
}

void tombo_pressure_rhs(nrs_tombo *t, int64_t off, const double *Ue, const double *U,
                        const double *BF, const double *mue, const double *irho, const double *div,
                        double g0idt, bool visc, double *rhs, hipStream_t s)
{
  const int64_t L = t->Nlocal;
  double *W = t->W.get(), *F = t->F.get();
  // the fused face pairs are summed by the curl kernels: the gather-scatters after
  // them leave those nodes out (F = BF copied in needs the full handle)
  const nrs_ogs *gsh = t->ogsOp ? t->ogsOp : t->ogs;
  if (visc) {
    const bool pf0 = tuned(TUNE_TOMBO, t->N, 0, t->E, 1) == 2;
    const bool pf1 = tuned(TUNE_TOMBO, t->N, 1, t->E, 1) == 2;
    by_order<CurlL>(t->N, t, off, L, false, pf0, Ue, W, s, t->fx.get());   // W = JW curl(Ue)
    ogs_gather_scatter<double>(gsh, W, 3, L, NRS_ADD, s);
    double *S = t->S.get();                                             // stress term (tombo.cpp:57-66)
    if (S) by_order<StressL>(t->N, t, off, Ue, mue, div, S, s);
    by_order<RhsL>(t->N, t, pf1, W, div, BF, off, mue, irho, F, s, t->fx.get(), S);   // F = rhsU
  } else {
    for (int c = 0; c < 3; ++c)                                         // F = BF
      NRS_HIP(hipMemcpyAsync(F + c * L, BF + c * off, L * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  ogs_gather_scatter<double>(visc ? gsh : t->ogs, F, 3, L, NRS_ADD, s);
  by_order<WdivL>(t->N, t, F, div, g0idt, rhs, s);
  by_order<SurfL>(t->N, t, g0idt, off, U, rhs, s);
  NRS_LAUNCH_CHECK();
}

}  // namespace nrs
