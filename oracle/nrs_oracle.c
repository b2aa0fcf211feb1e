/*
 * nrs_oracle.c -- CPU restatement of nekRS's pressure-Poisson hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP implementation.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library; the product path never calls it.
 *
 * Each function names the reference file:line it restates (nekRS v23 tree).
 * The pinning of every function (golden vectors from the compiled reference
 * kernels and gslib, closed forms, invariants) lives in tests/test_oracle_*.py.
 *
 * OpenMP: the element loops carry `omp parallel for` exactly where the
 * reference serial kernels carry them under __NEKRS__OMP__.
 */
#include "nrs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ========================================================================== */
/* 1D basis                                                                   */
/* ========================================================================== */

/* Legendre P_n(x) and P_n'(x) by the three-term recurrence. */
static void legendre(int n, double x, double *p, double *dp)
{
  double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
  if (n == 0) { *p = 1.0; *dp = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    const double d2 = d0 + (2.0 * k - 1.0) * p1;
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
  *p = p1; *dp = d1;
}

/* GLL nodes z (ascending, z0=-1, zN=1) and weights w.
 * Restates JacobiGLL (src/mesh/meshBasis1D.cpp:249-281): interior nodes are the
 * roots of P_N'(x) (the reference obtains them as Gauss-Jacobi(1,1) points);
 * weights are 2/(N(N+1)P_N(z)^2) (the reference row-sums the inverse
 * Vandermonde mass matrix, which is the same quantity). Newton iteration on
 * (1-x^2)P_N'(x) from Chebyshev-Gauss-Lobatto starts. */
void orc_gll(int N, double *z, double *w)
{
  if (N == 0) { z[0] = 0.0; w[0] = 2.0; return; }
  for (int i = 0; i <= N; ++i) {
    double x = -cos(M_PI * i / N);
    if (i > 0 && i < N) {
      for (int it = 0; it < 100; ++it) {
        double p, dp;
        legendre(N, x, &p, &dp);
        /* f = (1-x^2) P'_N ; f' = -N(N+1) P_N */
        const double f = (1.0 - x * x) * dp;
        const double df = -N * (N + 1.0) * p;
        const double dx = f / df;
        x -= dx;
        if (fabs(dx) < 1e-16) break;
      }
    }
    z[i] = x;
  }
  z[0] = -1.0; z[N] = 1.0;
  /* symmetrise */
  for (int i = 0; i <= N / 2; ++i) {
    const double a = 0.5 * (z[N - i] - z[i]);
    z[i] = -a; z[N - i] = a;
  }
  if (N % 2 == 0) z[N / 2] = 0.0;
  for (int i = 0; i <= N; ++i) {
    double p, dp;
    legendre(N, z[i], &p, &dp);
    w[i] = 2.0 / (N * (N + 1.0) * p * p);
  }
}

/* Barycentric weights of the nodes z[0..n-1]. */
static void bary_weights(int n, const double *z, double *lam)
{
  for (int j = 0; j < n; ++j) {
    double prod = 1.0;
    for (int k = 0; k < n; ++k)
      if (k != j) prod *= (z[j] - z[k]);
    lam[j] = 1.0 / prod;
  }
}

/* Differentiation matrix D[i*Nq+j] = l_j'(z_i).
 * Restates Dmatrix1D (src/mesh/meshBasis1D.cpp:99-119: D = Vr V^{-1}) with the
 * barycentric formula and the negative-sum diagonal. */
void orc_dmatrix(int N, const double *z, double *D)
{
  const int Nq = N + 1;
  double lam[32];
  bary_weights(Nq, z, lam);
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
}

/* Interpolation matrix I[i*Nin+j] = l_j(zout_i), l_j Lagrange on zin.
 * Restates InterpolationMatrix1D (src/mesh/meshBasis1D.cpp:132-151). */
void orc_interp(int Nin, const double *zin, int Nout, const double *zout, double *I)
{
  double lam[32];
  bary_weights(Nin, zin, lam);
  for (int i = 0; i < Nout; ++i) {
    int hit = -1;
    for (int j = 0; j < Nin; ++j)
      if (zout[i] == zin[j]) hit = j;
    if (hit >= 0) {
      for (int j = 0; j < Nin; ++j) I[i * Nin + j] = (j == hit) ? 1.0 : 0.0;
      continue;
    }
    double den = 0.0;
    for (int j = 0; j < Nin; ++j) den += lam[j] / (zout[i] - zin[j]);
    for (int j = 0; j < Nin; ++j) I[i * Nin + j] = (lam[j] / (zout[i] - zin[j])) / den;
  }
}

/* ========================================================================== */
/* box meshes                                                                 */
/* ========================================================================== */

/* examples/kershaw/kershaw.usr:107-127 (right, left, step) */
static double k_right(double eps, double x)
{
  return (x <= 0.5) ? (2.0 - eps) * x : 1.0 + eps * (x - 1.0);
}
static double k_left(double eps, double x) { return 1.0 - k_right(eps, 1.0 - x); }
static double k_step(double a, double b, double x)
{
  if (x <= 0.0) return a;
  if (x >= 1.0) return b;
  return a + (b - a) * x;
}

/* examples/kershaw/kershaw.usr:142-189 (kershaw) */
void orc_kershaw_map(double epsy, double epsz, double x, double y, double z,
                     double *xx, double *yy, double *zz)
{
  *xx = x;
  const int layer = (int)(x * 6.0);
  const double lambda = (x - layer / 6.0) * 6.0;
  if (layer == 0) {
    *yy = k_left(epsy, y);
    *zz = k_left(epsz, z);
  } else if (layer == 1 || layer == 4) {
    *yy = k_step(k_left(epsy, y), k_right(epsy, y), lambda);
    *zz = k_step(k_left(epsz, z), k_right(epsz, z), lambda);
  } else if (layer == 2) {
    const double t = lambda / 2.0;
    *yy = k_step(k_right(epsy, y), k_left(epsy, y), t);
    *zz = k_step(k_right(epsz, z), k_left(epsz, z), t);
  } else if (layer == 3) {
    const double t = (1.0 + lambda) / 2.0;
    *yy = k_step(k_right(epsy, y), k_left(epsy, y), t);
    *zz = k_step(k_right(epsz, z), k_left(epsz, z), t);
  } else {
    *yy = k_right(epsy, y);
    *zz = k_right(epsz, z);
  }
}

/* Structured box [0,1]^3 with nx*ny*nz hexes, lexicographic element order,
 * GLL nodes of order N inside each element ([k][j][i], i fastest).
 *  map NONE    : affine elements
 *  map SINE    : x_i += 0.05 sin(pi x) sin(pi y) sin(pi z) at every GLL node
 *  map KERSHAW : vertices mapped by kershaw(eps,eps) (kershaw.usr:191-226),
 *                nodes by trilinear interpolation of the mapped vertices
 *                (what Nek5000 does for elements without curved sides).
 * gid: 1-based global ids of the structured node grid (0 = excluded).
 * EToB[e*6+f]: DIRICHLET on boundary faces, 0 on interior faces; face order
 * follows FaceNodesHex3D (src/mesh/meshBasisHex3D.cpp:53-82):
 *   0:t=-1 1:s=-1 2:r=+1 3:s=+1 4:r=-1 5:t=+1. */
void orc_box_mesh(int nx, int ny, int nz, int N, int map, double eps,
                  double *x, double *y, double *z, long long *gid, int *EToB)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  double r[32], w[32];
  orc_gll(N, r, w);
  const long long gx = (long long)nx * N + 1, gy = (long long)ny * N + 1;
  for (int ez = 0; ez < nz; ++ez)
    for (int ey = 0; ey < ny; ++ey)
      for (int ex = 0; ex < nx; ++ex) {
        const long long e = ex + (long long)nx * (ey + (long long)ny * ez);
        double vx[8], vy[8], vz[8];
        for (int c = 0; c < 8; ++c) {
          const double X = (double)(ex + (c & 1)) / nx;
          const double Y = (double)(ey + ((c >> 1) & 1)) / ny;
          const double Z = (double)(ez + ((c >> 2) & 1)) / nz;
          if (map == ORC_MAP_KERSHAW)
            orc_kershaw_map(eps, eps, X, Y, Z, &vx[c], &vy[c], &vz[c]);
          else { vx[c] = X; vy[c] = Y; vz[c] = Z; }
        }
        for (int k = 0; k < Nq; ++k)
          for (int j = 0; j < Nq; ++j)
            for (int i = 0; i < Nq; ++i) {
              const long long n = e * Np + i + Nq * (j + Nq * k);
              const double a = 0.5 * (1.0 + r[i]), b = 0.5 * (1.0 + r[j]), c = 0.5 * (1.0 + r[k]);
              double px = 0, py = 0, pz = 0;
              for (int v = 0; v < 8; ++v) {
                const double wv = ((v & 1) ? a : 1.0 - a) * (((v >> 1) & 1) ? b : 1.0 - b) *
                                  (((v >> 2) & 1) ? c : 1.0 - c);
                px += wv * vx[v]; py += wv * vy[v]; pz += wv * vz[v];
              }
              if (map == ORC_MAP_SINE) {
                const double s = 0.05 * sin(M_PI * px) * sin(M_PI * py) * sin(M_PI * pz);
                px += s; py += s; pz += s;
              }
              x[n] = px; y[n] = py; z[n] = pz;
              const long long Gx = (long long)ex * N + i, Gy = (long long)ey * N + j,
                              Gz = (long long)ez * N + k;
              gid[n] = 1 + Gx + gx * (Gy + gy * Gz);
            }
        if (EToB) {
          EToB[e * 6 + 0] = (ez == 0) ? ORC_DIRICHLET : 0;
          EToB[e * 6 + 1] = (ey == 0) ? ORC_DIRICHLET : 0;
          EToB[e * 6 + 2] = (ex == nx - 1) ? ORC_DIRICHLET : 0;
          EToB[e * 6 + 3] = (ey == ny - 1) ? ORC_DIRICHLET : 0;
          EToB[e * 6 + 4] = (ex == 0) ? ORC_DIRICHLET : 0;
          EToB[e * 6 + 5] = (ez == nz - 1) ? ORC_DIRICHLET : 0;
        }
      }
}

/* ========================================================================== */
/* geometric factors: kernels/mesh/geometricFactorsHex3D.okl:27-143           */
/* ========================================================================== */
void orc_geofacs(int N, int E, const double *D, const double *w,
                 const double *x, const double *y, const double *z,
                 double *ggeo, double *Jout)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
#pragma omp parallel for
  for (int e = 0; e < E; ++e) {
    const double *X = x + (size_t)e * Np, *Y = y + (size_t)e * Np, *Z = z + (size_t)e * Np;
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          double xr = 0, yr = 0, zr = 0, xs = 0, ys = 0, zs = 0, xt = 0, yt = 0, zt = 0;
          for (int m = 0; m < Nq; ++m) {
            const double Dim = D[i * Nq + m], Djm = D[j * Nq + m], Dkm = D[k * Nq + m];
            const int nr = m + Nq * (j + Nq * k), ns = i + Nq * (m + Nq * k), nt = i + Nq * (j + Nq * m);
            xr += Dim * X[nr]; xs += Djm * X[ns]; xt += Dkm * X[nt];
            yr += Dim * Y[nr]; ys += Djm * Y[ns]; yt += Dkm * Y[nt];
            zr += Dim * Z[nr]; zs += Djm * Z[ns]; zt += Dkm * Z[nt];
          }
          const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
          const double Jinv = 1.0 / J;
          const double JW = J * w[i] * w[j] * w[k];
          const double rx = (ys * zt - zs * yt) * Jinv, ry = -(xs * zt - zs * xt) * Jinv,
                       rz = (xs * yt - ys * xt) * Jinv;
          const double sx = -(yr * zt - zr * yt) * Jinv, sy = (xr * zt - zr * xt) * Jinv,
                       sz = -(xr * yt - yr * xt) * Jinv;
          const double tx = (yr * zs - zr * ys) * Jinv, ty = -(xr * zs - zr * xs) * Jinv,
                       tz = (xr * ys - yr * xs) * Jinv;
          const int n = i + Nq * (j + Nq * k);
          double *G = ggeo + (size_t)e * 7 * Np;
          G[ORC_G00 * Np + n] = JW * (rx * rx + ry * ry + rz * rz);
          G[ORC_G01 * Np + n] = JW * (rx * sx + ry * sy + rz * sz);
          G[ORC_G02 * Np + n] = JW * (rx * tx + ry * ty + rz * tz);
          G[ORC_G11 * Np + n] = JW * (sx * sx + sy * sy + sz * sz);
          G[ORC_G12 * Np + n] = JW * (sx * tx + sy * ty + sz * tz);
          G[ORC_G22 * Np + n] = JW * (tx * tx + ty * ty + tz * tz);
          G[ORC_GWJ * Np + n] = JW;
          if (Jout) Jout[(size_t)e * Np + n] = J;
        }
  }
}

/* ========================================================================== */
/* gather-scatter: 3rd_party/gslib/ogs/src/ogsSetup.cpp:111-397 (local part)  */
/* ========================================================================== */
typedef struct { long long base; int64_t local; } node_t;

static int cmp_base_local(const void *a, const void *b)
{
  const node_t *x = (const node_t *)a, *y = (const node_t *)b;
  if (x->base != y->base) return (x->base < y->base) ? -1 : 1;
  return (x->local < y->local) ? -1 : (x->local > y->local);
}

/* Rows are ordered by the local index of their first member, members ascend
 * (ogsSetup.cpp:190-246: sort by base id then local id, number rows, re-sort
 * by local id, renumber rows in first-appearance order). invDegree is the
 * reciprocal multiplicity, 1 at excluded nodes (ogsSetup.cpp:364-392). */
orc_ogs *orc_ogs_setup(int64_t N, const long long *gids)
{
  orc_ogs *o = (orc_ogs *)calloc(1, sizeof(orc_ogs));
  o->N = N;
  int64_t n = 0;
  for (int64_t i = 0; i < N; ++i) n += (gids[i] != 0);
  node_t *nodes = (node_t *)malloc(sizeof(node_t) * (n ? n : 1));
  n = 0;
  for (int64_t i = 0; i < N; ++i)
    if (gids[i] != 0) { nodes[n].base = gids[i]; nodes[n].local = i; ++n; }
  qsort(nodes, n, sizeof(node_t), cmp_base_local);
  /* group id per sorted node, and first local index per group */
  int64_t ngroups = 0;
  int64_t *grp = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
  for (int64_t s = 0; s < n; ++s) {
    if (s == 0 || nodes[s].base != nodes[s - 1].base) ++ngroups;
    grp[s] = ngroups - 1;
  }
  int64_t *first = (int64_t *)malloc(sizeof(int64_t) * (ngroups ? ngroups : 1));
  int64_t *count = (int64_t *)calloc(ngroups ? ngroups : 1, sizeof(int64_t));
  for (int64_t s = 0; s < n; ++s) {
    if (s == 0 || grp[s] != grp[s - 1]) first[grp[s]] = nodes[s].local; /* sorted by local within group */
    count[grp[s]]++;
  }
  /* order groups by first appearance: a group's rank = order of its first local id */
  int64_t *rowOf = (int64_t *)malloc(sizeof(int64_t) * (ngroups ? ngroups : 1));
  {
    int64_t *owner = (int64_t *)malloc(sizeof(int64_t) * (N ? N : 1));
    for (int64_t i = 0; i < N; ++i) owner[i] = -1;
    for (int64_t g = 0; g < ngroups; ++g) owner[first[g]] = g;
    int64_t r = 0;
    for (int64_t i = 0; i < N; ++i)
      if (owner[i] >= 0) rowOf[owner[i]] = r++;
    free(owner);
  }
  o->Nrows = ngroups;
  o->Nids = n;
  o->offsets = (int64_t *)calloc(ngroups + 1, sizeof(int64_t));
  o->ids = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
  for (int64_t g = 0; g < ngroups; ++g) o->offsets[rowOf[g] + 1] = count[g];
  for (int64_t r = 0; r < ngroups; ++r) o->offsets[r + 1] += o->offsets[r];
  {
    int64_t *fill = (int64_t *)calloc(ngroups ? ngroups : 1, sizeof(int64_t));
    for (int64_t s = 0; s < n; ++s) {
      const int64_t r = rowOf[grp[s]];
      o->ids[o->offsets[r] + fill[r]++] = nodes[s].local;
    }
    free(fill);
  }
  o->invDegree = (double *)malloc(sizeof(double) * (N ? N : 1));
  for (int64_t i = 0; i < N; ++i) o->invDegree[i] = 1.0;
  for (int64_t r = 0; r < ngroups; ++r) {
    const int64_t c = o->offsets[r + 1] - o->offsets[r];
    for (int64_t m = o->offsets[r]; m < o->offsets[r + 1]; ++m) o->invDegree[o->ids[m]] = 1.0 / (double)c;
  }
  free(nodes); free(grp); free(first); free(count); free(rowOf);
  return o;
}

void orc_ogs_free(orc_ogs *o)
{
  if (!o) return;
  free(o->offsets); free(o->ids); free(o->invDegree); free(o);
}

int64_t orc_ogs_nrows(const orc_ogs *o) { return o->Nrows; }

void orc_ogs_get(const orc_ogs *o, int64_t *offsets, int64_t *ids, double *invDegree)
{
  if (offsets) memcpy(offsets, o->offsets, sizeof(int64_t) * (o->Nrows + 1));
  if (ids) memcpy(ids, o->ids, sizeof(int64_t) * o->Nids);
  if (invDegree) memcpy(invDegree, o->invDegree, sizeof(double) * o->N);
}

/* 3rd_party/gslib/ogs/okl/gatherScatterMany.okl:58-86: per row, reduce the
 * members in ascending order starting from the op identity (0 for add), then
 * write the result to every member; singleton rows are left untouched. */
#define GS_BODY(T, IDENT_ADD, IDENT_MUL, IDENT_MIN, IDENT_MAX)                          \
  for (int k = 0; k < nvec; ++k) {                                                      \
    T *q = v + (size_t)k * stride;                                                      \
    _Pragma("omp parallel for")                                                         \
    for (int64_t r = 0; r < o->Nrows; ++r) {                                            \
      const int64_t s = o->offsets[r], e = o->offsets[r + 1];                           \
      if (s + 1 == e) continue;                                                         \
      T g;                                                                              \
      switch (op) {                                                                     \
        case ORC_OP_ADD: g = IDENT_ADD; for (int64_t m = s; m < e; ++m) g += q[o->ids[m]]; break; \
        case ORC_OP_MUL: g = IDENT_MUL; for (int64_t m = s; m < e; ++m) g *= q[o->ids[m]]; break; \
        case ORC_OP_MIN: g = IDENT_MIN; for (int64_t m = s; m < e; ++m) { T a = q[o->ids[m]]; g = (a < g) ? a : g; } break; \
        default:         g = IDENT_MAX; for (int64_t m = s; m < e; ++m) { T a = q[o->ids[m]]; g = (a > g) ? a : g; } break; \
      }                                                                                 \
      for (int64_t m = s; m < e; ++m) q[o->ids[m]] = g;                                 \
    }                                                                                   \
  }

void orc_gs_f64(const orc_ogs *o, double *v, int nvec, int64_t stride, int op)
{
  GS_BODY(double, 0.0, 1.0, INFINITY, -INFINITY)
}
void orc_gs_f32(const orc_ogs *o, float *v, int nvec, int64_t stride, int op)
{
  GS_BODY(float, 0.0f, 1.0f, INFINITY, -INFINITY)
}
void orc_gs_i32(const orc_ogs *o, int *v, int nvec, int64_t stride, int op)
{
  GS_BODY(int, 0, 1, 2147483647, (-2147483647 - 1))
}

/* ========================================================================== */
/* Dirichlet mask: src/solvers/elliptic/ellipticOgs.cpp:17-60                  */
/* ========================================================================== */
int64_t orc_mask_ids(int N, int E, const orc_ogs *ogs, const int *EToB, int64_t *maskIds)
{
  const int Nq = N + 1, Nfp = Nq * Nq, Np = Nq * Nq * Nq;
  const int64_t L = (int64_t)E * Np;
  const int large = 1 << 20;
  int *mapB = (int *)malloc(sizeof(int) * (L ? L : 1));
  /* faceNodes per FaceNodesHex3D ordering */
  int *faceNodes = (int *)malloc(sizeof(int) * 6 * Nfp);
  {
    int cnt[6] = {0, 0, 0, 0, 0, 0};
    for (int n = 0; n < Np; ++n) {
      const int i = n % Nq, j = (n / Nq) % Nq, k = n / Nfp;
      if (k == 0) faceNodes[0 * Nfp + cnt[0]++] = n;
      if (j == 0) faceNodes[1 * Nfp + cnt[1]++] = n;
      if (i == N) faceNodes[2 * Nfp + cnt[2]++] = n;
      if (j == N) faceNodes[3 * Nfp + cnt[3]++] = n;
      if (i == 0) faceNodes[4 * Nfp + cnt[4]++] = n;
      if (k == N) faceNodes[5 * Nfp + cnt[5]++] = n;
    }
  }
  for (int64_t e = 0; e < E; ++e) {
    for (int n = 0; n < Np; ++n) mapB[e * Np + n] = large;
    for (int f = 0; f < 6; ++f) {
      const int bc = EToB[e * 6 + f];
      if (bc > 0)
        for (int n = 0; n < Nfp; ++n) {
          const int64_t id = e * Np + faceNodes[f * Nfp + n];
          if (bc < mapB[id]) mapB[id] = bc;
        }
    }
  }
  orc_gs_i32(ogs, mapB, 1, L, ORC_OP_MIN);
  int64_t Nm = 0;
  for (int64_t n = 0; n < L; ++n)
    if (mapB[n] == ORC_DIRICHLET) {
      if (maskIds) maskIds[Nm] = n;
      ++Nm;
    }
  free(mapB); free(faceNodes);
  return Nm;
}

/* kernels/core/mask.okl:27-35 */
void orc_mask_f64(int64_t Nmasked, const int64_t *maskIds, double *v)
{
  for (int64_t n = 0; n < Nmasked; ++n) v[maskIds[n]] = 0.0;
}
void orc_mask_f32(int64_t Nmasked, const int64_t *maskIds, float *v)
{
  for (int64_t n = 0; n < Nmasked; ++n) v[maskIds[n]] = 0.0f;
}

/* ========================================================================== */
/* Ax: kernels/elliptic/ellipticPartialAxCoeffHex3D.c:2-97                     */
/* qr = sum_m D[i][m] q(m,j,k) (the reference reads S=D^T as S[m*Nq+i]),        */
/* G-contraction scaled by lambda0, then Aq = sum_m D[m][i] Gqr(m,j,k) + ...,   */
/* plus GwJ*lambda1*q when not Poisson.  lambda_field=0 reads lambda[0]          */
/* (p_lambda = 0), 1 reads lambda[id] (coefficient field).                      */
/* ========================================================================== */
#define AX_BODY(T)                                                                        \
  const int Nq = N + 1, Np = Nq * Nq * Nq;                                                \
  _Pragma("omp parallel for")                                                             \
  for (int e = 0; e < E; ++e) {                                                           \
    T s_q[16 * 16 * 16], s_Gqr[16 * 16 * 16], s_Gqs[16 * 16 * 16], s_Gqt[16 * 16 * 16];   \
    const size_t el = elist ? (size_t)elist[e] : (size_t)e;                               \
    for (int n = 0; n < Np; ++n) s_q[n] = q[el * Np + n];                                 \
    for (int k = 0; k < Nq; ++k)                                                          \
      for (int j = 0; j < Nq; ++j)                                                        \
        for (int i = 0; i < Nq; ++i) {                                                    \
          const size_t gb = el * 7 * Np + i + Nq * (j + Nq * k);                          \
          const T G00 = ggeo[gb + ORC_G00 * Np], G01 = ggeo[gb + ORC_G01 * Np];           \
          const T G11 = ggeo[gb + ORC_G11 * Np], G12 = ggeo[gb + ORC_G12 * Np];           \
          const T G02 = ggeo[gb + ORC_G02 * Np], G22 = ggeo[gb + ORC_G22 * Np];           \
          const size_t id = el * Np + i + Nq * (j + Nq * k);                              \
          const T lam0 = lambda0[lambda_field ? id : 0];                                  \
          T qr = 0, qs = 0, qt = 0;                                                       \
          for (int m = 0; m < Nq; ++m) {                                                  \
            qr += D[i * Nq + m] * s_q[m + Nq * (j + Nq * k)];                             \
            qs += D[j * Nq + m] * s_q[i + Nq * (m + Nq * k)];                             \
            qt += D[k * Nq + m] * s_q[i + Nq * (j + Nq * m)];                             \
          }                                                                               \
          T Gqr = G00 * qr; Gqr += G01 * qs; Gqr += G02 * qt;                             \
          T Gqs = G01 * qr; Gqs += G11 * qs; Gqs += G12 * qt;                             \
          T Gqt = G02 * qr; Gqt += G12 * qs; Gqt += G22 * qt;                             \
          const int n = i + Nq * (j + Nq * k);                                            \
          s_Gqr[n] = lam0 * Gqr; s_Gqs[n] = lam0 * Gqs; s_Gqt[n] = lam0 * Gqt;            \
        }                                                                                 \
    for (int k = 0; k < Nq; ++k)                                                          \
      for (int j = 0; j < Nq; ++j)                                                        \
        for (int i = 0; i < Nq; ++i) {                                                    \
          const size_t gb = el * 7 * Np + i + Nq * (j + Nq * k);                          \
          const size_t id = el * Np + i + Nq * (j + Nq * k);                              \
          T r_Aq = 0;                                                                     \
          if (!poisson) {                                                                 \
            const T lam1 = lambda1[lambda_field ? id : 0];                                \
            r_Aq = ggeo[gb + ORC_GWJ * Np] * lam1 * s_q[i + Nq * (j + Nq * k)];           \
          }                                                                               \
          T Ar = 0, As = 0, At = 0;                                                       \
          for (int m = 0; m < Nq; ++m) {                                                  \
            Ar += D[m * Nq + i] * s_Gqr[m + Nq * (j + Nq * k)];                           \
            As += D[m * Nq + j] * s_Gqs[i + Nq * (m + Nq * k)];                           \
            At += D[m * Nq + k] * s_Gqt[i + Nq * (j + Nq * m)];                           \
          }                                                                               \
          Aq[id] = Ar + As + At + r_Aq;                                                   \
        }                                                                                 \
  }

void orc_ax_f64(int N, int E, const int *elist, const double *ggeo, const double *D,
                const double *lambda0, const double *lambda1, int poisson, int lambda_field,
                const double *q, double *Aq)
{
  AX_BODY(double)
}

void orc_ax_f32(int N, int E, const int *elist, const float *ggeo, const float *D,
                const float *lambda0, const float *lambda1, int poisson, int lambda_field,
                const float *q, float *Aq)
{
  AX_BODY(float)
}

/* kernels/elliptic/ellipticBlockPartialAxCoeffHex3D.c:2-147: nfld fields with
 * fieldOffset `offset`, per-field coefficient fields with stride `loffset`. */
void orc_block_ax_f64(int N, int E, int nfld, int64_t offset, int64_t loffset,
                      const double *ggeo, const double *D, const double *lambda0,
                      const double *lambda1, const double *q, double *Aq)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  for (int f = 0; f < nfld; ++f) {
#pragma omp parallel for
    for (int e = 0; e < E; ++e) {
      double s_q[16 * 16 * 16], s_Gqr[16 * 16 * 16], s_Gqs[16 * 16 * 16], s_Gqt[16 * 16 * 16];
      const double *qf = q + f * offset;
      const double *l0 = lambda0 + f * loffset, *l1 = lambda1 + f * loffset;
      for (int n = 0; n < Np; ++n) s_q[n] = qf[(size_t)e * Np + n];
      for (int n = 0; n < Np; ++n) {
        const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
        const size_t gb = (size_t)e * 7 * Np + n, id = (size_t)e * Np + n;
        double qr = 0, qs = 0, qt = 0;
        for (int m = 0; m < Nq; ++m) {
          qr += D[i * Nq + m] * s_q[m + Nq * (j + Nq * k)];
          qs += D[j * Nq + m] * s_q[i + Nq * (m + Nq * k)];
          qt += D[k * Nq + m] * s_q[i + Nq * (j + Nq * m)];
        }
        const double G00 = ggeo[gb + ORC_G00 * Np], G01 = ggeo[gb + ORC_G01 * Np],
                     G11 = ggeo[gb + ORC_G11 * Np], G12 = ggeo[gb + ORC_G12 * Np],
                     G02 = ggeo[gb + ORC_G02 * Np], G22 = ggeo[gb + ORC_G22 * Np];
        s_Gqr[n] = l0[id] * (G00 * qr + G01 * qs + G02 * qt);
        s_Gqs[n] = l0[id] * (G01 * qr + G11 * qs + G12 * qt);
        s_Gqt[n] = l0[id] * (G02 * qr + G12 * qs + G22 * qt);
      }
      for (int n = 0; n < Np; ++n) {
        const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
        const size_t gb = (size_t)e * 7 * Np + n, id = (size_t)e * Np + n;
        const double r0 = ggeo[gb + ORC_GWJ * Np] * l1[id] * s_q[n];
        double Ar = 0, As = 0, At = 0;
        for (int m = 0; m < Nq; ++m) Ar += D[m * Nq + i] * s_Gqr[m + Nq * (j + Nq * k)];
        for (int m = 0; m < Nq; ++m) As += D[m * Nq + j] * s_Gqs[i + Nq * (m + Nq * k)];
        for (int m = 0; m < Nq; ++m) At += D[m * Nq + k] * s_Gqt[i + Nq * (j + Nq * m)];
        Aq[f * offset + id] = Ar + As + At + r0;
      }
    }
  }
}

/* Element diagonal of A: kernels/elliptic/ellipticBlockBuildDiagonalHex3D.okl:27-117
 * (single field). */
void orc_build_diag_f64(int N, int E, const double *ggeo, const double *D,
                        const double *lambda0, const double *lambda1, int poisson,
                        int lambda_field, double *diag)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
#pragma omp parallel for
  for (int e = 0; e < E; ++e) {
    const double *G = ggeo + (size_t)e * 7 * Np;
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int n = i + Nq * (j + Nq * k);
          const size_t id = (size_t)e * Np + n;
          const double l0 = lambda0[lambda_field ? id : 0];
          double r = 0.0;
          r += 2.0 * G[ORC_G01 * Np + n] * l0 * D[i * Nq + i] * D[j * Nq + j];
          r += 2.0 * G[ORC_G02 * Np + n] * l0 * D[i * Nq + i] * D[k * Nq + k];
          r += 2.0 * G[ORC_G12 * Np + n] * l0 * D[j * Nq + j] * D[k * Nq + k];
          for (int m = 0; m < Nq; ++m) {
            const int nr = m + Nq * (j + Nq * k), ns = i + Nq * (m + Nq * k), nt = i + Nq * (j + Nq * m);
            const double lr = lambda0[lambda_field ? (size_t)e * Np + nr : 0];
            const double ls = lambda0[lambda_field ? (size_t)e * Np + ns : 0];
            const double lt = lambda0[lambda_field ? (size_t)e * Np + nt : 0];
            r += G[ORC_G00 * Np + nr] * lr * D[m * Nq + i] * D[m * Nq + i];
            r += G[ORC_G11 * Np + ns] * ls * D[m * Nq + j] * D[m * Nq + j];
            r += G[ORC_G22 * Np + nt] * lt * D[m * Nq + k] * D[m * Nq + k];
          }
          if (!poisson) r += G[ORC_GWJ * Np + n] * lambda1[lambda_field ? id : 0];
          diag[id] = r;
        }
  }
}

/* Transfers: kernels/elliptic/ellipticPreconCoarsenHex3D.c:26-100 and
 * ellipticPreconProlongateHex3D.c:27-114.  R is [NqC][NqF] (the transpose of
 * the coarse-to-fine interpolation matrix, MG/ellipticMultiGridLevelSetup.cpp:238-258).
 * coarsen:    qc = (R x R x R) qf
 * prolongate: qf += (R^T x R^T x R^T) qc */
void orc_coarsen(int Nf, int Nc, int E, const double *R, const double *qf, double *qc)
{
  const int NqF = Nf + 1, NqC = Nc + 1, NpF = NqF * NqF * NqF, NpC = NqC * NqC * NqC;
#pragma omp parallel for
  for (int e = 0; e < E; ++e) {
    double a[16 * 16 * 16], b[16 * 16 * 16];
    const double *u = qf + (size_t)e * NpF;
    /* contract k: a[kc][j][i] */
    for (int kc = 0; kc < NqC; ++kc)
      for (int j = 0; j < NqF; ++j)
        for (int i = 0; i < NqF; ++i) {
          double s = 0;
          for (int k = 0; k < NqF; ++k) s += R[kc * NqF + k] * u[i + NqF * (j + NqF * k)];
          a[i + NqF * (j + NqF * kc)] = s;
        }
    /* contract j: b[kc][jc][i] */
    for (int kc = 0; kc < NqC; ++kc)
      for (int jc = 0; jc < NqC; ++jc)
        for (int i = 0; i < NqF; ++i) {
          double s = 0;
          for (int j = 0; j < NqF; ++j) s += R[jc * NqF + j] * a[i + NqF * (j + NqF * kc)];
          b[i + NqF * (jc + NqC * kc)] = s;
        }
    for (int kc = 0; kc < NqC; ++kc)
      for (int jc = 0; jc < NqC; ++jc)
        for (int ic = 0; ic < NqC; ++ic) {
          double s = 0;
          for (int i = 0; i < NqF; ++i) s += R[ic * NqF + i] * b[i + NqF * (jc + NqC * kc)];
          qc[(size_t)e * NpC + ic + NqC * (jc + NqC * kc)] = s;
        }
  }
}

void orc_prolongate(int Nf, int Nc, int E, const double *R, const double *qc, double *qf)
{
  const int NqF = Nf + 1, NqC = Nc + 1, NpF = NqF * NqF * NqF, NpC = NqC * NqC * NqC;
#pragma omp parallel for
  for (int e = 0; e < E; ++e) {
    double a[16 * 16 * 16], b[16 * 16 * 16];
    const double *u = qc + (size_t)e * NpC;
    for (int k = 0; k < NqF; ++k)
      for (int jc = 0; jc < NqC; ++jc)
        for (int ic = 0; ic < NqC; ++ic) {
          double s = 0;
          for (int kc = 0; kc < NqC; ++kc) s += R[kc * NqF + k] * u[ic + NqC * (jc + NqC * kc)];
          a[ic + NqC * (jc + NqC * k)] = s;
        }
    for (int k = 0; k < NqF; ++k)
      for (int j = 0; j < NqF; ++j)
        for (int ic = 0; ic < NqC; ++ic) {
          double s = 0;
          for (int jc = 0; jc < NqC; ++jc) s += R[jc * NqF + j] * a[ic + NqC * (jc + NqC * k)];
          b[ic + NqC * (j + NqF * k)] = s;
        }
    for (int k = 0; k < NqF; ++k)
      for (int j = 0; j < NqF; ++j)
        for (int i = 0; i < NqF; ++i) {
          double s = 0;
          for (int ic = 0; ic < NqC; ++ic) s += R[ic * NqF + i] * b[ic + NqC * (j + NqF * k)];
          qf[(size_t)e * NpF + i + NqF * (j + NqF * k)] += s;
        }
  }
}

/* ========================================================================== */
/* BLAS-1: kernels/linAlg/weightedInnerProdMany.c:22-49                         */
/* ========================================================================== */
double orc_wdot(int64_t n, const double *w, const double *a, const double *b)
{
  double s = 0.0;
#pragma omp parallel for reduction(+ : s)
  for (int64_t i = 0; i < n; ++i) s += w[i] * a[i] * b[i];
  return s;
}

/* kernels/elliptic/ellipticBlockUpdatePCG.c:27-50 (one field):
 * r -= alpha Ap, returns sum w r^2 over the updated r. */
double orc_update_pcg(int64_t n, const double *w, const double *Ap, double alpha, double *r)
{
  double rr = 0.0;
#pragma omp parallel for reduction(+ : rr)
  for (int64_t i = 0; i < n; ++i) {
    const double rn = r[i] - alpha * Ap[i];
    rr += rn * rn * w[i];
    r[i] = rn;
  }
  return rr;
}

/* ========================================================================== */
/* operator and solvers                                                        */
/* ========================================================================== */

/* src/solvers/elliptic/ellipticOperator.cpp:117-172 (non-overlapped branch):
 * Ax over all elements, mask, gather-scatter add. */
void orc_operator(const orc_elliptic *el, const double *q, double *Aq)
{
  orc_ax_f64(el->N, el->E, NULL, el->ggeo, el->D, &el->lambda0, &el->lambda1, el->poisson, 0, q, Aq);
  orc_mask_f64(el->Nmasked, el->maskIds, Aq);
  orc_gs_f64(el->ogs, Aq, 1, el->ogs->N, ORC_OP_ADD);
}

void orc_zero_mean(int64_t n, double *q)
{
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += q[i];
  s /= (double)n;
  for (int64_t i = 0; i < n; ++i) q[i] -= s;
}

/* src/solvers/elliptic/linearSolver/PCG.cpp:33-83 (update) and 85-203 (pcg).
 * precon 0 = NONE (z = r), 1 = JACOBI (z = invDiag .* r).
 * reshist (optional, maxit+1 entries) receives rdotr after each iteration. */
int orc_pcg(const orc_elliptic *el, int precon, const double *invDiag, int flexible,
            double *r, double *x, double tol, int maxit, double *rdotr, double *reshist)
{
  const int64_t L = el->ogs->N;
  const double *w = el->ogs->invDegree;
  double *p = (double *)calloc(L, sizeof(double));
  double *z = (double *)calloc(L, sizeof(double));
  double *Ap = (double *)calloc(L, sizeof(double));
  const double resNormFactor = 1.0 / el->volume;
  double rdotz1 = 0.0, alpha = 0.0;
  int iter = 0;
  do {
    iter++;
    const double rdotz2 = rdotz1;
    if (precon == 1) {
      for (int64_t n = 0; n < L; ++n) z[n] = invDiag[n] * r[n];
      if (el->allNeumann) orc_zero_mean(L, z);      /* ellipticPreconditioner.cpp:82-83 */
      rdotz1 = orc_wdot(L, w, r, z);
    } else {
      /* PCG.cpp:101,128-130: with no preconditioner z aliases r and rdotz1 is
       * set to the caller's running residual norm `rdotr` (the normalised
       * sqrt(sum w r^2 / volume)), not to <r,r>_w.  Restated as is. */
      memcpy(z, r, sizeof(double) * L);
      rdotz1 = *rdotr;
    }
    double beta = 0.0;
    if (iter > 1) {
      beta = rdotz1 / rdotz2;
      if (flexible) {
        const double zdotAp = orc_wdot(L, w, z, Ap);
        beta = -alpha * zdotAp / rdotz2;
      }
    }
    for (int64_t n = 0; n < L; ++n) p[n] = z[n] + beta * p[n];
    orc_operator(el, p, Ap);
    const double pAp = orc_wdot(L, w, p, Ap);
    alpha = rdotz1 / (pAp + 1e-300);
    const double rr = orc_update_pcg(L, w, Ap, alpha, r);
    for (int64_t n = 0; n < L; ++n) x[n] += alpha * p[n];
    *rdotr = sqrt(rr * resNormFactor);
    if (reshist) reshist[iter - 1] = *rdotr;
    if (isnan(*rdotr)) { iter = -iter; break; }
  } while (*rdotr > tol && iter < maxit);
  free(p); free(z); free(Ap);
  return iter;
}

/* src/solvers/elliptic/ellipticSolve.cpp:32-190 (no projection):
 * r = rhs - A x0 (element Ax only), mask, gs; x0 saved, x = 0; res0 = ||r||_w
 * scaled by 1/volume; tol *= res0 if relative; pcg; x += x0. */
int orc_solve(const orc_elliptic *el, int precon, const double *invDiag, int flexible,
              double tol, int relative, int maxit, const double *rhs, double *x,
              double *res0, double *res, double *reshist)
{
  const int64_t L = el->ogs->N;
  double *r = (double *)malloc(sizeof(double) * L);
  double *Ax = (double *)malloc(sizeof(double) * L);
  double *x0 = (double *)malloc(sizeof(double) * L);
  orc_ax_f64(el->N, el->E, NULL, el->ggeo, el->D, &el->lambda0, &el->lambda1, el->poisson, 0, x, Ax);
  for (int64_t n = 0; n < L; ++n) r[n] = rhs[n] - Ax[n];
  if (el->allNeumann) orc_zero_mean(L, r);          /* ellipticSolve.cpp:101 */
  orc_mask_f64(el->Nmasked, el->maskIds, r);
  orc_gs_f64(el->ogs, r, 1, L, ORC_OP_ADD);
  memcpy(x0, x, sizeof(double) * L);
  memset(x, 0, sizeof(double) * L);
  const double r0 = sqrt(orc_wdot(L, el->ogs->invDegree, r, r) / el->volume);
  if (relative) tol *= r0;
  double rn = r0;
  const int it = orc_pcg(el, precon, invDiag, flexible, r, x, tol, maxit, &rn, reshist);
  for (int64_t n = 0; n < L; ++n) x[n] += x0[n];
  if (el->allNeumann) orc_zero_mean(L, x);          /* ellipticSolve.cpp:188-189 */
  if (res0) *res0 = r0;
  if (res) *res = rn;
  free(r); free(Ax); free(x0);
  return it;
}
