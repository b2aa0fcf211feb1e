// amg.hpp -- smoothed-aggregation AMG for the assembled order-1 coarse problem.
//
// The reference solves the order-1 multigrid level with one BoomerAMG V-cycle
// on an assembled FEM matrix (src/solvers/elliptic/MG/coarseLevel.cpp:47-222,
// MG/ellipticBuildFEM.cpp).  Here the matrix is assembled from the level's own
// matrix-free operator (so it equals the operator the V-cycle uses), a
// smoothed-aggregation hierarchy is built on the host (strength, three-pass
// aggregation, Jacobi-smoothed tentative prolongation, Galerkin products,
// dense inverse on the last level) and every level is applied on the device
// with CSR kernels in FP64.
#pragma once
#include <vector>

#include "common.hpp"

namespace nrs {

struct CsrHost {
  int n = 0, m = 0;              // rows, columns
  std::vector<int> rp, ci;       // rp[n+1]
  std::vector<double> v;
};

struct AmgParams {
  double theta = 0.08;           // strength threshold
  int coarsest = 256;            // dense solve at or below this size
  int maxLevels = 12;
  int sweeps = 2;                // damped Jacobi sweeps before and after the coarse correction
  int powerSteps = 20;           // for rho(D^-1 A)
  uint64_t seed = 7;
  // singular Poisson matrix with the constant vector as null space (all-Neumann):
  // the dense level stores the pseudo-inverse P (A + a 1 1^T)^-1 P, P = I - 1 1^T / n
  bool nullSpace = false;
};

// Host side of the hierarchy (kept for tests and the device upload).
struct AmgHostLevel {
  CsrHost A, P, R;               // P: n x nc, R = P^T
  std::vector<double> dinv;
  double omega = 0.0;            // 4 / (3 rho(D^-1 A))
  std::vector<double> inv;       // dense (pseudo-)inverse (last level only), row-major
};
std::vector<AmgHostLevel> amg_build(const CsrHost &A, const AmgParams &p);

struct AmgDevLevel {
  int n = 0, nc = 0;
  int laneA = 8, laneP = 8, laneR = 8;   // lanes per row of each matrix's SpMV
  bool dense = false;
  double omega = 0.0;
  DevBuf<int> arp, aci, prp, pci, rrp, rci;
  DevBuf<double> av, pv, rv, dinv, inv;
  DevBuf<double> b, xa, xb, r;   // level vectors (coarser levels)
};

struct Amg {
  AmgParams params;
  std::vector<AmgDevLevel> lv;
  int sweeps = 2;
  void setup(const CsrHost &A, const AmgParams &p);
  // z = one V-cycle applied to b (finest-level vectors of size lv[0].n);
  // tmp: a finest-size scratch vector.  Enqueued on s only.
  void apply(const double *b, double *z, double *tmp, hipStream_t s);
  int levels() const { return (int)lv.size(); }
};

// y = A x, and the fused forms used by the V-cycle and the coarse PCG
enum { SPMV_Y = 0, SPMV_RES = 1, SPMV_ADD = 2, SPMV_JAC = 3 };
void csr_spmv(int mode, int n, const int *rp, const int *ci, const double *v, const double *x,
              double *y, const double *b, const double *dinv, double omega, hipStream_t s, int lanes = 8);
// lanes per row suited to a matrix (8, or 64 for rows of 48+ entries on average)
int csr_lanes(const CsrHost &A);

// E-vector <-> unique-row vectors of the order-1 level: u[row] = e[first[row]];
// e[node] = row >= 0 ? u[row] : 0
template <typename T>
void rows_gather(int n, const int *first, const T *e, double *u, hipStream_t s);
template <typename T>
void rows_scatter(int64_t L, const int *rowOf, const double *u, T *e, hipStream_t s);

}  // namespace nrs
