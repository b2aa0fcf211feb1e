// mg.hpp -- p-multigrid preconditioner (MGSolver_t / pMGLevel analogue).
#pragma once
#include <memory>

#include "elliptic.hpp"

namespace nrs {

struct MGLevelInfo {
  int N;
  int64_t Nlocal;
  double rho;          // estimated lambda_max(S A)
  int degree;          // Chebyshev degree (down leg)
};

// Multiplicative V-cycle over p-levels; the level arithmetic type is chosen at
// setup (FP32 as the reference's pfloat, or FP64).
struct MGSolver {
  virtual ~MGSolver() = default;
  // z = V-cycle(r) on FP64 E-vectors (casts around the cycle as
  // ellipticPreconditioner.cpp:57-62 does).
  virtual void apply(const double *r, double *z, hipStream_t s) = 0;
  virtual std::vector<MGLevelInfo> info() const = 0;
  virtual int coarse_iterations() const = 0;   // iterations of the last coarse solve
  // out = S in on a smoothed level, in the level arithmetic type
  virtual void level_smoother(int level, const void *in, void *out, hipStream_t s) = 0;
  // multigridLevel::{Ax, residual, smooth, coarsen, prolongate} on one level
  // (op = nrs_mg_level_op), vectors in the level arithmetic type
  virtual void level_op(int level, int op, void *in, const void *x, void *out, bool xIsZero,
                        hipStream_t s) = 0;
  // true when apply() enqueues work only (coarse level smoothed, no host round trip)
  virtual bool device_only() const = 0;
};

std::unique_ptr<MGSolver> mg_create(nrs_elliptic *el, const nrs_elliptic_desc *d, hipStream_t s);

// Global ids of the order-Nc GLL nodes of a conforming hex mesh, derived from
// the order-Nf ids (vertices keep their ids; edge, face and interior nodes are
// numbered from the ids of the vertices that bound them).  Replaces Nek5000's
// numbering of the coarse meshes (src/mesh/meshGlobalIds.cpp via createMeshMG).
std::vector<long long> coarse_global_ids(int Nf, int E, const long long *gidF, int Nc);

// Order-Nc node coordinates interpolated from order-Nf ones (map_m_to_n,
// src/mesh/meshPhysicalNodesHex3D.cpp:48-50).
std::vector<double> interp_coords(int Nf, int E, const double *xf, int Nc);

}  // namespace nrs
