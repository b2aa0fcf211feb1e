// fdm.hpp -- overlapping Schwarz smoother (ASM / RAS) with fast diagonalisation
// on the extended (N+3)^3 element boxes (pMGLevel::build / smoothSchwarz analogue).
#pragma once
#include <vector>

#include "common.hpp"
#include "ogs.hpp"

namespace nrs {

// Face-to-face connection of a conforming hex mesh.  Faces are numbered as
// EToB: 0 k=0, 1 j=0, 2 i=N, 3 j=N, 4 i=0, 5 k=N; a face point has tangential
// coordinates (u, v) = (i, j) on faces 0/5, (i, k) on 1/3, (j, k) on 4/2.
// code = nbrFace | swap << 3 | flipU << 4 | flipV << 5: the neighbour's point
// matching (u, v) is (u', v') with (p, q) = swap ? (v, u) : (u, v),
// u' = flipU ? N - p : p, v' = flipV ? N - q : q, for any order N.
struct FaceLink {
  int nbr;    // neighbouring element, -1 on the boundary
  int code;
};
std::vector<FaceLink> face_links(int N, int E, const long long *gid);

// Element lengths [9][E]: left x,y,z, middle x,y,z, right x,y,z
// (ellipticMultiGridSchwarz.cpp:58-276) from order-N coordinates; the
// neighbour lengths come through the face links.
std::vector<double> element_lengths(int N, int E, const double *x, const double *y,
                                    const double *z, const std::vector<FaceLink> &links);

// 1D FDM pair on nl = N + 3 points (compute_1d_matrices, :308-556):
// S[nl][nl] as S[component][eigenvector] with S^T B S = I, lam ascending, rows
// of the ghost / Dirichlet end points zeroed.  Returns false if B is not
// positive definite.
bool fdm_1d(int N, const double *D, const double *w, int lbc, int rbc, double ll, double lm,
            double lr, double *S, double *lam);

template <typename T>
struct Schwarz {
  int N = 0, Nq = 0, Np = 0, nl = 0, E = 0;
  bool ras = false;
  int fdmPrecision = NRS_FDM_F32;
  DevBuf<T> S;              // [E][3][nl][nl]
  DevBuf<double> lam;       // [E][3][nl]
  DevBuf<uint16_t> S16;     // BF16/FP16 operators, rows in MFMA storage order
  DevBuf<float> lam32;
  DevBuf<int4> gmap;        // [E][6] neighbour second layer: base, du, dv (base < 0: none)
  DevBuf<int4> smap;        // [E][6] neighbour overlap-result slot (ASM)
  DevBuf<T> wts;            // ASM weights [E*Np] (RAS uses invDegree)
  DevBuf<T> boxOut;         // ASM: restricted FDM result [E*Np]
  DevBuf<T> ghostOut;       // ASM: overlap results [E][6][Nq][Nq]
  const T *invDegree = nullptr;
  const nrs_ogs *ogs = nullptr;
  int64_t Nmasked = 0;
  const int *maskIds = nullptr;
  // [E][ceil(Np/32)] Dirichlet mask bits of the level's nodes: the ASM
  // overlap kernel and the SIMT box solve store the masked nodes of Su as zero
  // (null, or RAS on the matrix cores: the mask runs as its own pass over
  // maskIds).  The masked nodes must be in no row of ogs, as the level's
  // handles are built.
  const uint32_t *maskBits = nullptr;
  double helm0 = 1.0, helm1 = 0.0;   // box solve inverts helm0 (lx + ly + lz) + helm1
  float2 helm32() const { return make_float2((float)helm0, (float)helm1); }

  // lengths: [9][E] of the finest level; EToB [E][6]; links of the mesh.
  // helmholtz: operator lambda0 * (stiffness) + lambda1 * (mass) with constant
  // coefficients; the box solve then inverts lambda0 (Lx + Ly + Lz) + lambda1
  // in the eigenbasis (S^T B S = I per direction).
  void setup(int N, int E, bool ras, int fdmPrecision, const std::vector<double> &lengths,
             const int *EToB, const std::vector<FaceLink> &links, const nrs_ogs *ogs,
             const T *invDegree, int64_t Nmasked, const int *maskIds, const uint32_t *maskBits,
             bool poisson, double lambda0, double lambda1, hipStream_t s);
  // Su = smoothSchwarz(u): FDM solves on the extended boxes, overlap exchange,
  // gather-scatter, mask.  Su must not alias u (the RAS kernel reads the
  // neighbours' u while writing its own Su).
  void apply(const T *u, T *Su, hipStream_t s);
  void finish(T *Su, bool masked, hipStream_t s);
  // choose the matrix-core kernel's elements per wavefront for this level
  // (FP32 vectors with BF16/FP16 operators; otherwise nothing to tune); returns it
  int autotune(hipStream_t s);
};

}  // namespace nrs
