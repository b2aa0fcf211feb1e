// tombo.hpp -- the pressure right-hand side of the TOMBO splitting
// (tombo::pressureSolve, src/navierStokes/tombo.cpp:11-121), the step that
// produces the pressure Poisson solve's input.
#pragma once
#include "common.hpp"
#include "ogs.hpp"

struct nrs_tombo {
  int N = 0, Nq = 0, Np = 0, E = 0;
  int64_t Nlocal = 0;
  nrs::DevBuf<double> D, w;
  nrs::DevBuf<double> vgeo;          // [E][10][Np]: rx ry rz sx sy sz tx ty tz JW
  nrs::DevBuf<double> invLMM;        // [Nlocal] 1 / gs(JW)
  nrs_ogs *ogs = nullptr;            // unmasked (nrs->gsh)
  // N = 7: faces i = N / i = 0 shared by consecutive elements (same block of
  // four, fx[e] = 1) are summed in the curl kernels; ogsOp leaves them out
  nrs_ogs *ogsOp = nullptr;
  nrs::DevBuf<uint8_t> fx;
  int64_t fusedPairs = 0;
  int Nb = 0;                        // elements with at least one flux face
  nrs::DevBuf<int> bElem, bType;     // [Nb] element, [Nb][6] face boundary type
  nrs::DevBuf<double> bGeo;          // [Nb][6][Nq*Nq][4]: nx ny nz WSJ
  nrs::DevBuf<double> W, F;          // [3][Nlocal] work: curl, then the assembled rhsU
  nrs::DevBuf<double> S;             // [3][Nlocal] stress term (NRS_TOMBO_STRESS), else empty
  // trilinear elements: the volume kernels form rx..tz, JW from the corners
  // [E][3][8] and the GLL nodes and weights zw [2][Nq] instead of reading vgeo
  bool tri = false;
  nrs::DevBuf<double> corners, zw;
  ~nrs_tombo();
};

namespace nrs {

// the reference's boundary type ids (src/bdry/bcType.h)
enum { BC_W = 1, BC_INT = 2, BC_V = 3, BC_SYMX = 4, BC_SYMY = 5, BC_SYMZ = 6, BC_SYM = 7 };

nrs_tombo *tombo_setup(int N, int E, const double *x, const double *y, const double *z,
                       const long long *gids, const int *bcType, int flags);
// W[c] = JW * curl(U)[c] (curlHex3D with scaleJW = 1); fields off apart
void tombo_curl(const nrs_tombo *t, int64_t off, const double *U, double *W, hipStream_t s);
void tombo_pressure_rhs(nrs_tombo *t, int64_t off, const double *Ue, const double *U,
                        const double *BF, const double *mue, const double *irho, const double *div,
                        double g0idt, bool visc, double *rhs, hipStream_t s);

}  // namespace nrs
