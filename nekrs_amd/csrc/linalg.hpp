// linalg.hpp -- BLAS-1 on E-vectors and device-resident reductions.
#pragma once
#include "common.hpp"

namespace nrs {

// Slots of the per-solver scalar block kept in device memory.  Reductions write
// their results here and the update kernels read alpha/beta from here, so an
// iteration needs no host round trip except the convergence check.
enum Scalar : int {
  S_RDOTZ1 = 0, S_RDOTZ2, S_ZDOTAP, S_ALPHA, S_BETA, S_PAP, S_RDOTR, S_TOL, S_RESNORM,
  S_ITER, S_TMP0, S_TMP1, S_TMP2, S_TMP3, S_COUNT = 16
};

// Deterministic two-level reduction: a fixed grid writes one partial per
// block, the last block to finish (atomic ticket) sums them in block order.
struct Reducer {
  static constexpr int kBlocks = 1024;
  static constexpr int kThreads = 256;
  DevBuf<double> partials;     // [kBlocks * 4]
  DevBuf<unsigned> ticket;     // [1], returns to 0 after every reduction
  Reducer();
};

// Weighted inner products into device scalars: out[slot] = sum w a b.
void wdot(Reducer &R, int64_t n, const double *w, const double *a, const double *b,
          double *scal, int slot, hipStream_t s);
void wdot_f32(Reducer &R, int64_t n, const float *w, const float *a, const float *b, double *scal,
              int slot, hipStream_t s);

// scal[slot] = sum_e sum_{n<Np} x[e*estride + off + n]
void plane_sum(Reducer &R, int64_t E, int Np, int64_t estride, int64_t off, const double *x,
               double *scal, int slot, hipStream_t s);

template <typename T> void fill(int64_t n, T v, T *x, hipStream_t s);
// y = a x + b y
template <typename T> void axpby(int64_t n, T a, const T *x, T b, T *y, hipStream_t s);
// z = x .* y
template <typename T> void axmyz(int64_t n, const T *x, const T *y, T *z, hipStream_t s);
// y = 1 / y
template <typename T> void reciprocal(int64_t n, T *y, hipStream_t s);
template <typename A, typename B> void convert(int64_t n, const A *x, B *y, hipStream_t s);

// ---- PCG pieces (src/solvers/elliptic/linearSolver/PCG.cpp:33-203) ----
// Jacobi preconditioner fused with the dots: z = invDiag .* r; rdotz = <r,z>_w,
// zdotAp = <z,Ap>_w (flexible); then beta from rdotz1/rdotz2 (or -alpha zdotAp/rdotz2).
void pcg_precon_dots(Reducer &R, int64_t n, const double *w, const double *invDiag,
                     const double *r, double *z, const double *Ap, bool jacobi, bool flexible,
                     bool none_precon, int iter, double *scal, hipStream_t s);
// p = z + beta p
void pcg_update_p(int64_t n, const double *z, double *p, const double *scal, hipStream_t s);
// alpha = rdotz1 / (<p,Ap>_w + 1e-300)
void pcg_alpha(Reducer &R, int64_t n, const double *w, const double *p, const double *Ap,
               double *scal, hipStream_t s);
// r -= alpha Ap; x += alpha p; rdotr = sqrt(<r,r>_w * resNormFactor)
void pcg_update_rx(Reducer &R, int64_t n, const double *w, const double *Ap, const double *p,
                   double *r, double *x, double resNormFactor, double *scal, hipStream_t s);
// scal[slot] = sqrt(<a,a>_w * factor)
void wnorm(Reducer &R, int64_t n, const double *w, const double *a, double factor, double *scal,
           int slot, hipStream_t s);

}  // namespace nrs
