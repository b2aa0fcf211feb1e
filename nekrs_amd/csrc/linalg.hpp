// linalg.hpp -- BLAS-1 on E-vectors and device-resident reductions.
#pragma once
#include "common.hpp"

namespace nrs {

// Slots of the per-solver scalar block kept in device memory.  Reductions write
// their results here and the update kernels read alpha/beta from here, so an
// iteration needs no host round trip except the convergence check.
enum Scalar : int {
  S_RDOTZ1 = 0, S_RDOTZ2, S_ZDOTAP, S_ALPHA, S_BETA, S_PAP, S_RDOTR, S_TOL, S_RESNORM,
  S_ITER,     // completed iterations
  S_ACTIVE,   // 1 while the iteration runs; kernels of a finished solve return at once
  S_MAXIT, S_TMP0, S_TMP1, S_TMP2, S_TMP3, S_COUNT = 16
};

// Deterministic two-launch reduction: a grid sized to the vector writes one
// partial per block, then a one-block kernel adds the partials in a fixed tree
// order and applies the scalar step (alpha, beta, norms).  No inter-block
// atomics or device-scope fences are needed (a device-scope release per block
// would write back the per-XCD L2 on every block).
struct Reducer {
  static constexpr int kBlocks = 1024;   // upper bound on the partials grid
  static constexpr int kThreads = 256;
  DevBuf<double> partials;               // [kBlocks * 2]
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
// The iteration state lives in device scalars: S_ITER counts completed
// iterations, S_ACTIVE turns every PCG kernel into a no-op once the residual
// reached S_TOL or S_ITER reached S_MAXIT, so any number of iterations can be
// enqueued ahead of the host's convergence check without changing the result.
void pcg_init(double *scal, double rdotr0, double tol, int maxit, hipStream_t s);
// pcg_init from the initial norm already in scal[S_RDOTR] (kept in S_TMP3)
void pcg_init_device(double *scal, double tol, bool relative, int maxit, hipStream_t s);
// Jacobi preconditioner fused with the dots: z = invDiag .* r; rdotz = <r,z>_w,
// zdotAp = <z,Ap>_w (flexible); then beta from rdotz1/rdotz2 (or -alpha zdotAp/rdotz2).
void pcg_precon_dots(Reducer &R, int64_t n, const double *w, const double *invDiag,
                     const double *r, double *z, const double *Ap, bool jacobi, bool flexible,
                     bool none_precon, double *scal, hipStream_t s);
// p = z + beta p
void pcg_update_p(int64_t n, const double *z, double *p, const double *scal, hipStream_t s);
// alpha = rdotz1 / (<p,Ap>_w + 1e-300)
void pcg_alpha(Reducer &R, int64_t n, const double *w, const double *p, const double *Ap,
               double *scal, hipStream_t s);
// r -= alpha Ap; x += alpha p; rdotr = sqrt(<r,r>_w * resNormFactor)
void pcg_update_rx(Reducer &R, int64_t n, const double *w, const double *Ap, const double *p,
                   double *r, double *x, double resNormFactor, double *scal, hipStream_t s);
// linAlg_t reductions over n entries of nf fields spaced off apart (op: nrs_reduction);
// partials: device scratch of at least reduce_partials_needed(n) doubles; out: one
// device double.  Deterministic: fixed per-thread order, fixed tree order.
int reduce_partials_needed(int64_t n);
constexpr int RED_WSUM = 16;   // internal: sum_i w_i x_i (not part of the C ABI)
void linalg_reduce(int op, int64_t n, int nf, int64_t off, const double *w, const double *x,
                   const double *y, double *partials, double *out, hipStream_t s);
// ellipticBlockUpdatePCG: r -= alpha Ap (nf fields), out = sum_i w_i sum_f r^2
void block_update_pcg(int64_t n, int nf, int64_t off, const double *w, const double *Ap, double alpha,
                      double *r, double *partials, double *out, hipStream_t s);
// y += factor * scal[slot]
void add_device_scalar(int64_t n, double factor, const double *scal, int slot, double *y, hipStream_t s);
// q -= (sum_i w_i q_i) / wsum (w = null: plain sum), the sum kept in scal[slot];
// enqueued work only (no host round trip)
void remove_mean(Reducer &R, int64_t n, const double *w, double wsum, double *q, double *scal, int slot,
                 hipStream_t s);
// scal[slot] = sqrt(<a,a>_w * factor)
void wnorm(Reducer &R, int64_t n, const double *w, const double *a, double factor, double *scal,
           int slot, hipStream_t s);

}  // namespace nrs
