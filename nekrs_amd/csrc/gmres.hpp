// gmres.hpp -- vector kernels of restarted (flexible) PGMRES
// (src/solvers/elliptic/linearSolver/PGMRES.cpp:111-340).  Vectors V_j, Z_j
// are ld apart; W is the inner-product weight (invDegree); results land in
// device memory (y[k], nw, nr) for one readback per iteration.
#pragma once
#include "common.hpp"

namespace nrs {

constexpr int kGmresMaxVectors = 32;
constexpr int kGmresPartials = kGmresMaxVectors * 4096;   // partial sums of one pass

// y_j = sum_i W_i V_j[i] x_i, j < k
void gmres_multi_dot(int64_t n, int k, int64_t ld, const double *w, const double *V, const double *x,
                     double *partials, double *y, hipStream_t s);
// x -= sum_{j<k} y_j V_j; nw = sum_i W_i x_i^2
void gmres_gram_schmidt(int64_t n, int k, int64_t ld, const double *w, const double *y, const double *V,
                        double *x, double *partials, double *nw, hipStream_t s);
// x += sum_{j<k} y_j Z_j
void gmres_update(int64_t n, int k, int64_t ld, const double *y, const double *Z, double *x, hipStream_t s);
// r = b - Ax; nr = sum_i W_i r_i^2
void gmres_residual(int64_t n, const double *w, const double *b, const double *Ax, double *r,
                    double *partials, double *nr, hipStream_t s);

}  // namespace nrs
