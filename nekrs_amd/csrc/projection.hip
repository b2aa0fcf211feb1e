// projection.hip -- initial guess from the projection on previous solutions.
//
// Restates SolutionProjection (src/solvers/elliptic/ellipticSolutionProjection.cpp)
// with the PGMRES pass kernels: the weighted multi-dot (weightedInnerProdMulti),
// the one-pass subtraction of a combination (multiScaledAddwOffset) and the
// accumulation (accumulate.okl).  The A-norm test of a new basis vector is
// done on the host after one readback of alpha, as in the reference.
#include <cmath>

#include "elliptic.hpp"
#include "gmres.hpp"

namespace nrs {

namespace {

bool classic(const nrs_elliptic *el) { return el->opts.initialGuess == NRS_GUESS_PROJECTION; }
int64_t fields_len(const nrs_elliptic *el) { return (int64_t)el->Nf * el->Nlocal; }
const double *weight(const nrs_elliptic *el)
{
  return el->Nf > 1 ? el->wRep.get() : el->ogs->invDegree.get();
}

// y = sum_{k<m} alpha_k X_k (accumulate.okl)
void accumulate(nrs_elliptic *el, int m, const double *X, double *y, hipStream_t s)
{
  const int64_t L = fields_len(el);
  fill<double>(L, 0.0, y, s);
  gmres_update(L, m, L, el->pjAlpha.get(), X, y, s);
}

// updateProjectionSpace (ellipticSolutionProjection.cpp:37-139): alpha_k =
// <xx_k, bb>; xx_{m-1} -= sum_{k<m-1} alpha_k xx_k (and bb likewise for
// CLASSIC); keep it scaled to unit A-norm, or drop it when the remaining norm
// is below 1e-7 of the original (the reference's test, kept as written).
void update_space(nrs_elliptic *el, hipStream_t s)
{
  const int64_t L = fields_len(el);
  const int m = el->pjM;
  if (m <= 0) return;
  double *XX = el->pjXX.get(), *BB = el->pjBB.get(), *alpha = el->pjAlpha.get();
  const double *bb = BB + (classic(el) ? (int64_t)(m - 1) * L : 0);
  double *nrm = el->gY.get() + kGmresMaxVectors;   // unused norm of the subtraction pass
  gmres_multi_dot(L, m, L, weight(el), XX, bb, el->gPart.get(), alpha, s);
  if (m > 1) {
    gmres_gram_schmidt(L, m - 1, L, weight(el), alpha, XX, XX + (int64_t)(m - 1) * L, el->gPart.get(),
                       nrm, s);
    if (classic(el))
      gmres_gram_schmidt(L, m - 1, L, weight(el), alpha, BB, BB + (int64_t)(m - 1) * L, el->gPart.get(),
                         nrm, s);
  }
  el->pjAlphaHost.resize(m);
  NRS_HIP(hipMemcpyAsync(el->pjAlphaHost.data(), alpha, m * sizeof(double), hipMemcpyDeviceToHost, s));
  NRS_HIP(hipStreamSynchronize(s));
  const double normOrig = el->pjAlphaHost[m - 1];
  double sumAlpha = 0.0;
  for (int k = 0; k < m - 1; ++k) sumAlpha += el->pjAlphaHost[k] * el->pjAlphaHost[k];
  const double normNew = std::sqrt(normOrig - sumAlpha);
  const double test = normNew / normOrig;
  if (test > 1e-7) {
    const double scale = 1.0 / normNew;
    axpby<double>(L, scale, XX + (int64_t)(m - 1) * L, 0.0, XX + (int64_t)(m - 1) * L, s);
    if (classic(el)) axpby<double>(L, scale, BB + (int64_t)(m - 1) * L, 0.0, BB + (int64_t)(m - 1) * L, s);
  } else {
    --el->pjM;   // linearly dependent: discard
  }
}

}  // namespace

// SolutionProjection::pre + computePreProjection (:141-192, :269-280)
bool projection_pre(nrs_elliptic *el, double *r, hipStream_t s)
{
  ++el->pjStep;
  if (el->pjStep < el->opts.projStart) return false;
  if (el->pjM <= 0) return false;
  el->pjPrev = el->pjM;
  const int64_t L = fields_len(el);
  const int m = el->pjM;
  gmres_multi_dot(L, m, L, weight(el), el->pjXX.get(), r, el->gPart.get(), el->pjAlpha.get(), s);
  accumulate(el, m, el->pjXX.get(), el->pjXbar.get(), s);            // xbar = sum alpha_k xx_k
  if (classic(el)) {
    double *rtmp = el->zv.get();
    accumulate(el, m, el->pjBB.get(), rtmp, s);                        // sum alpha_k bb_k
    axpby<double>(L, -1.0, rtmp, 1.0, r, s);
  } else {
    elliptic_operator(el, el->pjXbar.get(), el->pjBB.get(), true, s);  // bb = A xbar
    axpby<double>(L, -1.0, el->pjBB.get(), 1.0, r, s);
  }
  return true;
}

// SolutionProjection::post + computePostProjection (:194-235, :282-288)
void projection_post(nrs_elliptic *el, double *x, hipStream_t s)
{
  if (el->pjStep < el->opts.projStart) return;
  const int64_t L = fields_len(el);
  const size_t bytes = L * sizeof(double);
  double *XX = el->pjXX.get(), *BB = el->pjBB.get();
  if (el->pjM == 0) {                                  // first vector
    el->pjM = 1;
    NRS_HIP(hipMemcpyAsync(XX, x, bytes, hipMemcpyDeviceToDevice, s));
  } else if (el->pjM == el->opts.projVectors) {        // full: restart from this solution
    el->pjM = 1;
    axpby<double>(L, 1.0, el->pjXbar.get(), 1.0, x, s);
    NRS_HIP(hipMemcpyAsync(XX, x, bytes, hipMemcpyDeviceToDevice, s));
  } else {                                             // xx_{m-1} = x (the correction), x += xbar
    ++el->pjM;
    NRS_HIP(hipMemcpyAsync(XX + (int64_t)(el->pjM - 1) * L, x, bytes, hipMemcpyDeviceToDevice, s));
    axpby<double>(L, 1.0, el->pjXbar.get(), 1.0, x, s);
  }
  const int before = el->pjM;
  const int64_t bOff = classic(el) ? (int64_t)(el->pjM - 1) * L : 0;
  elliptic_operator(el, XX + (int64_t)(el->pjM - 1) * L, BB + bOff, true, s);
  update_space(el, s);
  if (el->pjM < before) {                              // dependent: reset the space to x
    el->pjM = 1;
    NRS_HIP(hipMemcpyAsync(XX, x, bytes, hipMemcpyDeviceToDevice, s));
    elliptic_operator(el, XX, BB, true, s);
    update_space(el, s);
  }
}

}  // namespace nrs
