// elliptic.hip -- elliptic_t setup, operator, preconditioner dispatch and the
// ellipticSolve/pcg driver (host C++ enqueueing HIP kernels on the caller's stream).
#include <chrono>
#include <cmath>
#include <cstring>

#include "elliptic.hpp"
#include "gmres.hpp"
#include "tune.hpp"
#include "kernels.hpp"
#include "mesh.hpp"
#include "mg.hpp"

nrs_elliptic::~nrs_elliptic()
{
  if (ogs) delete ogs;
  if (ogsOp) delete ogsOp;
  if (ogsFull) delete ogsFull;
  if (scalHost) (void)hipHostFree(scalHost);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (evPre0) (void)hipEventDestroy(evPre0);
  if (evPre1) (void)hipEventDestroy(evPre1);
  for (auto e : eventPool) (void)hipEventDestroy(e);
  for (auto e : chunkEv) if (e) (void)hipEventDestroy(e);
  if (auxDone) (void)hipEventDestroy(auxDone);
  if (aux) (void)hipStreamDestroy(aux);
}

hipEvent_t nrs_elliptic::next_event()
{
  if (eventNext == eventPool.size()) {
    hipEvent_t e;
    NRS_HIP(hipEventCreate(&e));
    eventPool.push_back(e);
  }
  return eventPool[eventNext++];
}

namespace nrs {

std::vector<uint32_t> mask_bits(int Np, int E, const std::vector<int> &maskIds)
{
  const int NW = (Np + 31) / 32;
  std::vector<uint32_t> bits((size_t)E * NW, 0u);
  for (int id : maskIds) {
    const int64_t e = id / Np;
    const int n = (int)(id - e * Np);
    bits[e * NW + (n >> 5)] |= (1u << (n & 31));
  }
  return bits;
}

// Dirichlet mask: restates ellipticOgs (src/solvers/elliptic/ellipticOgs.cpp:17-60):
// per node the smallest BC id of the faces it lies on (DIRICHLET=1 wins), reduced
// with gs-min over the unmasked handle, then every node equal to DIRICHLET.
std::vector<int> dirichlet_mask(int N, int E, const int *EToB, nrs_ogs *ogsFull, hipStream_t s)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  const int64_t L = (int64_t)E * Np;
  const int large = 1 << 20;
  std::vector<int> mapB(L, large);
  for (int64_t e = 0; e < E; ++e)
    for (int f = 0; f < 6; ++f) {
      const int bc = EToB[e * 6 + f];
      if (bc <= 0) continue;
      for (int b = 0; b < Nq; ++b)
        for (int a = 0; a < Nq; ++a) {
          int i, j, k;
          switch (f) {  // FaceNodesHex3D (src/mesh/meshBasisHex3D.cpp:53-82)
            case 0: i = a; j = b; k = 0; break;
            case 1: i = a; j = 0; k = b; break;
            case 2: i = N; j = a; k = b; break;
            case 3: i = a; j = N; k = b; break;
            case 4: i = 0; j = a; k = b; break;
            default: i = a; j = b; k = N; break;
          }
          int &m = mapB[e * Np + i + Nq * (j + Nq * k)];
          if (bc < m) m = bc;
        }
    }
  DevBuf<int> d;
  d.upload(mapB);
  ogs_gather_scatter<int>(ogsFull, d.get(), 1, L, NRS_MIN, s);
  NRS_HIP(hipStreamSynchronize(s));
  mapB = d.download();
  std::vector<int> ids;
  for (int64_t n = 0; n < L; ++n)
    if (mapB[n] == 1) ids.push_back((int)n);
  return ids;
}

std::vector<uint8_t> face_fusion(int N, int E, std::vector<long long> &ids, const std::vector<double> &invDegree,
                                 int nb)
{
  std::vector<uint8_t> fx;
  if (N != 7 || E < 2) return fx;
  constexpr int Nq = 8, Np = 512;
  fx.assign(E, 0);
  int64_t pairs = 0;
  for (int64_t e = 0; e + 1 < E; ++e) {
    if (e % nb == nb - 1) continue;                 // partner in the next block
    bool ok = true;
    for (int k = 1; k < Nq - 1 && ok; ++k)
      for (int j = 1; j < Nq - 1 && ok; ++j) {
        const int64_t ia = e * Np + k * 64 + j * 8 + 7, ib = (e + 1) * Np + k * 64 + j * 8;
        // the pair must be the node's only two copies: a third copy elsewhere would lose
        // its contribution once the pair leaves the operator's gather-scatter
        ok = ids[ia] != 0 && ids[ia] == ids[ib] && invDegree[ia] == 0.5;
      }
    if (!ok) continue;
    fx[e] = 1;
    ++pairs;
  }
  if (!pairs) { fx.clear(); return fx; }
  for (int64_t e = 0; e + 1 < E; ++e) {
    if (!fx[e]) continue;
    for (int k = 1; k < Nq - 1; ++k)
      for (int j = 1; j < Nq - 1; ++j) {
        ids[e * Np + k * 64 + j * 8 + 7] = 0;
        ids[(e + 1) * Np + k * 64 + j * 8] = 0;
      }
  }
  return fx;
}

// Elements per operator chunk: kOpChunks chunks, each a multiple of four elements
// so that chunks hold whole blocks of the fused Ax kernel.
int64_t op_chunk_elements(int64_t E)
{
  constexpr int F = nrs_elliptic::kOpChunks;
  return ((E + F - 1) / F + 3) / 4 * 4;
}

namespace {
// A form of the operator face fusion: blocks of nb elements (nb waves).
struct FusionForm { int nb; };
struct FusedOperator {
  std::unique_ptr<nrs_ogs> ogs;     // the operator's handle without the fused face nodes
  DevBuf<uint8_t> fx;
  int64_t pairs = 0;
};

// The handle and flags of a fusion form; false when no pair of elements qualifies.
bool make_fused(nrs_elliptic *el, const std::vector<long long> &maskedIds, FusionForm f, FusedOperator &c)
{
  std::vector<long long> ids(maskedIds);
  const std::vector<uint8_t> fx = face_fusion(el->N, el->E, ids, el->ogs->invDegreeHost, f.nb);
  if (fx.empty()) return false;
  c.ogs.reset(ogs_setup(el->Nlocal, ids.data()));
  ogs_set_chunks(c.ogs.get(), op_chunk_elements(el->E) * el->Np, nrs_elliptic::kOpChunks);
  c.fx.upload(fx);
  c.pairs = 0;
  for (uint8_t v : fx) c.pairs += v;
  return true;
}

void keep_fused(nrs_elliptic *el, FusionForm f, FusedOperator &c)
{
  delete el->ogsOp;
  el->ogsOp = c.ogs.release();
  el->fx = std::move(c.fx);
  el->fusedPairs = c.pairs;
  el->fusionBlock = f.nb;
}
}  // namespace

nrs_elliptic *elliptic_setup(const nrs_elliptic_desc *d, const nrs_elliptic_opts *o)
{
  NRS_CHECK(d && o, "elliptic setup: null descriptor/options");
  NRS_CHECK(d->N >= 1 && d->N <= 9, "elliptic setup: N must be 1..9");
  NRS_CHECK(d->Nelements > 0, "elliptic setup: no elements");
  NRS_CHECK(d->x && d->y && d->z && d->globalIds && d->EToB, "elliptic setup: null mesh arrays");
  NRS_CHECK(o->solver >= NRS_SOLVER_PCG && o->solver <= NRS_SOLVER_FGMRES, "unknown solver");
  NRS_CHECK(o->solver < NRS_SOLVER_GMRES || (o->gmresRestart >= 1 && o->gmresRestart <= kGmresMaxVectors),
            "PGMRES restart must be 1..32");
  NRS_CHECK(o->preconditioner >= NRS_PRECON_NONE && o->preconditioner <= NRS_PRECON_USER,
            "unknown preconditioner");
  NRS_CHECK(o->initialGuess >= NRS_GUESS_PREVIOUS && o->initialGuess <= NRS_GUESS_PROJECTION_ACONJ,
            "unknown initial guess");
  NRS_CHECK(o->initialGuess == NRS_GUESS_PREVIOUS ||
                (o->projVectors >= 1 && o->projVectors <= kGmresMaxVectors),
            "residual projection vectors must be 1..32");
  NRS_CHECK(o->maxIter >= 1, "maxIter must be >= 1");
  // ellipticSetup.cpp:59-73: multigrid with non-Jacobi smoothers and its coarse solve
  // are Poisson-only in the reference; kept as in the reference.
  auto el = std::make_unique<nrs_elliptic>();
  el->N = d->N; el->Nq = d->N + 1; el->Np = el->Nq * el->Nq * el->Nq; el->E = d->Nelements;
  el->Nf = d->Nfields <= 1 ? 1 : d->Nfields;
  NRS_CHECK(el->Nf == 1 || el->Nf == 3, "elliptic setup: Nfields must be 1 or 3");
  // a block solve with multigrid applies the (single-field) hierarchy to each
  // field; the reference's velocity solve uses Jacobi there
  el->Nlocal = (int64_t)el->E * el->Np;
  NRS_CHECK(el->Nlocal < (int64_t)INT32_MAX, "E*Np must fit in int32 (dlong)");
  el->poisson = d->poisson != 0;
  // allNeumann: Poisson and no face of a boundary type other than NEUMANN (2)
  // (ellipticSetup.cpp:187-206)
  if (el->poisson) {
    bool all = true;
    for (int64_t f = 0; f < (int64_t)el->E * 6 && all; ++f)
      if (d->EToB[f] > 0 && d->EToB[f] != 2) all = false;
    el->allNeumann = all;
  }
  if (el->allNeumann && el->Nf > 1)   // ellipticZeroMean.cpp:36-38 aborts for block solvers
    throw Error(NRS_ERR_UNSUPPORTED, "elliptic setup: the null space of an all-Neumann block solve "
                                     "is not supported (as in the reference)");
  el->lambda0 = d->lambda0;
  el->lambda1 = d->poisson ? 0.0 : d->lambda1;
  if (d->lambda0Field || d->lambda1Field) {
    NRS_CHECK(d->lambda0Field, "elliptic setup: lambda1Field without lambda0Field");
    NRS_CHECK(el->poisson || d->lambda1Field, "elliptic setup: a Helmholtz coefficient field needs lambda1Field");
    if (o->preconditioner == NRS_PRECON_MULTIGRID)
      throw Error(NRS_ERR_UNSUPPORTED, "elliptic setup: coefficient fields with the multigrid preconditioner "
                                       "(its levels take constant coefficients)");
    el->lf0 = d->lambda0Field;
    el->lf1 = el->poisson ? nullptr : d->lambda1Field;
  }
  el->opts = *o;
  hipStream_t s = nullptr;

  gll_nodes(el->N, el->z, el->w);
  el->Dh = dmatrix(el->z);
  el->D.upload(el->Dh);
  el->lam0.upload(std::vector<double>{el->lambda0});
  el->lam1.upload(std::vector<double>{el->lambda1});

  // geometric factors on the device (meshGeometricFactorsHex3D)
  {
    DevBuf<double> x, y, zc, wd;
    x.upload(d->x, el->Nlocal); y.upload(d->y, el->Nlocal); zc.upload(d->z, el->Nlocal);
    wd.upload(el->w);
    el->ggeo.alloc((size_t)el->Nlocal * 7);
    launch_geofacs(el->N, el->E, el->D.get(), wd.get(), x.get(), y.get(), zc.get(), el->ggeo.get(), s);
    NRS_HIP(hipStreamSynchronize(s));
  }
  el->red = std::make_unique<Reducer>();
  el->scal.alloc(S_COUNT);
  NRS_HIP(hipMemset(el->scal.get(), 0, S_COUNT * sizeof(double)));
  NRS_HIP(hipHostMalloc((void **)&el->scalHost, S_COUNT * sizeof(double), hipHostMallocDefault));
  plane_sum(*el->red, el->E, el->Np, 7 * (int64_t)el->Np, 6 * (int64_t)el->Np, el->ggeo.get(),
            el->scal.get(), S_TMP0, s);
  NRS_HIP(hipMemcpy(el->scalHost, el->scal.get(), S_COUNT * sizeof(double), hipMemcpyDeviceToHost));
  el->volume = el->scalHost[S_TMP0];

  // gather-scatter handles and the Dirichlet mask
  el->ogsFull = ogs_setup(el->Nlocal, d->globalIds);
  const std::vector<int> mids = dirichlet_mask(el->N, el->E, d->EToB, el->ogsFull, s);
  el->Nmasked = (int64_t)mids.size();
  el->maskIds.upload(mids);
  el->maskBits.upload(mask_bits(el->Np, el->E, mids));
  std::vector<long long> mg(d->globalIds, d->globalIds + el->Nlocal);   // Dirichlet ids zeroed
  std::vector<FusionForm> fusionForms;
  {
    for (int id : mids) mg[id] = 0;
    el->ogs = ogs_setup(el->Nlocal, mg.data());
    // operator face fusion forms to try (elements per block): with autotune on,
    // blocks of four and of two are timed against no fusion
    if (o->operatorFusion && el->Nf == 1 && !el->lfield() && el->N == 7) {
      if (o->operatorFusion == 2) fusionForms = {{2}};
      else if (o->autotune) fusionForms = {{4}, {2}};
      else fusionForms = {{4}};
    }
  }

  // operator chunks: rows of the operator's handle grouped by element chunk
  // (chunks of a multiple of four elements, whole blocks of the fused kernel)
  {
    constexpr int F = nrs_elliptic::kOpChunks;
    ogs_set_chunks(el->ogs, op_chunk_elements(el->E) * el->Np, F);
    NRS_HIP(hipStreamCreateWithFlags(&el->aux, hipStreamNonBlocking));
    for (auto &e : el->chunkEv) NRS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    NRS_HIP(hipEventCreateWithFlags(&el->auxDone, hipEventDisableTiming));
  }

  // operator launch variants (KERNEL AUTOTUNING, registerEllipticKernels.cpp:170-183)
  if (o->autotune) {
    autotune_ax<double>(el->N, el->E, el->ggeo.get(), el->D.get(), el->lam0.get(), el->lam1.get(),
                        el->poisson, s);
    // operator chunks for the unfused form, then -- where the face fusion applies -- for the
    // fused one; the fusion is kept only where it is faster (the results are the same)
    auto op_ms = [&](const nrs_ogs *h, const uint8_t *fx, int chunks) {
      DevBuf<double> q(el->Nlocal), out(el->Nlocal);
      fill_pattern<double>(el->Nlocal, q.get(), s);
      return time_launches([&] {
        chunked_operator<double>(el.get(), el->N, el->E, el->ggeo.get(), el->D.get(), el->lam0.get(),
                                 el->lam1.get(), el->poisson, el->maskBits.get(), h, q.get(), out.get(),
                                 chunks, s, false, fx);
      }, s);
    };
    const int cUnf = autotune_operator<double>(el.get(), el->N, el->E, el->ggeo.get(), el->D.get(),
                                               el->lam0.get(), el->lam1.get(), el->poisson,
                                               el->maskBits.get(), el->ogs, s);
    double tBest = fusionForms.empty() ? 0.0 : op_ms(el->ogs, nullptr, cUnf);
    int cBest = cUnf;
    for (const FusionForm &f : fusionForms) {
      FusedOperator c;
      if (!make_fused(el.get(), mg, f, c)) continue;
      // the kernels read the block size from the handle while this form is timed
      // (restored on exit, also when the timing throws)
      struct Form {
        nrs_elliptic *el; int nb;
        Form(nrs_elliptic *e, FusionForm f) : el(e), nb(e->fusionBlock) { el->fusionBlock = f.nb; }
        ~Form() { el->fusionBlock = nb; }
      } guard(el.get(), f);
      const int cf = autotune_operator<double>(el.get(), el->N, el->E, el->ggeo.get(), el->D.get(),
                                               el->lam0.get(), el->lam1.get(), el->poisson,
                                               el->maskBits.get(), c.ogs.get(), s, c.fx.get());
      const double tf = op_ms(c.ogs.get(), c.fx.get(), cf);
      if (tf < tBest) {
        tBest = tf;
        cBest = cf;
        keep_fused(el.get(), f, c);
        guard.nb = f.nb;
      }
    }
    set_tuned(TUNE_OP, el->N, NRS_F64, el->E, cBest);
  } else if (!fusionForms.empty()) {
    FusedOperator c;
    if (make_fused(el.get(), mg, fusionForms[0], c)) keep_fused(el.get(), fusionForms[0], c);
  }

  // workspace
  const int64_t Lt = (int64_t)el->Nf * el->Nlocal;
  el->p.alloc(Lt); el->zv.alloc(Lt); el->Ap.alloc(Lt); el->x0.alloc(Lt);
  if (o->solver == NRS_SOLVER_GMRES || o->solver == NRS_SOLVER_FGMRES) {
    // PGMRES workspace (GmresData, PGMRES.cpp:30-57): V [m][Lt], Z [m or 1][Lt]
    const int m = o->gmresRestart;
    el->gV.alloc((size_t)m * Lt);
    el->gZ.alloc((size_t)(o->solver == NRS_SOLVER_FGMRES ? m : 1) * Lt);
    el->gB.alloc(Lt);
  }
  if (o->solver == NRS_SOLVER_GMRES || o->solver == NRS_SOLVER_FGMRES ||
      o->initialGuess != NRS_GUESS_PREVIOUS) {
    el->gY.alloc(kGmresMaxVectors + 2);
    el->gPart.alloc(kGmresPartials);
  }
  if (o->initialGuess != NRS_GUESS_PREVIOUS) {
    // SolutionProjection constructor (ellipticSolutionProjection.cpp:237-267)
    const int mv = o->projVectors;
    el->pjXX.alloc((size_t)mv * Lt);
    el->pjBB.alloc((size_t)(o->initialGuess == NRS_GUESS_PROJECTION ? mv : 1) * Lt);
    el->pjXbar.alloc(Lt);
    el->pjAlpha.alloc(mv);
  }
  if (el->Nf > 1) {
    el->wRep.alloc(Lt);
    for (int f = 0; f < el->Nf; ++f)
      NRS_HIP(hipMemcpyAsync(el->wRep.get() + f * el->Nlocal, el->ogs->invDegree.get(),
                             el->Nlocal * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  NRS_HIP(hipEventCreate(&el->ev0));
  NRS_HIP(hipEventCreate(&el->ev1));
  NRS_HIP(hipEventCreate(&el->evPre0));
  NRS_HIP(hipEventCreate(&el->evPre1));

  if (o->preconditioner == NRS_PRECON_JACOBI) {
    el->invDiagA.alloc(Lt);
    elliptic_update_jacobi(el.get(), s);
  }
  if (o->preconditioner == NRS_PRECON_MULTIGRID) {
    el->mg = mg_create(el.get(), d, s);
  }
  NRS_HIP(hipStreamSynchronize(s));
  return el.release();
}

// ellipticUpdateJacobi (src/solvers/elliptic/ellipticUpdateJacobi.cpp:62-85): the
// element diagonal with the current coefficients, assembled, inverted; every
// field of a block solve has the same operator (loffset = 0)
void elliptic_update_jacobi(nrs_elliptic *el, hipStream_t s)
{
  if (el->opts.preconditioner != NRS_PRECON_JACOBI) return;
  launch_build_diag<double>(el->N, el->E, el->ggeo.get(), el->D.get(), el->l0ptr(), el->l1ptr(),
                            el->poisson, el->lfield(), el->invDiagA.get(), s);
  ogs_gather_scatter<double>(el->ogs, el->invDiagA.get(), 1, el->Nlocal, NRS_ADD, s);
  reciprocal<double>(el->Nlocal, el->invDiagA.get(), s);
  for (int f = 1; f < el->Nf; ++f)
    NRS_HIP(hipMemcpyAsync(el->invDiagA.get() + f * el->Nlocal, el->invDiagA.get(),
                           el->Nlocal * sizeof(double), hipMemcpyDeviceToDevice, s));
}

void elliptic_set_coefficients(nrs_elliptic *el, const double *l0, const double *l1, hipStream_t s)
{
  if (!l0 && !l1) {
    el->lf0 = el->lf1 = nullptr;
  } else {
    NRS_CHECK(l0, "set coefficients: lambda1 without lambda0");
    NRS_CHECK(el->poisson || l1, "set coefficients: a Helmholtz coefficient field needs lambda1");
    if (el->opts.preconditioner == NRS_PRECON_MULTIGRID)
      throw Error(NRS_ERR_UNSUPPORTED, "set coefficients: coefficient fields with the multigrid "
                                       "preconditioner (its levels take constant coefficients)");
    el->lf0 = l0;
    el->lf1 = el->poisson ? nullptr : l1;
  }
  elliptic_update_jacobi(el, s);
}

int operator_chunks(int N, int dtype, int64_t E) { return tuned(TUNE_OP, N, dtype, E, 1); }

template <typename T>
void chunked_operator(nrs_elliptic *el, int N, int E, const T *ggeo, const T *D, const T *l0,
                      const T *l1, bool poisson, const uint32_t *maskBits, const nrs_ogs *ogs,
                      const T *q, T *Aq, int nChunks, hipStream_t s, bool record, const uint8_t *fx,
                      bool lfield, const T *corners, const T *zw)
{
  NRS_CHECK(!(fx && lfield), "operator: the face fusion takes constant coefficients");
  NRS_CHECK(!(corners && (fx || lfield)), "operator: corner geometry takes constant coefficients, no fusion");
  auto ax = [&](int64_t a, int64_t n) {
    const int Np = (N + 1) * (N + 1) * (N + 1), NW = (Np + 31) / 32;
    const uint32_t *mb = maskBits ? maskBits + a * NW : nullptr;
    if (corners)
      launch_ax_trilin<T>(N, (int)n, corners + a * 24, zw, D, l0, l1, poisson, q + a * Np, Aq + a * Np, mb, s);
    else if (fx)
      launch_ax_fused<T>(N, (int)n, ggeo + a * 7 * Np, D, l0, l1, poisson, q + a * Np, Aq + a * Np, mb,
                         fx + a, s, el->fusionBlock);
    else if (lfield)
      launch_ax<T>(N, (int)n, nullptr, ggeo + a * 7 * Np, D, l0 + a * Np, poisson ? l1 : l1 + a * Np,
                   poisson, true, q + a * Np, Aq + a * Np, mb, s);
    else
      launch_ax<T>(N, (int)n, nullptr, ggeo + a * 7 * Np, D, l0, l1, poisson, false, q + a * Np,
                   Aq + a * Np, mb, s);
  };
  constexpr int F = nrs_elliptic::kOpChunks;
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
  record = record && el->profiling;
  if (record) {
    e0 = el->next_event(); e1 = el->next_event(); e2 = el->next_event();
    NRS_HIP(hipEventRecord(e0, s));
  }
  const int64_t L = (int64_t)E * (N + 1) * (N + 1) * (N + 1);
  if (nChunks <= 1 || ogs->nChunks != F || !el->aux) {
    ax(0, E);
    if (record) NRS_HIP(hipEventRecord(e1, s));
    ogs_gather_scatter<T>(ogs, Aq, 1, L, NRS_ADD, s);
  } else {
    NRS_CHECK(F % nChunks == 0, "operator: chunk count must divide 16");
    const int Np = (N + 1) * (N + 1) * (N + 1);
    const int64_t epc = ogs->nodesPerChunk / Np;
    NRS_CHECK(!fx || epc % el->fusionBlock == 0, "operator: fused chunks must hold whole blocks of elements");
    const int per = F / nChunks;
    for (int c = 0; c < nChunks; ++c) {
      const int64_t a = std::min<int64_t>(E, (int64_t)c * per * epc);
      const int64_t b = std::min<int64_t>(E, (int64_t)(c + 1) * per * epc);
      if (b > a) ax(a, b - a);
      NRS_HIP(hipEventRecord(el->chunkEv[c], s));
      NRS_HIP(hipStreamWaitEvent(el->aux, el->chunkEv[c], 0));
      ogs_gather_scatter_chunks<T>(ogs, Aq, c * per, (c + 1) * per, el->aux);
    }
    if (record) NRS_HIP(hipEventRecord(e1, s));
    NRS_HIP(hipEventRecord(el->auxDone, el->aux));
    NRS_HIP(hipStreamWaitEvent(s, el->auxDone, 0));
  }
  if (record) {
    NRS_HIP(hipEventRecord(e2, s));
    el->spans.push_back({NRS_PHASE_AX, e0, e1});
    el->spans.push_back({NRS_PHASE_GS, e1, e2});
  }
}
template void chunked_operator<double>(nrs_elliptic *, int, int, const double *, const double *,
                                       const double *, const double *, bool, const uint32_t *,
                                       const nrs_ogs *, const double *, double *, int, hipStream_t, bool,
                                       const uint8_t *, bool, const double *, const double *);
template void chunked_operator<float>(nrs_elliptic *, int, int, const float *, const float *,
                                      const float *, const float *, bool, const uint32_t *,
                                      const nrs_ogs *, const float *, float *, int, hipStream_t, bool,
                                      const uint8_t *, bool, const float *, const float *);

template <typename T>
int autotune_operator(nrs_elliptic *el, int N, int E, const T *ggeo, const T *D, const T *l0,
                      const T *l1, bool poisson, const uint32_t *maskBits, const nrs_ogs *ogs,
                      hipStream_t s, const uint8_t *fx)
{
  const int64_t L = (int64_t)E * (N + 1) * (N + 1) * (N + 1);
  DevBuf<T> q(L), ref(L), out(L);
  fill_pattern<T>(L, q.get(), s);
  auto run = [&](int k, T *o) {
    chunked_operator<T>(el, N, E, ggeo, D, l0, l1, poisson, maskBits, ogs, q.get(), o, k, s, false, fx);
  };
  run(1, ref.get());
  std::vector<T> hr(L), ho(L);
  NRS_HIP(hipMemcpyAsync(hr.data(), ref.get(), L * sizeof(T), hipMemcpyDeviceToHost, s));
  NRS_HIP(hipStreamSynchronize(s));
  int vals[8];
  const int nv = tune_variants(TUNE_OP, vals, 8);
  int best = 1;
  double bestMs = 1e300;
  for (int v = 0; v < nv; ++v) {
    run(vals[v], out.get());
    NRS_HIP(hipMemcpyAsync(ho.data(), out.get(), L * sizeof(T), hipMemcpyDeviceToHost, s));
    NRS_HIP(hipStreamSynchronize(s));
    // the same arithmetic in every form: bit-identical results required
    if (std::memcmp(ho.data(), hr.data(), L * sizeof(T)) != 0) continue;
    const double ms = time_launches([&] { run(vals[v], out.get()); }, s);
    if (ms < bestMs) { bestMs = ms; best = vals[v]; }
  }
  set_tuned(TUNE_OP, N, dtype_of<T>::value, E, best);
  return best;
}
template int autotune_operator<double>(nrs_elliptic *, int, int, const double *, const double *,
                                       const double *, const double *, bool, const uint32_t *,
                                       const nrs_ogs *, hipStream_t, const uint8_t *);
template int autotune_operator<float>(nrs_elliptic *, int, int, const float *, const float *,
                                      const float *, const float *, bool, const uint32_t *,
                                      const nrs_ogs *, hipStream_t, const uint8_t *);

// element part of the operator, all fields (no mask, no gather-scatter); three
// fields go through the block kernel, which reads the geometric factors once
static void element_operator(nrs_elliptic *el, const double *q, double *Aq, hipStream_t s)
{
  if (el->Nf > 1)
    launch_block_ax(el->N, el->E, el->Nlocal, 0, el->ggeo.get(), el->D.get(), el->l0ptr(),
                    el->l1ptr(), q, Aq, s, el->lfield());
  else
    launch_ax<double>(el->N, el->E, nullptr, el->ggeo.get(), el->D.get(), el->l0ptr(),
                      el->l1ptr(), el->poisson, el->lfield(), q, Aq, nullptr, s);
}

static void mask_fields(nrs_elliptic *el, double *v, hipStream_t s)
{
  for (int f = 0; f < el->Nf; ++f) launch_mask<double>(el->Nmasked, el->maskIds.get(), v + f * el->Nlocal, s);
}

void elliptic_operator(nrs_elliptic *el, const double *q, double *Aq, bool masked, hipStream_t s)
{
  if (el->Nf > 1) {
    // three fields: the block apply (ellipticBlockPartialAxCoeffHex3D), the
    // mask of each field, one gather-scatter of the three vectors
    // (oogs::startFinish with k = 3)
    launch_block_ax(el->N, el->E, el->Nlocal, 0, el->ggeo.get(), el->D.get(), el->l0ptr(),
                    el->l1ptr(), q, Aq, s, el->lfield());
    if (masked) mask_fields(el, Aq, s);
    ogs_gather_scatter<double>(el->ogs, Aq, el->Nf, el->Nlocal, NRS_ADD, s);
    return;
  }
  // coefficient fields take the unfused kernel and the full masked handle
  const bool fused = el->ogsOp && !el->lfield();
  chunked_operator<double>(el, el->N, el->E, el->ggeo.get(), el->D.get(), el->l0ptr(),
                           el->l1ptr(), el->poisson, masked ? el->maskBits.get() : nullptr,
                           fused ? el->ogsOp : el->ogs, q, Aq, operator_chunks(el->N, NRS_F64, el->E),
                           s, true, fused ? el->fx.get() : nullptr, el->lfield());
}

void zero_mean(nrs_elliptic *el, double *q, hipStream_t s)
{
  remove_mean(*el->red, el->Nlocal, nullptr, (double)el->Nlocal, q, el->scal.get(), S_TMP1, s);
}

void elliptic_preconditioner(nrs_elliptic *el, const double *r, double *z, hipStream_t s)
{
  switch (el->opts.preconditioner) {
    case NRS_PRECON_JACOBI: axmyz<double>(el->Nf * el->Nlocal, el->invDiagA.get(), r, z, s); break;
    case NRS_PRECON_MULTIGRID:   // one V-cycle per field (the hierarchy is single-field)
      for (int f = 0; f < el->Nf; ++f) el->mg->apply(r + f * el->Nlocal, z + f * el->Nlocal, s);
      break;
    case NRS_PRECON_USER:        // ellipticPreconditioner.cpp:72-74
      NRS_CHECK(el->userPrecon, "PRECONDITIONER = USER without a user preconditioner");
      el->userPrecon(el->userCtx, r, z, (void *)s);
      break;
    default:
      NRS_HIP(hipMemcpyAsync(z, r, el->Nf * el->Nlocal * sizeof(double), hipMemcpyDeviceToDevice, s));
      break;
  }
  if (el->allNeumann) zero_mean(el, z, s);   // ellipticPreconditioner.cpp:82-83
}

// Restarted (flexible) PGMRES (src/solvers/elliptic/linearSolver/PGMRES.cpp:111-340):
// classical Gram-Schmidt against all basis vectors in one pass, Givens
// rotations and the small triangular solve on the host; one readback of
// (y, |w|) per iteration.  r: the masked, assembled initial residual; x = 0;
// tol: absolute on the 1/volume-scaled norm.
static void gmres_solve(nrs_elliptic *el, double *r, double *x, double res0, double tol, int maxIter,
                        bool timed, int &iter, double &rdotr, double &precon_s, int &status,
                        hipStream_t s)
{
  const int64_t L = (int64_t)el->Nf * el->Nlocal;
  const double *wgt = el->Nf > 1 ? el->wRep.get() : el->ogs->invDegree.get();
  const auto &o = el->opts;
  const bool flexible = o.solver == NRS_SOLVER_FGMRES;
  const int m = o.gmresRestart, ldH = m + 1;
  const double sqrtRNF = std::sqrt(1.0 / el->volume);
  double *V = el->gV.get(), *Z = el->gZ.get(), *b = el->gB.get(), *w = el->p.get(), *Ax = el->Ap.get();
  double *yd = el->gY.get(), *part = el->gPart.get();
  std::vector<double> H((size_t)ldH * ldH, 0.0), cs(m, 0.0), sn(m, 0.0), sv(m + 1, 0.0), y(m + 1, 0.0);
  NRS_HIP(hipMemcpyAsync(b, r, L * sizeof(double), hipMemcpyDeviceToDevice, s));
  double nr = res0 / sqrtRNF;
  double error = res0;
  iter = 0;
  rdotr = res0;
  if (!(nr > 0.0)) return;   // zero residual: x = 0 is the solution
  // timing: inside the Arnoldi loop by events read after that iteration's
  // readback (no extra synchronisation), elsewhere by a synchronisation
  bool pending = false;
  auto precon = [&](const double *in, double *out, bool deferred) {
    if (timed && deferred) {
      NRS_HIP(hipEventRecord(el->evPre0, s));
      elliptic_preconditioner(el, in, out, s);
      NRS_HIP(hipEventRecord(el->evPre1, s));
      pending = true;
      return;
    }
    auto p0 = std::chrono::steady_clock::now();
    elliptic_preconditioner(el, in, out, s);
    if (timed) {
      NRS_HIP(hipStreamSynchronize(s));
      precon_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - p0).count();
    }
  };
  auto collect = [&]() {
    if (!pending) return;
    float ms = 0.0f;
    NRS_HIP(hipEventElapsedTime(&ms, el->evPre0, el->evPre1));
    precon_s += 1e-3 * ms;
    pending = false;
  };
  // gmresUpdate (PGMRES.cpp:59-104): back substitution, then x += Z y (flexible)
  // or x += M^-1 (V y)
  auto update = [&](int k) {
    for (int a = k - 1; a >= 0; --a) {
      y[a] = sv[a];
      for (int c = a + 1; c < k; ++c) y[a] -= H[a + c * ldH] * y[c];
      y[a] /= H[a + a * ldH];
    }
    NRS_HIP(hipMemcpyAsync(yd, y.data(), k * sizeof(double), hipMemcpyHostToDevice, s));
    if (flexible) {
      gmres_update(L, k, L, yd, Z, x, s);
    } else {
      double *zt = el->zv.get();
      fill<double>(L, 0.0, zt, s);
      gmres_update(L, k, L, yd, V, zt, s);
      precon(zt, w, false);
      axpby<double>(L, 1.0, w, 1.0, x, s);
    }
  };
  std::vector<double> hy(kGmresMaxVectors + 2);
  while (iter < maxIter) {
    sv.assign(m + 1, 0.0);
    sv[0] = nr;
    axpby<double>(L, 1.0 / nr, r, 0.0, V, s);                       // V_0 = r / |r|
    int used = m;
    bool done = false;
    for (int i = 0; i < m; ++i) {
      double *Mv = flexible ? Z + (int64_t)i * L : Z;
      precon(V + (int64_t)i * L, Mv, true);                         // z = M^-1 V_i
      elliptic_operator(el, Mv, w, true, s);                        // w = A z
      gmres_multi_dot(L, i + 1, L, wgt, V, w, part, yd, s);        // y_j = <V_j, w>
      gmres_gram_schmidt(L, i + 1, L, wgt, yd, V, w, part, yd + kGmresMaxVectors, s);
      NRS_HIP(hipMemcpyAsync(hy.data(), yd, (kGmresMaxVectors + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
      NRS_HIP(hipStreamSynchronize(s));
      collect();
      const double nw = std::sqrt(hy[kGmresMaxVectors]);
      H[i + 1 + i * ldH] = nw;
      if (i < m - 1) axpby<double>(L, 1.0 / nw, w, 0.0, V + (int64_t)(i + 1) * L, s);
      for (int k = 0; k <= i; ++k) H[k + i * ldH] = hy[k];
      for (int k = 0; k < i; ++k) {                                 // previous rotations
        const double h1 = H[k + i * ldH], h2 = H[k + 1 + i * ldH];
        H[k + i * ldH] = cs[k] * h1 + sn[k] * h2;
        H[k + 1 + i * ldH] = -sn[k] * h1 + cs[k] * h2;
      }
      const double h1 = H[i + i * ldH], h2 = H[i + 1 + i * ldH];
      const double hr = std::sqrt(h1 * h1 + h2 * h2);
      cs[i] = h1 / hr;
      sn[i] = h2 / hr;
      H[i + i * ldH] = cs[i] * h1 + sn[i] * h2;
      H[i + 1 + i * ldH] = 0.0;
      sv[i + 1] = -sn[i] * sv[i];
      sv[i] = cs[i] * sv[i];
      ++iter;
      error = std::fabs(sv[i + 1]) * sqrtRNF;
      rdotr = error;
      if (std::isnan(error)) { status = NRS_ERR_NAN; return; }
      if (error < tol || iter == maxIter) {
        update(i + 1);
        used = i + 1;
        done = true;
        break;
      }
    }
    (void)used;
    if (done) break;
    update(m);
    // restart: r = b - A x (fusedResidualAndNorm)
    elliptic_operator(el, x, Ax, true, s);
    gmres_residual(L, wgt, b, Ax, r, part, yd + kGmresMaxVectors, s);
    NRS_HIP(hipMemcpyAsync(hy.data() + kGmresMaxVectors, yd + kGmresMaxVectors, sizeof(double),
                           hipMemcpyDeviceToHost, s));
    NRS_HIP(hipStreamSynchronize(s));
    nr = std::sqrt(hy[kGmresMaxVectors]);
    error = nr * sqrtRNF;
    rdotr = error;
    if (std::isnan(error)) { status = NRS_ERR_NAN; return; }
    if (error <= tol) return;
  }
}

// pcg (linearSolver/PCG.cpp:85-203) or pgmres (PGMRES.cpp:111-340): the
// iteration on the masked, assembled residual r, x += the increment.
int elliptic_krylov(nrs_elliptic *el, double *r, double *x, double res0, double tol, int maxIter,
                    bool timed, int &iter, double &rdotr, double &precon_s, hipStream_t s)
{
  const int64_t L = (int64_t)el->Nf * el->Nlocal;
  const double *w = el->Nf > 1 ? el->wRep.get() : el->ogs->invDegree.get();
  double *scal = el->scal.get();
  Reducer &R = *el->red;
  const auto &o = el->opts;
  const bool flexible = o.solver == NRS_SOLVER_FCG;
  const bool none = o.preconditioner == NRS_PRECON_NONE;
  // Jacobi fused into the dot kernel, unless the mean of z must be removed first
  const bool jacobi = o.preconditioner == NRS_PRECON_JACOBI && !el->allNeumann;
  const double resNormFactor = 1.0 / el->volume;
  NRS_CHECK(maxIter >= 1, "maxIter must be >= 1");
  NRS_CHECK(o.preconditioner != NRS_PRECON_USER || el->userPrecon,
            "PRECONDITIONER = USER without a user preconditioner");
  iter = 0;
  rdotr = res0;
  precon_s = 0.0;
  int status = NRS_OK;
  if (o.solver == NRS_SOLVER_GMRES || o.solver == NRS_SOLVER_FGMRES) {
    gmres_solve(el, r, x, res0, tol, maxIter, timed, iter, rdotr, precon_s, status, s);
    return status;
  }

  double *p = el->p.get(), *Ap = el->Ap.get();
  double *z = none ? r : el->zv.get();
  fill<double>(L, 0.0, p, s);
  pcg_init(scal, res0, tol, maxIter, s);
  // Without a preconditioner or with Jacobi an iteration is cheap, and its
  // convergence check (a readback) costs as much as its kernels: enqueue four
  // between readbacks.  Every update kernel is a no-op once the device-side
  // flag says the iteration has stopped (converged or maxIter), so x, r and
  // the count are those of checking after each iteration.
  const int perCheck = (none || o.preconditioner == NRS_PRECON_JACOBI) ? 4 : 1;
  for (;;) {
    for (int b = 0; b < perCheck; ++b) {
      if (!none && !jacobi) {   // timed only with perCheck = 1: the events are read after the readback
        if (timed && perCheck == 1) NRS_HIP(hipEventRecord(el->evPre0, s));
        elliptic_preconditioner(el, r, z, s);
        if (timed && perCheck == 1) NRS_HIP(hipEventRecord(el->evPre1, s));
      }
      pcg_precon_dots(R, L, w, el->invDiagA.get(), r, z, Ap, jacobi, flexible, none, scal, s);
      pcg_update_p(L, z, p, scal, s);
      elliptic_operator(el, p, Ap, true, s);
      pcg_alpha(R, L, w, p, Ap, scal, s);
      pcg_update_rx(R, L, w, Ap, p, r, x, resNormFactor, scal, s);
    }
    NRS_HIP(hipMemcpyAsync(el->scalHost, scal, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
    NRS_HIP(hipStreamSynchronize(s));
    if (timed && !none && !jacobi && perCheck == 1) {
      float ms = 0.0f;
      NRS_HIP(hipEventElapsedTime(&ms, el->evPre0, el->evPre1));
      precon_s += 1e-3 * ms;
    }
    rdotr = el->scalHost[S_RDOTR];
    iter = (int)el->scalHost[S_ITER];
    if (std::isnan(rdotr)) return NRS_ERR_NAN;
    if (el->scalHost[S_ACTIVE] == 0.0) return NRS_OK;   // rdotr <= tol or iter == maxIter (PCG.cpp:200)
  }
}

// ellipticSolve (src/solvers/elliptic/ellipticSolve.cpp:32-190) + pcg (PCG.cpp:85-203).
int elliptic_solve(nrs_elliptic *el, double *rhs, double *x, nrs_solve_stats *st, hipStream_t s)
{
  const int64_t L = (int64_t)el->Nf * el->Nlocal;   // all fields
  const double *w = el->Nf > 1 ? el->wRep.get() : el->ogs->invDegree.get();
  double *r = rhs;
  double *scal = el->scal.get();
  Reducer &R = *el->red;
  const auto &o = el->opts;
  const double resNormFactor = 1.0 / el->volume;
  NRS_CHECK(o.preconditioner != NRS_PRECON_USER || el->userPrecon,
            "PRECONDITIONER = USER without a user preconditioner");
  NRS_CHECK(o.preconditioner != NRS_PRECON_MULTIGRID || o.coarseSolver != NRS_COARSE_USER || el->userCoarse,
            "multigrid: coarse solver USER without a coarse solve");
  auto t0 = std::chrono::steady_clock::now();
  NRS_HIP(hipEventRecord(el->ev0, s));
  // ELLIPTIC PRECO COEFF FIELD: the Jacobi diagonal from the fields' current values
  // (ellipticSolve.cpp:79-88)
  if (el->lfield() && o.precoCoeffField) elliptic_update_jacobi(el, s);

  // r = rhs - A x0 (element apply only), mask, gs   (ellipticSolve.cpp:91-103)
  element_operator(el, x, el->Ap.get(), s);
  axpby<double>(L, -1.0, el->Ap.get(), 1.0, r, s);
  if (el->allNeumann) zero_mean(el, r, s);
  mask_fields(el, r, s);
  ogs_gather_scatter<double>(el->ogs, r, el->Nf, el->Nlocal, NRS_ADD, s);
  // x0 = x, x = 0   (:105-106)
  NRS_HIP(hipMemcpyAsync(el->x0.get(), x, L * sizeof(double), hipMemcpyDeviceToDevice, s));
  fill<double>(L, 0.0, x, s);
  // projection of r on the previous solutions (:107-128); res00 before it
  const bool project = el->opts.initialGuess != NRS_GUESS_PREVIOUS;
  bool projected = false;
  double res00 = 0.0;
  if (project) {
    wnorm(R, L, w, r, resNormFactor, scal, S_TMP0, s);
    projected = projection_pre(el, r, s);
  }
  // res0 = ||r||_w / sqrt(volume)   (:130-139)
  wnorm(R, L, w, r, resNormFactor, scal, S_RDOTR, s);
  NRS_HIP(hipMemcpyAsync(el->scalHost, scal, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
  NRS_HIP(hipStreamSynchronize(s));
  const double res0 = el->scalHost[S_RDOTR];
  res00 = project ? el->scalHost[S_TMP0] : res0;
  if (std::isnan(res0) || std::isnan(res00)) return NRS_ERR_NAN;
  double tol = o.tol;
  if (o.relative) tol *= res0;
  // after the iteration: x += xbar and the basis update (:171-176), then x += x0
  auto finish = [&](int status) {
    if (project && status == NRS_OK) projection_post(el, x, s);
    axpby<double>(L, 1.0, el->x0.get(), 1.0, x, s);
    if (el->allNeumann) zero_mean(el, x, s);   // ellipticSolve.cpp:188-189
  };
  auto fill_stats = [&](int iter, double rdotr, double precon_s) {
    if (!st) return;
    st->iterations = iter;
    st->converged = (rdotr <= tol) ? 1 : 0;
    st->res0 = res0;
    st->res = rdotr;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    st->precon_seconds = precon_s;
    st->res00 = res00;
    st->projVectors = projected ? el->pjPrev : 0;
  };

  int iter = 0;
  double rdotr = res0, precon_s = 0.0;
  const int status = elliptic_krylov(el, r, x, res0, tol, o.maxIter, st != nullptr, iter, rdotr,
                                     precon_s, s);

  // x += xbar, basis update, x += x0   (:171-186)
  finish(status);
  NRS_HIP(hipEventRecord(el->ev1, s));
  NRS_HIP(hipStreamSynchronize(s));
  fill_stats(iter, rdotr, precon_s);
  return status;
}

// ---- the same solve as enqueued work only (no host synchronisation), for
// callers that capture iterations into a graph.  begin: r = gs(mask(rhs - A x)),
// x0 = x, x = 0, res0 and tolerance on the device; iterate: one (flexible) PCG
// iteration, a no-op once the solve has converged or reached maxIter (the
// update kernels test the device flag; the preconditioner and operator still
// run); end: x += x0.  The preconditioner must not synchronise (multigrid with a
// smoothed coarse level, Jacobi, or none).
void elliptic_solve_begin(nrs_elliptic *el, double *rhs, double *x, hipStream_t s)
{
  NRS_CHECK(el->Nf == 1, "device-driven solve: one field");
  NRS_CHECK(el->opts.initialGuess == NRS_GUESS_PREVIOUS,
            "device-driven solve: the projection initial guess needs host readbacks");
  const int64_t L = el->Nlocal;
  NRS_CHECK(!(el->mg && !el->mg->device_only()),
            "device-driven solve: the multigrid coarse level must be smoothed, not solved");
  if (el->lfield() && el->opts.precoCoeffField) elliptic_update_jacobi(el, s);
  launch_ax<double>(el->N, el->E, nullptr, el->ggeo.get(), el->D.get(), el->l0ptr(),
                    el->l1ptr(), el->poisson, el->lfield(), x, el->Ap.get(), nullptr, s);
  axpby<double>(L, -1.0, el->Ap.get(), 1.0, rhs, s);
  if (el->allNeumann) zero_mean(el, rhs, s);
  launch_mask<double>(el->Nmasked, el->maskIds.get(), rhs, s);
  ogs_gather_scatter<double>(el->ogs, rhs, 1, L, NRS_ADD, s);
  axpby<double>(L, 1.0, x, 0.0, el->x0.get(), s);
  fill<double>(L, 0.0, x, s);
  wnorm(*el->red, L, el->ogs->invDegree.get(), rhs, 1.0 / el->volume, el->scal.get(), S_RDOTR, s);
  pcg_init_device(el->scal.get(), el->opts.tol, el->opts.relative != 0, el->opts.maxIter, s);
  fill<double>(L, 0.0, el->p.get(), s);
}

void elliptic_iterate(nrs_elliptic *el, double *r, double *x, hipStream_t s)
{
  const int64_t L = el->Nlocal;
  const double *w = el->ogs->invDegree.get();
  const auto &o = el->opts;
  const bool none = o.preconditioner == NRS_PRECON_NONE;
  const bool jacobi = o.preconditioner == NRS_PRECON_JACOBI && !el->allNeumann;
  double *z = none ? r : el->zv.get();
  if (!none && !jacobi) elliptic_preconditioner(el, r, z, s);
  pcg_precon_dots(*el->red, L, w, el->invDiagA.get(), r, z, el->Ap.get(), jacobi,
                  o.solver == NRS_SOLVER_FCG, none, el->scal.get(), s);
  pcg_update_p(L, z, el->p.get(), el->scal.get(), s);
  elliptic_operator(el, el->p.get(), el->Ap.get(), true, s);
  pcg_alpha(*el->red, L, w, el->p.get(), el->Ap.get(), el->scal.get(), s);
  pcg_update_rx(*el->red, L, w, el->Ap.get(), el->p.get(), r, x, 1.0 / el->volume, el->scal.get(), s);
}

void elliptic_solve_end(nrs_elliptic *el, double *x, hipStream_t s)
{
  axpby<double>(el->Nlocal, 1.0, el->x0.get(), 1.0, x, s);
  if (el->allNeumann) zero_mean(el, x, s);
}

}  // namespace nrs
