// elliptic.hpp -- solver state (internal layout of nrs_elliptic).
#pragma once
#include <memory>

#include "common.hpp"
#include "linalg.hpp"
#include "ogs.hpp"

namespace nrs {
struct MGSolver;
}

// Mirrors elliptic_t (src/solvers/elliptic/elliptic.h:73-181) for one field.
struct nrs_elliptic {
  int N = 0, Nq = 0, Np = 0, E = 0;
  int Nf = 1;                        // fields (3: block solve, vectors [3][Nlocal])
  int64_t Nlocal = 0;                // per field
  bool poisson = true;
  // a Poisson problem without Dirichlet faces (elliptic_t::allNeumann,
  // ellipticSetup.cpp:187-206): the mean is removed from the residual, the
  // preconditioned residual and the solution (ellipticZeroMean)
  bool allNeumann = false;
  double lambda0 = 1.0, lambda1 = 0.0;
  // ELLIPTIC COEFF FIELD: per-node coefficients [Nlocal] owned by the caller
  // (elliptic_t::o_lambda0 / o_lambda1); null: the constants above
  const double *lf0 = nullptr, *lf1 = nullptr;
  bool lfield() const { return lf0 != nullptr; }
  const double *l0ptr() const { return lf0 ? lf0 : lam0.get(); }
  const double *l1ptr() const { return lf0 ? (lf1 ? lf1 : lam1.get()) : lam1.get(); }
  double volume = 0.0;
  nrs_elliptic_opts opts{};

  std::vector<double> z, w, Dh;      // GLL nodes/weights, D (host)
  nrs::DevBuf<double> ggeo, D, lam0, lam1;
  nrs::DevBuf<float> ggeo32, D32, lam032, lam132;   // FP32 copies (multigrid levels)

  nrs_ogs *ogsFull = nullptr;        // unmasked (mesh->ogs)
  nrs_ogs *ogs = nullptr;            // masked   (elliptic->ogs)
  // Operator face fusion (N = 7): fx[e] = 1 when elements e, e+1 (same block
  // of fusionBlock elements) share face i = N / i = 0 node for node; their
  // face-interior nodes are summed inside Ax and left out of ogsOp, the
  // operator's handle.
  nrs_ogs *ogsOp = nullptr;
  nrs::DevBuf<uint8_t> fx;
  int fusionBlock = 4;               // elements per block of the fused Ax kernel (2 or 4)
  int64_t fusedPairs = 0;
  int64_t Nmasked = 0;
  nrs::DevBuf<int> maskIds;
  nrs::DevBuf<uint32_t> maskBits;    // [E][ceil(Np/32)]
  nrs::DevBuf<double> invDiagA;      // Jacobi (assembled), fine level, [Nf][Nlocal]
  nrs::DevBuf<double> wRep;          // invDegree repeated per field (Nf > 1)
  nrs::DevBuf<double> gV, gZ, gB, gY, gPart;   // PGMRES basis, preconditioned basis, rhs, y, partials

  // Solution projection (SolutionProjection, ellipticSolutionProjection.h):
  // basis xx [projVectors][Nf*Nlocal], A xx ([projVectors] for CLASSIC, one
  // vector for ACONJ), xbar = the projected guess, alpha [projVectors].
  nrs::DevBuf<double> pjXX, pjBB, pjXbar, pjAlpha;
  int pjM = 0;                       // numVecsProjection
  int pjStep = 0;                    // solves so far (timestep)
  int pjPrev = 0;                    // basis size used by the last pre()
  std::vector<double> pjAlphaHost;

  // PRECONDITIONER = USER: elliptic_t::userPreconditioner (elliptic.h:178)
  nrs_user_precon_fn userPrecon = nullptr;
  void *userCtx = nullptr;
  // MULTIGRID coarse solver = USER: coarseLevel_t::solvePtr (MG/MGSolver.hpp:99)
  nrs_user_coarse_fn userCoarse = nullptr;
  void *userCoarseCtx = nullptr;

  // workspace (ellipticUpdateWorkspace, elliptic.h:240-262)
  nrs::DevBuf<double> p, zv, Ap, x0;
  nrs::DevBuf<double> scal;          // nrs::S_COUNT device scalars
  double *scalHost = nullptr;        // pinned mirror
  std::unique_ptr<nrs::Reducer> red;
  std::unique_ptr<nrs::MGSolver> mg;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evPre0 = nullptr, evPre1 = nullptr;   // around one preconditioner apply (timing)

  // Chunked operator: the masked handle's rows are grouped into kOpChunks
  // element chunks; Ax of chunk c+1 runs on the caller's stream while the
  // gather-scatter of the rows completed by chunk c runs on `aux`, reading
  // values that are still in cache.
  static constexpr int kOpChunks = 16;
  hipStream_t aux = nullptr;
  hipEvent_t chunkEv[kOpChunks] = {};
  hipEvent_t auxDone = nullptr;

  // Per-phase device timing (platform->timer analogue, src/core/timer.cpp:131-230):
  // when enabled, events bracket each Ax and gather-scatter launch of the
  // operator on the caller's stream; totals are read after a synchronisation.
  struct Span { int phase; hipEvent_t a, b; };
  bool profiling = false;
  std::vector<Span> spans;
  std::vector<hipEvent_t> eventPool;
  size_t eventNext = 0;
  hipEvent_t next_event();

  ~nrs_elliptic();
};

namespace nrs {
nrs_elliptic *elliptic_setup(const nrs_elliptic_desc *d, const nrs_elliptic_opts *o);
void elliptic_operator(nrs_elliptic *el, const double *q, double *Aq, bool masked, hipStream_t s);
void elliptic_preconditioner(nrs_elliptic *el, const double *r, double *z, hipStream_t s);
// ellipticUpdateJacobi: invDiagA from the current coefficients (JACOBI only)
void elliptic_update_jacobi(nrs_elliptic *el, hipStream_t s);
// repoint the coefficient fields (both null: the constants) and update the Jacobi diagonal
void elliptic_set_coefficients(nrs_elliptic *el, const double *l0, const double *l1, hipStream_t s);
int elliptic_solve(nrs_elliptic *el, double *rhs, double *x, nrs_solve_stats *st, hipStream_t s);
// pcg / pgmres on the masked, assembled residual r, x += increment; res0 the
// normalised initial norm, tol absolute; returns NRS_OK or NRS_ERR_NAN
int elliptic_krylov(nrs_elliptic *el, double *r, double *x, double res0, double tol, int maxIter,
                    bool timed, int &iter, double &rdotr, double &precon_s, hipStream_t s);
// Device-driven pieces of the same solve (no host synchronisation).
void elliptic_solve_begin(nrs_elliptic *el, double *rhs, double *x, hipStream_t s);
void elliptic_iterate(nrs_elliptic *el, double *r, double *x, hipStream_t s);
void elliptic_solve_end(nrs_elliptic *el, double *x, hipStream_t s);
// SolutionProjection::pre / post (ellipticSolutionProjection.cpp:269-288):
// pre removes the projection of the assembled residual r (and returns whether
// it ran); post adds it back to the solution correction x and extends the basis.
bool projection_pre(nrs_elliptic *el, double *r, hipStream_t s);
void projection_post(nrs_elliptic *el, double *x, hipStream_t s);

nrs_batch *batch_setup(int n, const nrs_elliptic_desc *descs, const nrs_elliptic_opts *opts,
                       int itersPerLaunch);
int batch_solve(nrs_batch *b, double *const *rhs, double *const *x, nrs_solve_stats *st, hipStream_t s);
void batch_set_mode(nrs_batch *b, int mode);

// Aq = gs(mask(Ax(q))) with the Ax split into nChunks (1, 2, 4, 8, 16) element
// chunks overlapped with the gather-scatter of completed rows (ogs must have
// been chunked into kOpChunks chunks of ceil(E / kOpChunks) elements); nChunks
// = 1 or an unchunked handle runs Ax then gs.  Same results in every form.
// record: add the Ax / gs spans to the handle's device timers when enabled.
// fx: the face-fusion flags (launch_ax_fused; ogs then is the reduced handle) or null.
// lfield: l0 / l1 are per-node fields [E*Np] (fx must then be null).
template <typename T>
void chunked_operator(nrs_elliptic *el, int N, int E, const T *ggeo, const T *D, const T *l0,
                      const T *l1, bool poisson, const uint32_t *maskBits, const nrs_ogs *ogs,
                      const T *q, T *Aq, int nChunks, hipStream_t s, bool record = false,
                      const uint8_t *fx = nullptr, bool lfield = false, const T *corners = nullptr,
                      const T *zw = nullptr);
// the operator's chunk count for (N, type, E): tuned, else 1
int operator_chunks(int N, int dtype, int64_t E);
// time nChunks = 1, 2, 4, 8, 16 on this operator, record the fastest
template <typename T>
int autotune_operator(nrs_elliptic *el, int N, int E, const T *ggeo, const T *D, const T *l0,
                      const T *l1, bool poisson, const uint32_t *maskBits, const nrs_ogs *ogs,
                      hipStream_t s, const uint8_t *fx = nullptr);
// Face fusion of an order-7 operator: from the masked ids [E*512] find the
// pairs (e, e+1), e % nb != nb - 1, that share face i = N / i = 0 node for node
// and hold the only two copies of each of those nodes (invDegree = 1/2, from
// the handle of the same ids), return their flags [E] and the ids with those
// face-interior nodes zeroed (the operator's gather-scatter ids).  Empty flags
// when no pair qualifies.
std::vector<uint8_t> face_fusion(int N, int E, std::vector<long long> &ids, const std::vector<double> &invDegree,
                                 int nb = 4);

// ellipticZeroMean (src/solvers/elliptic/ellipticZeroMean.cpp:31-45): q -= the
// mean of its E-vector entries (sum over E*Np, divided by E*Np); enqueued work only.
void zero_mean(nrs_elliptic *el, double *q, hipStream_t s);

// Shared helpers for the multigrid levels.
std::vector<uint32_t> mask_bits(int Np, int E, const std::vector<int> &maskIds);
std::vector<int> dirichlet_mask(int N, int E, const int *EToB, nrs_ogs *ogsFull, hipStream_t s);
}  // namespace nrs
