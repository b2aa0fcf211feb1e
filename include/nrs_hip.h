/*
 * nrs_hip.h -- C ABI of libnekrs_hip, an MI355X (gfx950) implementation of
 * nekRS's pressure-Poisson hot path: the element stiffness apply (Ax), the
 * ogs gather-scatter, the p-multigrid preconditioner and PCG / flexible CG.
 *
 * Conventions (same as the reference):
 *   - E-vectors are laid out [field][e][k][j][i] (i fastest), Np = (N+1)^3.
 *   - ggeo is [e][7][Np] with ids G00,G01,G11,G12,G02,G22,GWJ = 0..6
 *     (src/mesh/mesh3D.h:96-102).
 *   - D is row-major, D[i*(N+1)+j] = d(l_j)/dr at r_i (GLL nodes).
 *   - dlong = int32 (element and node indices), hlong = int64 (global ids)
 *     (3rd_party/gslib/ogs/include/ogstypes.h:21-45).
 *   - Every array argument documented "device" is a device pointer owned by
 *     the caller; the library never frees caller memory.  "host" arrays are
 *     read during the call only.
 *   - `stream` is a hipStream_t (NULL = the default stream).  All work of a
 *     call is enqueued on that stream; calls that return scalars synchronise it.
 *   - Every function returns NRS_OK (0) or a negative nrs_status; the
 *     reference aborts instead (nrsCheck -> MPI_Abort).  nrs_last_error()
 *     gives the message of the last failure on the calling thread.
 *   - A handle is used by one host thread at a time.
 */
#ifndef NRS_HIP_H
#define NRS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nrs_status {
  NRS_OK = 0,
  NRS_ERR_INVALID = -1,      /* bad argument or unsupported configuration      */
  NRS_ERR_HIP = -2,          /* HIP runtime error                              */
  NRS_ERR_NAN = -3,          /* NaN residual norm (PCG.cpp:193-195)            */
  NRS_ERR_UNSUPPORTED = -4,  /* valid for the reference, not built here        */
  NRS_ERR_NO_DEVICE = -5
};

enum nrs_dtype { NRS_F64 = 0, NRS_F32 = 1, NRS_I32 = 2 };
/* gather-scatter reductions: ogsAdd / ogsMul / ogsMin / ogsMax
 * (3rd_party/gslib/ogs/ogs.hpp:124-136) */
enum nrs_op { NRS_ADD = 0, NRS_MUL = 1, NRS_MIN = 2, NRS_MAX = 3 };

typedef struct nrs_ogs nrs_ogs;
typedef struct nrs_elliptic nrs_elliptic;

const char *nrs_last_error(void);
/* Select the HIP device for the calling thread and check it is gfx950. */
int nrs_init(int device);
const char *nrs_version(void);

/* ------------------------------------------------------------------------- */
/* Setup helpers                                                              */
/* ------------------------------------------------------------------------- */

/* Structured box [0,1]^3, nx*ny*nz hexes in lexicographic order, GLL nodes of
 * order N.  map: 0 affine, 1 sine bump x_i += 0.05 sin(pi x)sin(pi y)sin(pi z)
 * at every node, 2 Kershaw(eps) applied to element vertices with trilinear
 * elements (examples/kershaw/kershaw.usr:107-226).  All outputs host arrays:
 * x,y,z[E*Np], globalIds[E*Np] (1-based), EToB[E*6] (1 = Dirichlet, 0 interior,
 * faces ordered as FaceNodesHex3D, src/mesh/meshBasisHex3D.cpp:53-82).
 * Replaces the Nek5000 genbox/.re2 + usrdat path, which this build omits. */
int nrs_box_mesh(int nx, int ny, int nz, int N, int map, double eps,
                 double *x, double *y, double *z, long long *globalIds, int *EToB);

/* GLL nodes/weights and differentiation matrix (host arrays of N+1, (N+1)^2).
 * Replaces JacobiGLL / Dmatrix1D (src/mesh/meshBasis1D.cpp:99-119,249-281). */
int nrs_gll(int N, double *z, double *w, double *D);

/* Geometric factors from node coordinates, on the GPU.
 * Replaces geometricFactorsHex3D (kernels/mesh/geometricFactorsHex3D.okl:27-143).
 * x,y,z: device [E*Np]; ggeo: device [E*7*Np]. */
int nrs_geometric_factors(int N, int Nelements, const double *x, const double *y,
                          const double *z, double *ggeo, void *stream);

/* ------------------------------------------------------------------------- */
/* Kernel-level seams                                                         */
/* ------------------------------------------------------------------------- */

/* Element stiffness apply, Aq_e = D^T (G . (D q_e)) [+ GwJ lambda1 q_e].
 * Replaces ellipticPartialAxCoeffHex3D (ABI kernels/elliptic/
 * ellipticPartialAxCoeffHex3D.c:2-12, registered as
 * "poisson-ellipticPartialAxCoeffHex3D", src/core/registerEllipticKernels.cpp:183).
 *   dtype      NRS_F64 or NRS_F32 (every floating array in that type)
 *   elementList device int32 [Nelements] or NULL for 0..Nelements-1
 *   lambda0/1  device; lambda_field=0: lambda0[0] scalar (p_lambda=0),
 *              lambda_field=1: per-node fields indexed like q (p_lambda=1);
 *              lambda1 ignored when poisson=1.
 *   q, Aq      device E-vectors (Aq written only for listed elements). */
int nrs_ax(int N, int dtype, int Nelements, const int *elementList, const void *ggeo,
           const void *D, const void *lambda0, const void *lambda1, int poisson,
           int lambda_field, const void *q, void *Aq, void *stream);

/* Three-field Helmholtz apply with coefficient fields.
 * Replaces ellipticBlockPartialAxCoeffHex3D (kernels/elliptic/
 * ellipticBlockPartialAxCoeffHex3D.c:2-147). q,Aq: [3][offset]; lambda0/1:
 * [3][loffset]. FP64. */
int nrs_block_ax(int N, int Nelements, int64_t offset, int64_t loffset, const double *ggeo,
                 const double *D, const double *lambda0, const double *lambda1,
                 const double *q, double *Aq, void *stream);

/* Element diagonal of A (before assembly).
 * Replaces ellipticBlockBuildDiagonalHex3D (kernels/elliptic/
 * ellipticBlockBuildDiagonalHex3D.okl:27-117), single field. */
int nrs_build_diagonal(int N, int dtype, int Nelements, const void *ggeo, const void *D,
                       const void *lambda0, const void *lambda1, int poisson, int lambda_field,
                       void *diag, void *stream);

/* v[maskIds[n]] = 0.  Replaces mask (kernels/core/mask.okl:27-35) behind
 * ellipticApplyMask (src/solvers/elliptic/ellipticApplyMask.cpp:8-27). */
int nrs_mask(int dtype, int64_t Nmasked, const int *maskIds, void *v, void *stream);

/* Tensor-product transfers between orders Nf > Nc.  R: device [(Nc+1)*(Nf+1)],
 * R = transpose(InterpolationMatrix1D(Nc -> Nf))
 * (MG/ellipticMultiGridLevelSetup.cpp:238-258).
 * coarsen:    qc  = (R x R x R) qf        (ellipticPreconCoarsenHex3D.c:26-100)
 * prolongate: qf += (R^T x R^T x R^T) qc  (ellipticPreconProlongateHex3D.c:27-114) */
int nrs_coarsen(int Nf, int Nc, int dtype, int Nelements, const void *R, const void *qf,
                void *qc, void *stream);
int nrs_prolongate(int Nf, int Nc, int dtype, int Nelements, const void *R, const void *qc,
                   void *qf, void *stream);

/* ------------------------------------------------------------------------- */
/* Gather-scatter (ogs)                                                       */
/* ------------------------------------------------------------------------- */

/* Build a gather-scatter handle from global ids (host int64 [N], 0 = node
 * excluded).  Replaces ogsSetup (3rd_party/gslib/ogs/ogs.hpp:183,
 * src/ogsSetup.cpp:111-397) on one rank: rows are the groups of equal ids,
 * ordered by first local appearance, members ascending; singleton rows are
 * dropped from the device structure (the reference kernel skips them,
 * gatherScatterMany.okl:71). */
int nrs_ogs_setup(int64_t N, const long long *globalIds, nrs_ogs **ogs);
int nrs_ogs_free(nrs_ogs *ogs);

/* In-place QQ^T reduction on nvec vectors spaced `stride` elements apart.
 * Replaces oogs::startFinish(o_v, k, stride, type, op, h)
 * (3rd_party/gslib/ogs/ogs.hpp:282-285, src/oogs.cpp:682-837) and
 * ogsGatherScatterMany. v: device. */
int nrs_ogs_gather_scatter(nrs_ogs *ogs, void *v, int nvec, int64_t stride, int dtype, int op,
                           void *stream);

/* The split form of the same call, the overlap seam of oogs::start /
 * oogs::finish (3rd_party/gslib/ogs/ogs.hpp:282-283, src/oogs.cpp:682-821) as
 * ellipticOperator uses it (ellipticOperator.cpp:129-157): start, then the
 * caller's work on v, then finish with the same arguments.  The reference's
 * start packs and posts the rows shared with other ranks; on one device every
 * row is local, so start only checks and records its arguments and does not
 * read v (the caller may still write any entry of v), and finish performs the
 * whole gather-scatter on `stream` -- the same result, bit for bit, as
 * nrs_ogs_gather_scatter of v as it is at finish.  A finish without a start,
 * a second start before finish, or a finish whose arguments differ from its
 * start's is NRS_ERR_INVALID. */
int nrs_ogs_start(nrs_ogs *ogs, void *v, int nvec, int64_t stride, int dtype, int op, void *stream);
int nrs_ogs_finish(nrs_ogs *ogs, void *v, int nvec, int64_t stride, int dtype, int op, void *stream);

/* 1/multiplicity per node (1 at excluded nodes), FP64 device [N]
 * (ogs->o_invDegree, ogsSetup.cpp:364-392). */
int nrs_ogs_inv_degree(nrs_ogs *ogs, double *invDegree, void *stream);

/* Row statistics: rows with more than one member, their members, and the
 * number of singleton rows. */
int nrs_ogs_info(nrs_ogs *ogs, int64_t *Nrows, int64_t *Nmembers, int64_t *Nsingletons);

/* ------------------------------------------------------------------------- */
/* linAlg_t reductions and the PCG residual update                            */
/* ------------------------------------------------------------------------- */
/* The reductions of linAlg_t (src/linAlg/linAlg.hpp:156-215, the kernels under kernels/linAlg):
 *   NRS_RED_SUM     sum x                    (sum / sumMany)
 *   NRS_RED_MAX     max x                    (max)
 *   NRS_RED_MIN     min x                    (min)
 *   NRS_RED_NORM2   sqrt(sum x^2)            (norm2 / norm2Many)
 *   NRS_RED_NORM1   sum |x|                  (norm1 / norm1Many)
 *   NRS_RED_AMAX    max |x|                  (amax / amaxMany)
 *   NRS_RED_WNORM2  sqrt(sum_i w_i sum_f x^2)  (weightedNorm2Many)
 *   NRS_RED_WDOT    sum_i w_i sum_f x y        (weightedInnerProdMany)
 *   NRS_RED_DOT     sum x y                  (innerProd)
 * over N entries of Nfields fields spaced fieldOffset apart (x, y device
 * [Nfields][fieldOffset]; w device [N], read only by the weighted forms; y only
 * by the inner products).  The result is written to the host double *result;
 * the call synchronises `stream`.  N = 0 gives 0 (sums, norms, amax), -inf
 * (max) or +inf (min).  The order of the additions is fixed (one partial per
 * block of a grid that depends only on N, combined in a fixed tree), so a result
 * is reproducible bit for bit. */
enum nrs_reduction { NRS_RED_SUM = 0, NRS_RED_MAX = 1, NRS_RED_MIN = 2, NRS_RED_NORM2 = 3,
                     NRS_RED_NORM1 = 4, NRS_RED_AMAX = 5, NRS_RED_WNORM2 = 6, NRS_RED_WDOT = 7,
                     NRS_RED_DOT = 8 };
int nrs_linalg_reduce(int op, int64_t N, int Nfields, int64_t fieldOffset, const double *w,
                      const double *x, const double *y, double *result, void *stream);

/* nrs_linalg_reduce without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops that must not call hipMalloc, harnesses that forbid it).  Same
 * arguments, result and error behaviour as nrs_linalg_reduce, bit for bit; the
 * partial sums live in a fixed device array of 1056 doubles (8.25 KiB) that
 * every GPU the library runs on keeps until the process ends (its address is
 * looked up in nrs_init, or on the first call on a device).  Every N fits
 * (the partials grid is capped at 1024 blocks); a grid beyond the array is
 * refused with NRS_ERR_INVALID.  All calls on one GPU share the array: calls
 * that overlap in time on one GPU (other streams or threads) race.  Like
 * nrs_linalg_reduce it synchronises `stream`, so it cannot be captured. */
int nrs_linalg_reduce_noalloc(int op, int64_t N, int Nfields, int64_t fieldOffset, const double *w,
                              const double *x, const double *y, double *result, void *stream);

/* r -= alpha Ap on Nfields fields spaced `offset` apart and *rdotr (host) =
 * sum_i invDegree_i sum_f r[i + f offset]^2 of the updated r (no square root).
 * Replaces ellipticBlockUpdatePCG (kernels/elliptic/ellipticBlockUpdatePCG.c:27-50,
 * called from PCG.cpp:67-83).  invDegree device [N]; Ap, r device
 * [Nfields][offset].  Synchronises `stream`. */
int nrs_block_update_pcg(int64_t N, int Nfields, int64_t offset, const double *invDegree,
                         const double *Ap, double alpha, double *r, double *rdotr, void *stream);

/* ------------------------------------------------------------------------- */
/* Elliptic solver (elliptic_t + ellipticSolve)                               */
/* ------------------------------------------------------------------------- */

typedef struct {
  int N;                         /* polynomial order                            */
  int Nelements;
  const double *x, *y, *z;       /* host [E*Np] node coordinates                */
  const long long *globalIds;    /* host [E*Np], 1-based                        */
  const int *EToB;               /* host [E*6], 1 = DIRICHLET (elliptic.h:45)   */
  int poisson;                   /* 1: Poisson (lambda1 ignored)                */
  double lambda0, lambda1;       /* constant coefficients (p_lambda = 0)        */
  int Nfields;                   /* 0 or 1: one field; 3: block solve of three
                                  * fields [3][E*Np] with one operator (the
                                  * velocity solve, elliptic_t::Nfields = 3,
                                  * elliptic.h:77; Jacobi or no preconditioner) */
  /* ELLIPTIC COEFF FIELD = TRUE (ellipticOperator.cpp:44, elliptic_t::o_lambda0 /
   * o_lambda1, elliptic.h:168-169): per-node coefficients, device [E*Np] FP64,
   * owned by the caller and read at every operator apply, so the caller may
   * change their values between solves (nekRS points them into
   * nrs->o_ellipticCoeff, setup.cpp:784-785).  One field serves every field
   * of a block solve (elliptic_t::loffset = 0).  NULL: the constants above.
   * lambda1Field is required for a Helmholtz problem with lambda0Field and
   * ignored for Poisson.  Preconditioners NONE, JACOBI and USER (multigrid
   * levels take constant coefficients: NRS_ERR_UNSUPPORTED). */
  const double *lambda0Field;
  const double *lambda1Field;
} nrs_elliptic_desc;

/* SOLVER = PCG, PCG+FLEXIBLE, PGMRES, PGMRES+FLEXIBLE (PGMRES.cpp:111-340) */
enum nrs_solver { NRS_SOLVER_PCG = 0, NRS_SOLVER_FCG = 1, NRS_SOLVER_GMRES = 2, NRS_SOLVER_FGMRES = 3 };
/* PRECONDITIONER = NONE, JACOBI, MULTIGRID, USER (ellipticPreconditioner.cpp:33-84);
 * USER calls the function given to nrs_elliptic_set_user_preconditioner. */
enum nrs_precon { NRS_PRECON_NONE = 0, NRS_PRECON_JACOBI = 1, NRS_PRECON_MULTIGRID = 2,
                  NRS_PRECON_USER = 3 };
enum nrs_smoother {
  NRS_SMOOTHER_CHEBYSHEV_JACOBI = 0,  /* DAMPEDJACOBI,CHEBYSHEV (1st kind)       */
  NRS_SMOOTHER_OPT4_CHEBYSHEV_JACOBI = 1,
  NRS_SMOOTHER_CHEBYSHEV_ASM = 2,     /* CHEBYSHEV+ASM (FDM/Schwarz)             */
  NRS_SMOOTHER_OPT4_CHEBYSHEV_ASM = 3,
  NRS_SMOOTHER_ASM = 4,               /* ASM (no post-smoothing)                 */
  NRS_SMOOTHER_RAS = 5,               /* RAS (no post-smoothing)                 */
  NRS_SMOOTHER_CHEBYSHEV_RAS = 6,     /* CHEBYSHEV+RAS                           */
  NRS_SMOOTHER_OPT4_CHEBYSHEV_RAS = 7 /* FOURTHOPTCHEBYSHEV+RAS                  */
};
/* Coarsest level: FP64 Jacobi-PCG solve; smoothed with the level smoother
 * (MULTIGRID COARSE SOLVE = FALSE, ellipticMultiGridSetup.cpp:315-320); or
 * smoothed with Chebyshev-Jacobi even when the finer levels use Schwarz (no
 * counterpart in the reference: its Schwarz boxes on an order-1 level with
 * free box ends make a poor smoother there). */
enum nrs_coarse {
  NRS_COARSE_PCG_F64 = 0,        /* FP64 Jacobi-PCG to coarseTol                       */
  NRS_COARSE_SMOOTH = 1,
  NRS_COARSE_SMOOTH_JACOBI = 2,
  NRS_COARSE_USER = 4,           /* the caller's coarse solve, coarseLevel_t::solvePtr
                                  * (MG/MGSolver.hpp:99): the function given to
                                  * nrs_elliptic_set_user_coarse_solver; the V-cycle
                                  * is then not replayed as one graph             */
  NRS_COARSE_AMG = 3             /* FP64 PCG on the assembled order-1 matrix,
                                  * preconditioned by one smoothed-aggregation AMG
                                  * V-cycle (in place of BoomerAMG,
                                  * coarseLevel.cpp:47-222), to coarseTol; with
                                  * coarseTol <= 0 exactly coarseMaxIter (1..64)
                                  * iterations on the device, no host round trip,
                                  * so the whole V-cycle replays as one graph      */
};
enum nrs_mg_precision { NRS_MG_F32 = 0, NRS_MG_F64 = 1 };
/* INITIAL GUESS (parReader.cpp:1195-1246, ellipticSetup.cpp:306-321): the
 * caller's x (PREVIOUS), or the previous solutions' projection
 * (SolutionProjection, ellipticSolutionProjection.cpp): CLASSIC keeps A x_k for
 * every basis vector, ACONJ keeps an A-orthonormal basis and applies A once. */
enum nrs_initial_guess { NRS_GUESS_PREVIOUS = 0, NRS_GUESS_PROJECTION = 1, NRS_GUESS_PROJECTION_ACONJ = 2 };
enum nrs_fdm_precision { NRS_FDM_F32 = 0, NRS_FDM_BF16 = 1, NRS_FDM_F16 = 2 };

typedef struct {
  int solver;                    /* nrs_solver                                  */
  int preconditioner;            /* nrs_precon                                  */
  double tol;                    /* SOLVER TOLERANCE                            */
  int relative;                  /* LINEAR SOLVER STOPPING CRITERION=RELATIVE   */
  int maxIter;                   /* MAXIMUM ITERATIONS                          */
  /* multigrid */
  int nLevels;                   /* 0: default schedule (determineMGLevels)     */
  int levels[8];                 /* orders, finest first, e.g. {7,3,1}          */
  int smoother;                  /* nrs_smoother                                */
  int chebyDegree;               /* MULTIGRID CHEBYSHEV DEGREE                  */
  double chebyMinFactor;         /* default 0.1                                 */
  double chebyMaxFactor;         /* default 1.1                                 */
  int coarseSolver;              /* nrs_coarse                                  */
  double coarseTol;              /* relative tolerance of the coarse PCG        */
  int coarseMaxIter;
  int mgPrecision;               /* nrs_mg_precision                            */
  int fdmPrecision;              /* nrs_fdm_precision                           */
  unsigned long long seed;       /* Arnoldi start vector seed                   */
  int useGraphs;                 /* replay captured HIP graphs for the coarse PCG */
  int autotune;                  /* KERNEL AUTOTUNING: time launch variants at setup */
  int gmresRestart;              /* PGMRES RESTART (default 15, at most 32)     */
  int initialGuess;              /* nrs_initial_guess (default PREVIOUS)        */
  int projVectors;               /* RESIDUAL PROJECTION VECTORS (default 8, 1..32) */
  int projStart;                 /* RESIDUAL PROJECTION START: solve index from
                                  * which the projection runs (default 5)       */
  int operatorFusion;            /* 1 (default): at N = 7, sum the shared face
                                  * of consecutive elements that meet i = N to
                                  * i = 0 inside the Ax kernel and leave those
                                  * nodes out of the operator's gather-scatter
                                  * (same results bit for bit), in blocks of four
                                  * elements, or of two where that times faster;
                                  * with autotune on it is kept only where it
                                  * times faster than no fusion; 2: blocks of two;
                                  * 0: off */
  int preDegree[8];              /* per multigrid level (0: chebyDegree): Chebyshev
                                  * degree of the smoothing before restriction
                                  * (x = 0), DownLegChebyshevDegree               */
  int postDegree[8];             /* ... after prolongation, UpLegChebyshevDegree
                                  * (ellipticMultiGridLevel.cpp:146,214); the
                                  * coarsest level, when smoothed, keeps degree 3 */
  int precoCoeffField;           /* ELLIPTIC PRECO COEFF FIELD (parReader.cpp:999):
                                  * with coefficient fields and JACOBI, every solve
                                  * first rebuilds the Jacobi diagonal from the
                                  * fields' current values (ellipticSolve.cpp:79-88,
                                  * ellipticUpdateJacobi.cpp:91-115); default 1 */
  int mgGeometry;                /* 1 (default): when every element is trilinear
                                  * (its nodes the trilinear map of its corners),
                                  * the multigrid levels' operators form their
                                  * geometric factors from the corners in the
                                  * kernel instead of streaming stored ones (FP32
                                  * at every order, FP64 up to order 7); 0: the
                                  * stored factors, as the reference.  The outer
                                  * operator always uses the stored factors. */
} nrs_elliptic_opts;

typedef struct {
  int iterations;
  int converged;
  double res0;                   /* initial normalised residual norm            */
  double res;                    /* final normalised residual norm              */
  double seconds;                /* wall time of the solve                      */
  double precon_seconds;
  double res00;                  /* normalised residual before the projection
                                  * (= res0 without one)                       */
  int projVectors;               /* basis vectors the projection used (0: none) */
} nrs_solve_stats;

/* MULTIGRID SCHEDULE (the .par key pMGSchedule), e.g.
 * "p=7+degree=2, p=3+degree=2, p=1, p=3+degree=2, p=7+degree=2": the rules of
 * parseMultigridSchedule (MG/parseMultigridSchedule.cpp:7-131) and the levels
 * determineMGLevels takes from it (MG/determineMGLevels.cpp:22-40).  Sets
 * opts->nLevels and opts->levels (the distinct down-leg orders, finest first),
 * and preDegree / postDegree as the reference's level setup reads the schedule
 * (MG/ellipticMultiGridLevelSetup.cpp:160-167): the smoothing before
 * restriction takes the degree written on the schedule's up leg for that order,
 * the smoothing after prolongation the down leg's (the reference's assignment,
 * kept).  defaultDegree fills entries of orders above the coarsest that give
 * none; N = the mesh order, which the finest entry must equal.  A degree given
 * for the coarsest order is refused (the coarsest level, when smoothed, keeps
 * degree 3; the reference's level setup would take it as that level's degree
 * after prolongation and read the absent up-leg entry as 0 before restriction).
 * Errors (NRS_ERR_INVALID, the reasons in nrs_last_error): unknown token,
 * missing order or degree, different down- and up-leg orders, an order < 1,
 * finest order != N, more than 8 levels, a degree on the coarsest order.  As in the reference, a token is an
 * order when it contains 'p' (so "sweeps=2" reads as order 2).  Host only. */
int nrs_parse_mg_schedule(const char *schedule, int N, int defaultDegree, nrs_elliptic_opts *opts);

/* Fill opts with the reference's defaults for this path (PCG, Jacobi,
 * tol 1e-8 relative, maxIter 500, Chebyshev factors 0.1/1.1, degree 2). */
void nrs_elliptic_default_opts(nrs_elliptic_opts *opts);

/* Replaces ellipticSolveSetup (src/solvers/elliptic/ellipticSetup.cpp:116-327):
 * geometric factors, unmasked and masked ogs, Dirichlet mask
 * (ellipticOgs.cpp:4-134), invDegree, Jacobi diagonal, p-multigrid levels. */
int nrs_elliptic_setup(const nrs_elliptic_desc *desc, const nrs_elliptic_opts *opts,
                       nrs_elliptic **out);
int nrs_elliptic_free(nrs_elliptic *el);

/* Sizes: Nlocal = E*Np, Nmasked, volume = sum(GwJ). */
int nrs_elliptic_info(nrs_elliptic *el, int64_t *Nlocal, int64_t *Nmasked, double *volume);

/* Face pairs summed inside the Ax kernel by the operator face fusion (0 when
 * off or not applicable). */
int nrs_elliptic_fusion_info(nrs_elliptic *el, int64_t *fusedPairs);

/* Device buffers owned by the handle: 0 ggeo (f64), 1 D (f64), 2 invDegree
 * (f64), 3 maskIds (int32), 4 invDiagA (f64). */
int nrs_elliptic_buffer(nrs_elliptic *el, int which, void **ptr, int64_t *count);
/* Copy `count` entries of buffer `which` into the caller's device array dst. */
int nrs_elliptic_copy_buffer(nrs_elliptic *el, int which, void *dst, int64_t count, void *stream);

/* Point the handle's coefficients at new fields (the assignment of
 * elliptic->o_lambda0 / o_lambda1, setup.cpp:784-785): lambda0 / lambda1 device
 * [E*Np] FP64, kept by reference as in nrs_elliptic_desc; both NULL: back to
 * the descriptor's constants.  Rebuilds the Jacobi diagonal from the new values
 * (ellipticUpdateJacobi). */
int nrs_elliptic_set_coefficients(nrs_elliptic *el, const double *lambda0, const double *lambda1,
                                  void *stream);

/* ellipticUpdateJacobi(elliptic) (src/solvers/elliptic/ellipticUpdateJacobi.cpp:
 * 32-89,91-115): invDiagA = 1 / gs(diag(A)) from the coefficients' current
 * values, enqueued on `stream` (a no-op without the JACOBI preconditioner).
 * nrs_elliptic_solve calls it first when opts.precoCoeffField is set and the
 * handle has coefficient fields (ellipticSolve.cpp:79-88). */
int nrs_elliptic_update_jacobi(nrs_elliptic *el, void *stream);

/* Aq = gs(mask(Ax(q))) in FP64 on device E-vectors.
 * Replaces ellipticOperator(elliptic, o_q, o_Aq, "double", masked)
 * (src/solvers/elliptic/ellipticOperator.cpp:117-172). */
int nrs_elliptic_operator(nrs_elliptic *el, const double *q, double *Aq, int masked,
                          void *stream);

/* z = M^{-1} r.  Replaces ellipticPreconditioner
 * (src/solvers/elliptic/ellipticPreconditioner.cpp:33-84). */
int nrs_elliptic_preconditioner(nrs_elliptic *el, const double *r, double *z, void *stream);

/* The user preconditioner hook, elliptic_t::userPreconditioner (elliptic.h:178),
 * called for PRECONDITIONER = USER (ellipticPreconditioner.cpp:72-74) in place of
 * the built-in ones: fn(user, r, z, stream) must enqueue z = M^{-1} r on
 * `stream` (r and z are device vectors [Nfields][E*Np] of the handle; fn must not
 * write r).  The solve calls it from the calling host thread, once per iteration
 * (twice per PGMRES restart for the non-flexible form).  fn = NULL clears the hook;
 * a USER handle without one fails its solve with NRS_ERR_INVALID. */
typedef void (*nrs_user_precon_fn)(void *user, const double *r, double *z, void *stream);
int nrs_elliptic_set_user_preconditioner(nrs_elliptic *el, nrs_user_precon_fn fn, void *user);

/* The coarse-solve hook of NRS_COARSE_USER, coarseLevel_t::solvePtr
 * (MG/MGSolver.hpp:99, called by MGSolver_t::runVcycle, MGSolver.cpp:174-177):
 * fn(user, rhs, x, stream) must enqueue on `stream` the coarsest level's solve,
 * writing all of x; rhs and x are device E-vectors of the coarsest level
 * (nrs_elliptic_mg_info gives its length) in the multigrid precision.  Called
 * from the host thread of the preconditioner, once per V-cycle.  fn = NULL
 * clears it; a NRS_COARSE_USER handle without one fails its solve and its
 * preconditioner with NRS_ERR_INVALID before any work. */
typedef void (*nrs_user_coarse_fn)(void *user, const void *rhs, void *x, void *stream);
int nrs_elliptic_set_user_coarse_solver(nrs_elliptic *el, nrs_user_coarse_fn fn, void *user);

/* The Krylov iteration alone, the seam between ellipticSolve and its linear
 * solver: pcg(elliptic, o_r, o_x, tol, MAXIT, rdotr) (linearSolver/PCG.cpp:85-203,
 * elliptic.h:195-196) for opts.solver = PCG / FCG, pgmres (PGMRES.cpp:111-340,
 * elliptic.h:199-200) for GMRES / FGMRES.  r: the masked, assembled residual
 * (overwritten); x: the increment, updated in place (PGMRES's restarts assume it
 * starts at zero, as ellipticSolve passes it, ellipticSolve.cpp:105-106); tol:
 * absolute, on the normalised norm; maxIter >= 1 replaces opts.maxIter; *rdotr:
 * on entry the normalised initial norm ||r||_w / sqrt(volume), on exit the final
 * one; *iterations: the count pcg returns.  No projection, no mask of x0.
 * Returns NRS_OK (also when maxIter is reached) or NRS_ERR_NAN. */
int nrs_elliptic_krylov(nrs_elliptic *el, double *r, double *x, double tol, int maxIter,
                        double *rdotr, int *iterations, void *stream);

/* Solve A x = rhs.  rhs is the unassembled right-hand side (the solve applies
 * mask and gather-scatter to rhs - A x0 itself); x holds the initial guess on
 * entry and the solution on exit.  rhs is overwritten with the final residual.
 * Replaces ellipticSolve (src/solvers/elliptic/ellipticSolve.cpp:32-190) with
 * pcg (linearSolver/PCG.cpp:85-203) or PGMRES.  With opts.initialGuess =
 * PROJECTION or PROJECTION_ACONJ the residual's projection on the previous
 * solutions is removed before the iteration and added back after it, and the
 * new solution extends the basis (SolutionProjection::pre/post,
 * ellipticSolve.cpp:121-174); every call counts as one time step.  Returns
 * NRS_OK also when maxIter is hit (stats->converged = 0), NRS_ERR_NAN on a NaN
 * residual. */
int nrs_elliptic_solve(nrs_elliptic *el, double *rhs, double *x, nrs_solve_stats *stats,
                       void *stream);

/* p-multigrid level report (MGSolver_t::Report analogue,
 * MG/ellipticMultiGridSetup.cpp:326-342): for each level i < *nLevels (<= max)
 * its order, E-vector length, estimated lambda_max(S A) and Chebyshev degree;
 * coarseIterations = PCG iterations of the last coarse solve. */
int nrs_elliptic_mg_info(nrs_elliptic *el, int max, int *nLevels, int *orders, int64_t *Nlocal,
                         double *rho, int *degree, int *coarseIterations);

/* out = S in on multigrid level `level` (0 = finest, not the coarsest): the
 * level's smoother operator, 1/diag(A) (Jacobi) or the overlapping Schwarz
 * solve with fast diagonalisation (ASM / RAS), including its gather-scatter
 * and mask.  in/out are device E-vectors of that level in the multigrid
 * precision (float for NRS_MG_F32, double for NRS_MG_F64); out != in.
 * Replaces pMGLevel::smoother (MG/ellipticMultiGridLevel.cpp:104-111) and
 * pMGLevel::smoothSchwarz (MG/ellipticMultiGridSchwarz.cpp:1056-1156). */
int nrs_elliptic_mg_smoother(nrs_elliptic *el, int level, const void *in, void *out, void *stream);

/* One operation of multigrid level `level` (0 = finest), the per-level
 * interface multigridLevel::{Ax, residual, smooth, coarsen, prolongate}
 * (MG/MGSolver.hpp:59-69, pMGLevel in MG/ellipticMultiGridLevel.cpp).  Device
 * E-vectors of the level's order in the multigrid precision (float for
 * NRS_MG_F32, double for NRS_MG_F64):
 *   NRS_MG_AX          out = A in (the level operator: Ax, mask, gather-scatter)
 *   NRS_MG_RESIDUAL    out = in - A x
 *   NRS_MG_SMOOTH      Chebyshev / Schwarz smoothing of out (the iterate) for the
 *                      right-hand side in; xIsZero = 1: the smoothing before
 *                      restriction (out starts from 0), 0: after prolongation;
 *                      not on a coarsest level that is solved
 *   NRS_MG_COARSEN     `level` >= 1 is the coarse level: in (of level - 1) is
 *                      weighted by that level's 1/multiplicity in place, as the
 *                      reference does, then out = R in, gather-scatter, mask
 *   NRS_MG_PROLONGATE  `level` >= 1 is the coarse level: out (of level - 1) += P in
 * Enqueued on `stream`; vectors of different roles must not alias. */
enum nrs_mg_level_op { NRS_MG_AX = 0, NRS_MG_RESIDUAL = 1, NRS_MG_SMOOTH = 2, NRS_MG_COARSEN = 3,
                       NRS_MG_PROLONGATE = 4 };
int nrs_elliptic_mg_level(nrs_elliptic *el, int level, int op, void *in, const void *x, void *out,
                          int xIsZero, void *stream);

/* Host-only FDM setup pieces of the Schwarz smoother (no device work):
 * the 1D operator pair of one element direction on N+3 points, with
 * boundary flags lbc/rbc (0 interior, 1 Dirichlet, 2 Neumann) and the
 * left/middle/right element lengths: S[nl*nl] (row = point, column =
 * eigenvector, S^T B S = I, ghost/Dirichlet rows zeroed) and ascending
 * lam[nl].  Replaces compute_1d_matrices (MG/ellipticMultiGridSchwarz.cpp:505-556,
 * LAPACK dsygv). */
int nrs_fdm_1d(int N, int lbc, int rbc, double ll, double lm, double lr, double *S, double *lam);
/* Element lengths [9][E] (left x,y,z, middle x,y,z, right x,y,z) from the
 * order-N node coordinates and ids.  Replaces compute_element_lengths
 * (MG/ellipticMultiGridSchwarz.cpp:58-276). */
int nrs_fdm_element_lengths(int N, int Nelements, const double *x, const double *y, const double *z,
                            const long long *globalIds, double *lengths);

/* Global ids of the order-Nc nodes of a conforming hexahedral mesh derived from
 * its order-Nf ids (host arrays [E*(Nf+1)^3] -> [E*(Nc+1)^3]); replaces Nek5000's
 * numbering of the multigrid meshes (createMeshMG -> meshGlobalIds). */
int nrs_coarse_global_ids(int Nf, int Nelements, const long long *idsFine, int Nc, long long *idsCoarse);

/* ------------------------------------------------------------------------- */
/* Batched independent solves                                                 */
/* ------------------------------------------------------------------------- */
/* No counterpart in the reference (one elliptic_t per field and rank there).
 * n independent problems (any order, size, lambda, options), each set up as by
 * nrs_elliptic_setup; multigrid problems need a coarse level without host
 * round trips: NRS_COARSE_SMOOTH, NRS_COARSE_SMOOTH_JACOBI, or NRS_COARSE_AMG with
 * coarseTol <= 0 (the solve is device-driven).  Setup captures itersPerLaunch iterations of each problem
 * into a HIP graph per problem (NRS_BATCH_GRAPH_PER_PROBLEM, the default; see
 * nrs_batch_set_mode for the single-graph form). */
typedef struct nrs_batch nrs_batch;
int nrs_batch_setup(int n, const nrs_elliptic_desc *descs, const nrs_elliptic_opts *opts,
                    int itersPerLaunch, nrs_batch **out);
/* Solve all problems: rhs[p], x[p] device arrays as in nrs_elliptic_solve
 * (x: initial guess in, solution out; rhs: final residual out); stats[n].
 * Rounds of graph launches run until every problem has converged or reached
 * maxIter. */
int nrs_batch_solve(nrs_batch *b, double *const *rhs, double *const *x, nrs_solve_stats *stats,
                    void *stream);
/* Problem count, nodes of the graph(s) of the current mode, rounds of graph
 * launches of the last solve. */
int nrs_batch_info(nrs_batch *b, int *n, int64_t *graphNodes, int *lastLaunches);
int nrs_batch_free(nrs_batch *b);
/* The handle of problem p (owned by the batch: do not free it).  A stand-alone
 * nrs_elliptic_solve on it runs the same iteration one problem at a time. */
int nrs_batch_problem(nrs_batch *b, int p, nrs_elliptic **el);
/* How solves advance the problems: NRS_BATCH_ONE_GRAPH replays the single
 * graph of all problems (converged problems' nodes switched off);
 * NRS_BATCH_GRAPH_PER_PROBLEM captures one graph per problem (at the first
 * solve in this mode) and launches only the active problems' graphs, spread
 * over a pool of streams, per round.  Results are identical. */
enum nrs_batch_mode { NRS_BATCH_ONE_GRAPH = 0, NRS_BATCH_GRAPH_PER_PROBLEM = 1 };
int nrs_batch_set_mode(nrs_batch *b, int mode);

/* ------------------------------------------------------------------------- */
/* Kernel launch variants (KERNEL AUTOTUNING; the reference times p_knl       */
/* variants at setup: src/bench/axHelm/benchmarkAx.cpp:268-383,               */
/* src/bench/fdm/benchmarkFDM.cpp:180-210)                                    */
/* ------------------------------------------------------------------------- */
/* Variants: NRS_TUNE_AX = elements-per-block multiplier {1, 2, 4} of the
 * operator kernel, keyed by dtype (nrs_dtype); NRS_TUNE_FDM = elements per
 * wavefront {1, 2, 4, 8} of the matrix-core Schwarz kernel, keyed by the
 * nrs_fdm_precision; NRS_TUNE_OPERATOR = element chunks {1, 2, 4, 8, 16} of
 * the solver operator (Ax of chunk c+1 overlapped with the gather-scatter of
 * the rows chunk c completes), keyed by dtype; NRS_TUNE_TOMBO = load schedule
 * {1: at use, 2: one plane ahead} of the pressure right-hand side's curl
 * kernels, type 0 = JW curl, 1 = the fused second curl + gradient (timed at
 * nrs_tombo_setup).
 * Keys are (kernel, N, type, floor(log2 Nelements)); a
 * setup with opts.autotune records the fastest variant whose result matches
 * the default's.  variant 0 = none recorded (the default is used). */
enum nrs_tune_kernel { NRS_TUNE_AX = 0, NRS_TUNE_FDM = 1, NRS_TUNE_OPERATOR = 2, NRS_TUNE_TOMBO = 3 };
int nrs_tuned_variant(int kernel, int N, int type, int64_t Nelements, int *variant);
/* Record a variant by hand (0 clears it); later launches with this key use it. */
int nrs_set_variant(int kernel, int N, int type, int64_t Nelements, int variant);

/* ------------------------------------------------------------------------- */
/* Device timers (replaces platform->timer device tags, src/core/timer.cpp)   */
/* ------------------------------------------------------------------------- */
enum nrs_phase { NRS_PHASE_AX = 0, NRS_PHASE_GS = 1, NRS_PHASE_COUNT = 2 };

/* Enable (1) or disable (0) per-launch event timing of the operator's Ax and
 * gather-scatter launches; enabling clears previous records. */
int nrs_elliptic_profile_enable(nrs_elliptic *el, int on);
/* Synchronise on the recorded events and return, per phase, the total device
 * milliseconds and the number of launches recorded since enabling. */
int nrs_elliptic_profile_read(nrs_elliptic *el, double ms[NRS_PHASE_COUNT],
                              int64_t launches[NRS_PHASE_COUNT]);

/* ------------------------------------------------------------------------- */
/* Pressure right-hand side of the TOMBO splitting (the step before the solve) */
/* ------------------------------------------------------------------------- */
typedef struct nrs_tombo nrs_tombo;

/* Geometry for tombo::pressureSolve's right-hand side: volume factors
 * (geometricFactorsHex3D.okl:27-143: rx..tz, JW), face normals and WSJ
 * (surfaceGeometricFactorsHex3D.okl) of the faces that carry a flux, the
 * unmasked gather-scatter (nrs->gsh) and invLMM = 1/gs(JW)
 * (meshComputeInvLMM.cpp).  x, y, z, globalIds: host [E*Np] as in
 * nrs_elliptic_desc; bcType: host [E*6] velocity boundary type of each face
 * with the ids of src/bdry/bcType.h (1 W, 2 INT, 3 V, 4-7 SYM*, 16 O, ...;
 * 0 = interior face).  Replaces the mesh and boundary setup those kernels read
 * (meshSetup.cpp, meshSurfaceGeometricFactorsHex3D.cpp, bcMap). */
enum nrs_tombo_flags {
  NRS_TOMBO_FUSION = 1,          /* N = 7: faces shared by consecutive elements summed in the
                                  * curl kernels (bit-identical), gather-scatters without them */
  NRS_TOMBO_AUTOTUNE = 2,        /* time the load schedules, keep the fusion only if faster  */
  NRS_TOMBO_STRESS = 4,          /* VELOCITY STRESSFORMULATION = TRUE (tombo.cpp:57-66): the
                                  * second curl less JW (2 sym(grad Ue) grad(mue)/mue
                                  * - 2/3 div grad(mue)/mue) (pressureStressHex3D.okl)        */
  NRS_TOMBO_STORED_GEOMETRY = 8  /* read the stored rx..tz, JW even where every element is
                                  * trilinear (by default such meshes have the volume kernels
                                  * form them from the element corners)                     */
};
int nrs_tombo_setup(int N, int Nelements, const double *x, const double *y, const double *z,
                    const long long *globalIds, const int *bcType, int flags, nrs_tombo **out);
int nrs_tombo_free(nrs_tombo *h);

/* rhs = the pressure Poisson right-hand side of tombo::pressureSolve
 * (src/navierStokes/tombo.cpp:11-121; with NRS_TOMBO_STRESS the stress term is
 * taken from JW curl W first, tombo.cpp:57-66):
 *   W = gs(JW curl Ue) / gs(JW);  rhsU = BF - mue irho (JW curl W - 4/3 JW grad div)
 *   (rhsU = BF when viscContribution = 0, i.e. nBDF = 1);  rhsU = gs(rhsU) / gs(JW);
 *   rhs = wdiv(rhsU) + g0idt JW div - WSJ n.(g0idt U) on V/INT faces - WSJ n.rhsU on SYM faces.
 * Ue, U, BF: device [3][fieldOffset]; mue, irho, div: device [E*Np]; rhs: device
 * [E*Np] (the input of nrs_elliptic_solve's rhs).  FP64. */
int nrs_tombo_pressure_rhs(nrs_tombo *h, int64_t fieldOffset, const double *Ue, const double *U,
                           const double *BF, const double *mue, const double *irho,
                           const double *div, double g0idt, int viscContribution, double *rhs,
                           void *stream);

/* Face pairs summed in the curl kernels (0: fusion off or dropped by the
 * autotuner) and elements with a flux face. */
int nrs_tombo_info(nrs_tombo *h, int64_t *fusedPairs, int64_t *fluxElements);

/* W = JW curl(U) on [3][fieldOffset] device fields (curlHex3D.okl, scaleJW = 1). */
int nrs_tombo_curl(nrs_tombo *h, int64_t fieldOffset, const double *U, double *W, void *stream);

/* Device buffers owned by the handle: 0 vgeo (f64 [E][10][Np]: rx ry rz sx sy
 * sz tx ty tz JW), 1 invLMM (f64 [E*Np]), 2 the element corners the volume
 * kernels form their factors from (f64 [E][3][8]; count 0 when they stream vgeo). */
int nrs_tombo_buffer(nrs_tombo *h, int which, void **ptr, int64_t *count);
/* Copy `count` entries of buffer `which` into the caller's device array dst. */
int nrs_tombo_copy_buffer(nrs_tombo *h, int which, double *dst, int64_t count, void *stream);

/* ------------------------------------------------------------------------- */
/* Mesh files: .re2 and the numbering of unstructured hexahedral meshes        */
/* ------------------------------------------------------------------------- */
/* Header of a .re2 file (re2::nelg, src/io/re2Reader.cpp:3-41, versions
 * #v002/#v003/#v004): element counts, curved-side and boundary record counts. */
int nrs_re2_info(const char *path, int *nelgt, int *nelgv, int *ncurve, int *nbc);
/* Element corners and boundary codes (3rd_party/nek5000/core/reader_re2.f:
 * readp_re2_mesh/buf_to_xyz :68-157, :430-471, readp_re2_bc/buf_to_bc
 * :503-541).  vx, vy, vz: host [nelgt*8], corners in lexicographic order
 * i + 2j + 4k (the file's cyclic order mapped as meshNekReader.cpp:30);
 * group: host [nelgt] or NULL; cbc: host [nelgt*6*3] three-character codes of
 * the first field ("W  ", "v  ", "O  ", "E  ", ...; blanks where the file has
 * none) in the solver's face order (meshNekReader.cpp:31). Curved sides are
 * read by nrs_re2_curves. */
int nrs_re2_read(const char *path, double *vx, double *vy, double *vz, int *group, char *cbc);
/* The curved-side records of a .re2 file, in file order (readp_re2_curve /
 * buf_to_curve, 3rd_party/nek5000/core/reader_re2.f:159-284,473-498): elem
 * 0-based element, side the file's edge (1..12) or face number, params the
 * record's curve(1:5) (for a midside node 'm': its x, y, z), type its
 * ccurve character.  Host arrays of ncurve (nrs_re2_info), params [ncurve*5]. */
int nrs_re2_curves(const char *path, int *elem, int *side, double *params, char *type);
/* Applies midside-node curved edges ('m') to GLL coordinates made by
 * nrs_mesh_from_vertices: each element with an 'm' edge becomes the
 * triquadratic Gordon-Hall map through its corners and edge midpoints
 * (Nek5000's xyzquad, 3rd_party/nek5000/core/genxyz.f:1334-1410, with
 * gh_face_extend, navier5.f:2501-2590; linquad keeps order 1 trilinear),
 * overwriting its x, y, z [E*(N+1)^3].  Records with a blank type are
 * skipped; circle arcs ('C'), spheres ('s') and 'e' surfaces give
 * NRS_ERR_UNSUPPORTED.  Global ids are unchanged (the topology is). */
int nrs_mesh_midside(int N, int Nelements, const double *vx, const double *vy, const double *vz,
                     int ncurve, const int *elem, const int *side, const double *params,
                     const char *type, double *x, double *y, double *z);
/* GLL node coordinates (trilinear in the corners) and 1-based global ids of
 * an order-N conforming hexahedral mesh from its corners [E*8] (lexicographic):
 * corners closer than tol are one vertex (tol <= 0: 1e-3 of the shortest
 * edge; replaces nek's set_glo_num / parCon connectivity), then every node is
 * numbered by its vertex, edge (ordered by its end vertices), face (ordered
 * from its smallest vertex) or element interior.  Periodic faces are not
 * connected here (nrs_mesh_from_vertices_periodic connects them).
 * x, y, z, globalIds: host [E*(N+1)^3]. */
int nrs_mesh_from_vertices(int N, int Nelements, const double *vx, const double *vy,
                           const double *vz, double tol, double *x, double *y, double *z,
                           long long *globalIds);

/* Periodic face pairs of a .re2 file: for every 'P' boundary record (buf_to_bc,
 * reader_re2.f:503-541: bc(1) = the connected element, bc(2) = its face),
 * peer[(e*6 + f)*2 + 0/1] = the connected element (0-based) and face (solver
 * order); -1 where face f of element e is not periodic.  peer: host [nelgt*6*2]. */
int nrs_re2_periodic(const char *path, int *peer);
/* nrs_mesh_from_vertices with periodic connectivity (nek's periodic vertex
 * numbering in set_glo_num / parCon): the corners of each face pair in `peer`
 * (as nrs_re2_periodic gives it) are one vertex, matched by position relative
 * to the face centres, so the nodes of paired faces share global ids.  Fails
 * (NRS_ERR_INVALID) when the identification joins two corners of one element
 * or two different edges (fewer than three elements across a periodic
 * direction). */
int nrs_mesh_from_vertices_periodic(int N, int Nelements, const double *vx, const double *vy,
                                    const double *vz, double tol, const int *peer, double *x, double *y,
                                    double *z, long long *globalIds);

#ifdef __cplusplus
}
#endif
#endif
