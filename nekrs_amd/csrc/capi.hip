// capi.hip -- the extern "C" boundary of libnekrs_hip (declared in include/nrs_hip.h).
// Every entry point validates its arguments, runs, and converts exceptions into
// status codes; nothing here aborts the process.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "batch.hpp"
#include "elliptic.hpp"
#include "tune.hpp"
#include "fdm.hpp"
#include "kernels.hpp"
#include "mesh.hpp"
#include "mg.hpp"
#include "re2.hpp"
#include "schedule.hpp"
#include "tombo.hpp"

namespace nrs {
static thread_local std::string g_last_error;
void set_last_error(const std::string &m) { g_last_error = m; }
}  // namespace nrs

// Scratch of the allocation-free reduction (nrs_linalg_reduce_noalloc): the
// per-block partials and the result slot, in a file-scope device array that
// each GPU holds for the life of the process.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in symbol): the
// device address is looked up once per device with hipGetSymbolAddress, with
// that device current, and kept in g_scratch_ptr.
namespace {
constexpr int kScratchPartials = nrs::Reducer::kBlocks;           // the reduction grid's cap
constexpr int kScratchDoubles = kScratchPartials + 32;            // + the result, 256-B aligned
__device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxScratchDevices] = {};

// Looks up (once) the address of `scratch` on `device`, which must be current.
double *scratch_setup(int device)
{
  if (device < 0 || device >= kMaxScratchDevices)
    throw nrs::Error(NRS_ERR_UNSUPPORTED, "scratch: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    NRS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw nrs::Error(NRS_ERR_HIP, "scratch: hipGetSymbolAddress returned null");
    g_scratch_ptr[device] = static_cast<double *>(ptr);
  }
  return g_scratch_ptr[device];
}
}  // namespace

#define NRS_API_BEGIN try {
#define NRS_API_END                                   \
  return NRS_OK;                                      \
  }                                                   \
  catch (const nrs::Error &e) {                       \
    nrs::set_last_error(e.what());                    \
    return e.code;                                    \
  }                                                   \
  catch (const std::exception &e) {                   \
    nrs::set_last_error(e.what());                    \
    return NRS_ERR_INVALID;                           \
  }

using nrs::as_stream;

extern "C" {

const char *nrs_last_error(void) { return nrs::g_last_error.c_str(); }
const char *nrs_version(void) { return "nekrs_amd 0.1 (gfx950)"; }

int nrs_init(int device)
{
  NRS_API_BEGIN
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    throw nrs::Error(NRS_ERR_NO_DEVICE, "no HIP device");
  NRS_CHECK(device >= 0 && device < n, "device index out of range");
  NRS_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  NRS_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    throw nrs::Error(NRS_ERR_UNSUPPORTED, std::string("library built for gfx950, device is ") + prop.gcnArchName);
  scratch_setup(device);
  NRS_API_END
}

int nrs_box_mesh(int nx, int ny, int nz, int N, int map, double eps, double *x, double *y, double *z,
                 long long *globalIds, int *EToB)
{
  NRS_API_BEGIN
  NRS_CHECK(x && y && z, "box mesh: null coordinate arrays");
  NRS_CHECK(map >= 0 && map <= 2, "box mesh: map must be 0 (affine), 1 (sine) or 2 (Kershaw)");
  nrs::box_mesh(nx, ny, nz, N, map, eps, x, y, z, globalIds, EToB);
  NRS_API_END
}

int nrs_gll(int N, double *z, double *w, double *D)
{
  NRS_API_BEGIN
  NRS_CHECK(N >= 1 && N <= 32, "gll: N must be 1..32");
  std::vector<double> zz, ww;
  nrs::gll_nodes(N, zz, ww);
  if (z) std::memcpy(z, zz.data(), zz.size() * sizeof(double));
  if (w) std::memcpy(w, ww.data(), ww.size() * sizeof(double));
  if (D) {
    const auto DD = nrs::dmatrix(zz);
    std::memcpy(D, DD.data(), DD.size() * sizeof(double));
  }
  NRS_API_END
}

int nrs_geometric_factors(int N, int Nelements, const double *x, const double *y, const double *z,
                          double *ggeo, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(N >= 1 && N <= 9, "geometric factors: N must be 1..9");
  NRS_CHECK(Nelements >= 0, "geometric factors: negative element count");
  NRS_CHECK(Nelements == 0 || (x && y && z && ggeo), "geometric factors: null array");
  std::vector<double> zz, ww;
  nrs::gll_nodes(N, zz, ww);
  nrs::DevBuf<double> D, w;
  D.upload(nrs::dmatrix(zz));
  w.upload(ww);
  nrs::launch_geofacs(N, Nelements, D.get(), w.get(), x, y, z, ggeo, as_stream(stream));
  NRS_HIP(hipStreamSynchronize(as_stream(stream)));
  NRS_API_END
}

int nrs_ax(int N, int dtype, int Nelements, const int *elementList, const void *ggeo, const void *D,
           const void *lambda0, const void *lambda1, int poisson, int lambda_field, const void *q,
           void *Aq, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(N >= 1 && N <= 9, "Ax: N must be 1..9");
  NRS_CHECK(Nelements >= 0, "Ax: negative element count");
  if (Nelements == 0) return NRS_OK;
  NRS_CHECK(ggeo && D && lambda0 && q && Aq, "Ax: null array");
  NRS_CHECK(poisson || lambda1, "Ax: Helmholtz needs lambda1");
  const hipStream_t s = as_stream(stream);
  if (dtype == NRS_F64)
    nrs::launch_ax<double>(N, Nelements, elementList, (const double *)ggeo, (const double *)D,
                           (const double *)lambda0, (const double *)lambda1, poisson != 0,
                           lambda_field != 0, (const double *)q, (double *)Aq, nullptr, s);
  else if (dtype == NRS_F32)
    nrs::launch_ax<float>(N, Nelements, elementList, (const float *)ggeo, (const float *)D,
                          (const float *)lambda0, (const float *)lambda1, poisson != 0,
                          lambda_field != 0, (const float *)q, (float *)Aq, nullptr, s);
  else
    throw nrs::Error(NRS_ERR_INVALID, "Ax: dtype must be NRS_F64 or NRS_F32");
  NRS_API_END
}

int nrs_block_ax(int N, int Nelements, int64_t offset, int64_t loffset, const double *ggeo,
                 const double *D, const double *lambda0, const double *lambda1, const double *q,
                 double *Aq, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(N >= 1 && N <= 9, "block Ax: N must be 1..9");
  NRS_CHECK(Nelements >= 0, "block Ax: negative element count");
  if (Nelements == 0) return NRS_OK;
  const int64_t L = (int64_t)Nelements * (N + 1) * (N + 1) * (N + 1);
  NRS_CHECK(offset >= L && loffset >= L, "block Ax: field offsets smaller than E*Np");
  NRS_CHECK(ggeo && D && lambda0 && lambda1 && q && Aq, "block Ax: null array");
  nrs::launch_block_ax(N, Nelements, offset, loffset, ggeo, D, lambda0, lambda1, q, Aq, as_stream(stream));
  NRS_API_END
}

int nrs_build_diagonal(int N, int dtype, int Nelements, const void *ggeo, const void *D,
                       const void *lambda0, const void *lambda1, int poisson, int lambda_field,
                       void *diag, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(N >= 1 && N <= 9, "diagonal: N must be 1..9");
  NRS_CHECK(Nelements >= 0, "diagonal: negative element count");
  if (Nelements == 0) return NRS_OK;
  NRS_CHECK(ggeo && D && lambda0 && diag && (poisson || lambda1), "diagonal: null array");
  const hipStream_t s = as_stream(stream);
  if (dtype == NRS_F64)
    nrs::launch_build_diag<double>(N, Nelements, (const double *)ggeo, (const double *)D,
                                   (const double *)lambda0, (const double *)lambda1, poisson != 0,
                                   lambda_field != 0, (double *)diag, s);
  else if (dtype == NRS_F32)
    nrs::launch_build_diag<float>(N, Nelements, (const float *)ggeo, (const float *)D,
                                  (const float *)lambda0, (const float *)lambda1, poisson != 0,
                                  lambda_field != 0, (float *)diag, s);
  else
    throw nrs::Error(NRS_ERR_INVALID, "diagonal: dtype must be NRS_F64 or NRS_F32");
  NRS_API_END
}

int nrs_mask(int dtype, int64_t Nmasked, const int *maskIds, void *v, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(Nmasked >= 0, "mask: negative count");
  if (Nmasked == 0) return NRS_OK;
  NRS_CHECK(maskIds && v, "mask: null array");
  const hipStream_t s = as_stream(stream);
  if (dtype == NRS_F64) nrs::launch_mask<double>(Nmasked, maskIds, (double *)v, s);
  else if (dtype == NRS_F32) nrs::launch_mask<float>(Nmasked, maskIds, (float *)v, s);
  else throw nrs::Error(NRS_ERR_INVALID, "mask: dtype must be NRS_F64 or NRS_F32");
  NRS_API_END
}

int nrs_coarsen(int Nf, int Nc, int dtype, int Nelements, const void *R, const void *qf, void *qc,
                void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(Nelements >= 0, "coarsen: negative element count");
  if (Nelements == 0) return NRS_OK;
  NRS_CHECK(R && qf && qc, "coarsen: null array");
  const hipStream_t s = as_stream(stream);
  if (dtype == NRS_F64) nrs::launch_coarsen<double>(Nf, Nc, Nelements, (const double *)R, (const double *)qf, (double *)qc, s);
  else if (dtype == NRS_F32) nrs::launch_coarsen<float>(Nf, Nc, Nelements, (const float *)R, (const float *)qf, (float *)qc, s);
  else throw nrs::Error(NRS_ERR_INVALID, "coarsen: dtype must be NRS_F64 or NRS_F32");
  NRS_API_END
}

int nrs_prolongate(int Nf, int Nc, int dtype, int Nelements, const void *R, const void *qc, void *qf,
                   void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(Nelements >= 0, "prolongate: negative element count");
  if (Nelements == 0) return NRS_OK;
  NRS_CHECK(R && qf && qc, "prolongate: null array");
  const hipStream_t s = as_stream(stream);
  if (dtype == NRS_F64) nrs::launch_prolongate<double>(Nf, Nc, Nelements, (const double *)R, (const double *)qc, (double *)qf, s);
  else if (dtype == NRS_F32) nrs::launch_prolongate<float>(Nf, Nc, Nelements, (const float *)R, (const float *)qc, (float *)qf, s);
  else throw nrs::Error(NRS_ERR_INVALID, "prolongate: dtype must be NRS_F64 or NRS_F32");
  NRS_API_END
}

int nrs_ogs_setup(int64_t N, const long long *globalIds, nrs_ogs **ogs)
{
  NRS_API_BEGIN
  NRS_CHECK(ogs != nullptr, "ogs setup: null output handle");
  *ogs = nullptr;
  *ogs = nrs::ogs_setup(N, globalIds);
  NRS_API_END
}

int nrs_ogs_free(nrs_ogs *ogs)
{
  NRS_API_BEGIN
  delete ogs;
  NRS_API_END
}

int nrs_ogs_gather_scatter(nrs_ogs *ogs, void *v, int nvec, int64_t stride, int dtype, int op, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(ogs && (v || ogs->N == 0), "gather-scatter: null argument");
  NRS_CHECK(op >= NRS_ADD && op <= NRS_MAX, "gather-scatter: unknown op");
  const hipStream_t s = as_stream(stream);
  if (dtype == NRS_F64) nrs::ogs_gather_scatter<double>(ogs, (double *)v, nvec, stride, op, s);
  else if (dtype == NRS_F32) nrs::ogs_gather_scatter<float>(ogs, (float *)v, nvec, stride, op, s);
  else if (dtype == NRS_I32) nrs::ogs_gather_scatter<int>(ogs, (int *)v, nvec, stride, op, s);
  else throw nrs::Error(NRS_ERR_INVALID, "gather-scatter: unknown dtype");
  NRS_API_END
}

int nrs_ogs_start(nrs_ogs *ogs, void *v, int nvec, int64_t stride, int dtype, int op, void *stream)
{
  NRS_API_BEGIN
  (void)stream;
  NRS_CHECK(ogs && (v || ogs->N == 0), "gather-scatter start: null argument");
  NRS_CHECK(op >= NRS_ADD && op <= NRS_MAX, "gather-scatter start: unknown op");
  NRS_CHECK(dtype == NRS_F64 || dtype == NRS_F32 || dtype == NRS_I32, "gather-scatter start: unknown dtype");
  NRS_CHECK(nvec >= 1 && (nvec == 1 || stride >= ogs->N), "gather-scatter start: bad nvec / stride");
  NRS_CHECK(!ogs->pending, "gather-scatter start: the previous start has no finish");
  // one device holds every row: no halo to pack or exchange, v is not read here
  ogs->pend = {v, nvec, stride, dtype, op};
  ogs->pending = true;
  NRS_API_END
}

int nrs_ogs_finish(nrs_ogs *ogs, void *v, int nvec, int64_t stride, int dtype, int op, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(ogs, "gather-scatter finish: null handle");
  NRS_CHECK(ogs->pending, "gather-scatter finish: no matching start");
  const auto &p = ogs->pend;
  NRS_CHECK(p.v == v && p.nvec == nvec && p.stride == stride && p.dtype == dtype && p.op == op,
            "gather-scatter finish: arguments differ from the start");
  ogs->pending = false;
  return nrs_ogs_gather_scatter(ogs, v, nvec, stride, dtype, op, stream);
  NRS_API_END
}

int nrs_ogs_inv_degree(nrs_ogs *ogs, double *invDegree, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(ogs && (invDegree || ogs->N == 0), "invDegree: null argument");
  if (ogs->N)
    NRS_HIP(hipMemcpyAsync(invDegree, ogs->invDegree.get(), ogs->N * sizeof(double),
                           hipMemcpyDeviceToDevice, as_stream(stream)));
  NRS_API_END
}

int nrs_ogs_info(nrs_ogs *ogs, int64_t *Nrows, int64_t *Nmembers, int64_t *Nsingletons)
{
  NRS_API_BEGIN
  NRS_CHECK(ogs, "ogs info: null handle");
  if (Nrows) *Nrows = ogs->Nrows;
  if (Nmembers) *Nmembers = ogs->Nmembers;
  if (Nsingletons) *Nsingletons = ogs->Nsingletons;
  NRS_API_END
}

int nrs_linalg_reduce(int op, int64_t N, int Nfields, int64_t fieldOffset, const double *w,
                      const double *x, const double *y, double *result, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(result, "linalg reduce: null result");
  NRS_CHECK(op >= NRS_RED_SUM && op <= NRS_RED_DOT, "linalg reduce: unknown reduction");
  NRS_CHECK(N >= 0 && Nfields >= 1, "linalg reduce: N must be >= 0 and Nfields >= 1");
  NRS_CHECK(Nfields == 1 || fieldOffset >= N, "linalg reduce: fieldOffset must be >= N");
  NRS_CHECK(N == 0 || x, "linalg reduce: null x");
  NRS_CHECK(N == 0 || !(op == NRS_RED_DOT || op == NRS_RED_WDOT) || y, "linalg reduce: null y");
  NRS_CHECK(N == 0 || !(op == NRS_RED_WNORM2 || op == NRS_RED_WDOT) || w, "linalg reduce: null weight");
  const hipStream_t s = as_stream(stream);
  nrs::DevBuf<double> part((size_t)nrs::reduce_partials_needed(N) + 1);
  nrs::linalg_reduce(op, N, Nfields, fieldOffset, w, x, y, part.get() + 1, part.get(), s);
  NRS_HIP(hipMemcpyAsync(result, part.get(), sizeof(double), hipMemcpyDeviceToHost, s));
  NRS_HIP(hipStreamSynchronize(s));
  NRS_API_END
}

int nrs_linalg_reduce_noalloc(int op, int64_t N, int Nfields, int64_t fieldOffset, const double *w,
                              const double *x, const double *y, double *result, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(result, "linalg reduce (no alloc): null result");
  NRS_CHECK(op >= NRS_RED_SUM && op <= NRS_RED_DOT, "linalg reduce (no alloc): unknown reduction");
  NRS_CHECK(N >= 0 && Nfields >= 1, "linalg reduce (no alloc): N must be >= 0 and Nfields >= 1");
  NRS_CHECK(Nfields == 1 || fieldOffset >= N, "linalg reduce (no alloc): fieldOffset must be >= N");
  NRS_CHECK(N == 0 || x, "linalg reduce (no alloc): null x");
  NRS_CHECK(N == 0 || !(op == NRS_RED_DOT || op == NRS_RED_WDOT) || y,
            "linalg reduce (no alloc): null y");
  NRS_CHECK(N == 0 || !(op == NRS_RED_WNORM2 || op == NRS_RED_WDOT) || w,
            "linalg reduce (no alloc): null weight");
  const int partials = nrs::reduce_partials_needed(N);
  if (partials > kScratchPartials)
    throw nrs::Error(NRS_ERR_INVALID, "linalg reduce (no alloc): " + std::to_string(partials) +
                                          " partial sums exceed the scratch of " +
                                          std::to_string(kScratchPartials));
  int device = 0;
  NRS_HIP(hipGetDevice(&device));
  double *const sp = scratch_setup(device);      // the device address, never `scratch` itself
  const hipStream_t s = as_stream(stream);
  // every partial slot the finish kernel reads is written by the streaming kernel
  // of the same call, so nothing is cleared first
  nrs::linalg_reduce(op, N, Nfields, fieldOffset, w, x, y, sp, sp + kScratchPartials, s);
  NRS_HIP(hipMemcpyAsync(result, sp + kScratchPartials, sizeof(double), hipMemcpyDeviceToHost, s));
  NRS_HIP(hipStreamSynchronize(s));
  NRS_API_END
}

int nrs_block_update_pcg(int64_t N, int Nfields, int64_t offset, const double *invDegree,
                         const double *Ap, double alpha, double *r, double *rdotr, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(rdotr, "update pcg: null rdotr");
  NRS_CHECK(N >= 0 && Nfields >= 1, "update pcg: N must be >= 0 and Nfields >= 1");
  NRS_CHECK(Nfields == 1 || offset >= N, "update pcg: offset must be >= N");
  NRS_CHECK(N == 0 || (invDegree && Ap && r), "update pcg: null argument");
  const hipStream_t s = as_stream(stream);
  nrs::DevBuf<double> part((size_t)nrs::reduce_partials_needed(N) + 1);
  nrs::block_update_pcg(N, Nfields, offset, invDegree, Ap, alpha, r, part.get() + 1, part.get(), s);
  NRS_HIP(hipMemcpyAsync(rdotr, part.get(), sizeof(double), hipMemcpyDeviceToHost, s));
  NRS_HIP(hipStreamSynchronize(s));
  NRS_API_END
}

void nrs_elliptic_default_opts(nrs_elliptic_opts *o)
{
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->solver = NRS_SOLVER_PCG;
  o->preconditioner = NRS_PRECON_JACOBI;
  o->tol = 1e-8;
  o->relative = 1;
  o->maxIter = 500;
  o->nLevels = 0;
  o->smoother = NRS_SMOOTHER_CHEBYSHEV_JACOBI;
  o->chebyDegree = 2;
  o->chebyMinFactor = 0.1;
  o->chebyMaxFactor = 1.1;
  o->coarseSolver = NRS_COARSE_PCG_F64;
  o->coarseTol = 1e-2;
  o->coarseMaxIter = 200;
  o->mgPrecision = NRS_MG_F32;
  o->fdmPrecision = NRS_FDM_F32;
  o->seed = 0;
  o->useGraphs = 1;
  o->autotune = 1;
  o->gmresRestart = 15;
  o->initialGuess = NRS_GUESS_PREVIOUS;
  o->projVectors = 8;
  o->projStart = 5;
  o->operatorFusion = 1;
  o->mgGeometry = 1;
  o->precoCoeffField = 1;
}

int nrs_elliptic_setup(const nrs_elliptic_desc *desc, const nrs_elliptic_opts *opts, nrs_elliptic **out)
{
  NRS_API_BEGIN
  NRS_CHECK(out != nullptr, "elliptic setup: null output handle");
  *out = nullptr;
  *out = nrs::elliptic_setup(desc, opts);
  NRS_API_END
}

int nrs_elliptic_free(nrs_elliptic *el)
{
  NRS_API_BEGIN
  delete el;
  NRS_API_END
}

int nrs_elliptic_info(nrs_elliptic *el, int64_t *Nlocal, int64_t *Nmasked, double *volume)
{
  NRS_API_BEGIN
  NRS_CHECK(el, "elliptic info: null handle");
  if (Nlocal) *Nlocal = el->Nlocal;
  if (Nmasked) *Nmasked = el->Nmasked;
  if (volume) *volume = el->volume;
  NRS_API_END
}

int nrs_elliptic_buffer(nrs_elliptic *el, int which, void **ptr, int64_t *count)
{
  NRS_API_BEGIN
  NRS_CHECK(el && ptr, "elliptic buffer: null argument");
  void *p = nullptr;
  int64_t n = 0;
  switch (which) {
    case 0: p = el->ggeo.get(); n = (int64_t)el->ggeo.size(); break;
    case 1: p = el->D.get(); n = (int64_t)el->D.size(); break;
    case 2: p = el->ogs->invDegree.get(); n = (int64_t)el->ogs->invDegree.size(); break;
    case 3: p = el->maskIds.get(); n = el->Nmasked; break;
    case 4: p = el->invDiagA.get(); n = (int64_t)el->invDiagA.size(); break;
    default: throw nrs::Error(NRS_ERR_INVALID, "elliptic buffer: unknown buffer id");
  }
  *ptr = p;
  if (count) *count = n;
  NRS_API_END
}

int nrs_elliptic_copy_buffer(nrs_elliptic *el, int which, void *dst, int64_t count, void *stream)
{
  NRS_API_BEGIN
  void *p = nullptr;
  int64_t n = 0;
  const int rc = nrs_elliptic_buffer(el, which, &p, &n);
  if (rc != NRS_OK) return rc;
  NRS_CHECK(count >= 0 && count <= n, "copy buffer: count exceeds the buffer");
  NRS_CHECK(dst || count == 0, "copy buffer: null destination");
  const size_t esz = (which == 3) ? sizeof(int) : sizeof(double);
  if (count)
    NRS_HIP(hipMemcpyAsync(dst, p, count * esz, hipMemcpyDeviceToDevice, as_stream(stream)));
  NRS_API_END
}

int nrs_elliptic_set_coefficients(nrs_elliptic *el, const double *lambda0, const double *lambda1,
                                  void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(el, "set coefficients: null handle");
  nrs::elliptic_set_coefficients(el, lambda0, lambda1, as_stream(stream));
  NRS_API_END
}

int nrs_elliptic_update_jacobi(nrs_elliptic *el, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(el, "update jacobi: null handle");
  nrs::elliptic_update_jacobi(el, as_stream(stream));
  NRS_API_END
}

int nrs_elliptic_operator(nrs_elliptic *el, const double *q, double *Aq, int masked, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(el && q && Aq, "operator: null argument");
  nrs::elliptic_operator(el, q, Aq, masked != 0, as_stream(stream));
  NRS_API_END
}

int nrs_elliptic_preconditioner(nrs_elliptic *el, const double *r, double *z, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(el && r && z, "preconditioner: null argument");
  nrs::elliptic_preconditioner(el, r, z, as_stream(stream));
  NRS_API_END
}

int nrs_elliptic_mg_info(nrs_elliptic *el, int max, int *nLevels, int *orders, int64_t *Nlocal,
                         double *rho, int *degree, int *coarseIterations)
{
  NRS_API_BEGIN
  NRS_CHECK(el && nLevels, "mg info: null argument");
  NRS_CHECK(el->mg != nullptr, "mg info: the preconditioner is not multigrid");
  const auto info = el->mg->info();
  *nLevels = (int)info.size();
  for (int i = 0; i < (int)info.size() && i < max; ++i) {
    if (orders) orders[i] = info[i].N;
    if (Nlocal) Nlocal[i] = info[i].Nlocal;
    if (rho) rho[i] = info[i].rho;
    if (degree) degree[i] = info[i].degree;
  }
  if (coarseIterations) *coarseIterations = el->mg->coarse_iterations();
  NRS_API_END
}

int nrs_batch_setup(int n, const nrs_elliptic_desc *descs, const nrs_elliptic_opts *opts,
                    int itersPerLaunch, nrs_batch **out)
{
  NRS_API_BEGIN
  NRS_CHECK(descs && opts && out, "batch setup: null argument");
  *out = nullptr;
  *out = nrs::batch_setup(n, descs, opts, itersPerLaunch);
  NRS_API_END
}

int nrs_batch_solve(nrs_batch *b, double *const *rhs, double *const *x, nrs_solve_stats *stats,
                    void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(b && rhs && x, "batch solve: null argument");
  return nrs::batch_solve(b, rhs, x, stats, nrs::as_stream(stream));
  NRS_API_END
}

int nrs_batch_info(nrs_batch *b, int *n, int64_t *graphNodes, int *lastLaunches)
{
  NRS_API_BEGIN
  NRS_CHECK(b, "batch info: null handle");
  if (n) *n = (int)b->probs.size();
  if (graphNodes) *graphNodes = (int64_t)(b->mode == 0 ? b->graphNodes : b->perProblemNodes);
  if (lastLaunches) *lastLaunches = b->lastLaunches;
  NRS_API_END
}

int nrs_batch_free(nrs_batch *b)
{
  NRS_API_BEGIN
  delete b;
  NRS_API_END
}

int nrs_batch_set_mode(nrs_batch *b, int mode)
{
  NRS_API_BEGIN
  NRS_CHECK(b, "batch mode: null handle");
  nrs::batch_set_mode(b, mode);
  NRS_API_END
}

int nrs_batch_problem(nrs_batch *b, int p, nrs_elliptic **el)
{
  NRS_API_BEGIN
  NRS_CHECK(b && el, "batch problem: null argument");
  NRS_CHECK(p >= 0 && p < (int)b->probs.size(), "batch problem: index out of range");
  *el = b->probs[p];
  NRS_API_END
}

int nrs_tuned_variant(int kernel, int N, int type, int64_t Nelements, int *variant)
{
  NRS_API_BEGIN
  NRS_CHECK(variant != nullptr, "tuned variant: null output");
  NRS_CHECK(kernel >= NRS_TUNE_AX && kernel <= NRS_TUNE_TOMBO, "tuned variant: unknown kernel");
  NRS_CHECK(N >= 1 && N <= 9 && type >= 0 && type < nrs::TUNE_TYPES && Nelements >= 1,
            "tuned variant: bad key");
  *variant = nrs::tuned(kernel, N, type, Nelements, 0);
  NRS_API_END
}

int nrs_set_variant(int kernel, int N, int type, int64_t Nelements, int variant)
{
  NRS_API_BEGIN
  NRS_CHECK(kernel >= NRS_TUNE_AX && kernel <= NRS_TUNE_TOMBO, "set variant: unknown kernel");
  NRS_CHECK(N >= 1 && N <= 9 && type >= 0 && type < nrs::TUNE_TYPES && Nelements >= 1,
            "set variant: bad key");
  nrs::set_tuned(kernel, N, type, Nelements, variant);
  NRS_API_END
}

int nrs_elliptic_mg_smoother(nrs_elliptic *el, int level, const void *in, void *out, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(el && in && out, "mg smoother: null argument");
  NRS_CHECK(el->mg != nullptr, "mg smoother: the preconditioner is not multigrid");
  el->mg->level_smoother(level, in, out, nrs::as_stream(stream));
  NRS_API_END
}

int nrs_fdm_1d(int N, int lbc, int rbc, double ll, double lm, double lr, double *S, double *lam)
{
  NRS_API_BEGIN
  NRS_CHECK(N >= 1 && N <= 15, "fdm 1d: order out of range");
  NRS_CHECK(S && lam, "fdm 1d: null argument");
  NRS_CHECK(lbc >= 0 && lbc <= 2 && rbc >= 0 && rbc <= 2, "fdm 1d: boundary flags are 0, 1 or 2");
  NRS_CHECK(ll > 0 && lm > 0 && lr > 0, "fdm 1d: lengths must be positive");
  std::vector<double> z, w;
  nrs::gll_nodes(N, z, w);
  const auto D = nrs::dmatrix(z);
  NRS_CHECK(nrs::fdm_1d(N, D.data(), w.data(), lbc, rbc, ll, lm, lr, S, lam),
            "fdm 1d: mass matrix not positive definite");
  NRS_API_END
}

int nrs_fdm_element_lengths(int N, int Nelements, const double *x, const double *y, const double *z,
                            const long long *globalIds, double *lengths)
{
  NRS_API_BEGIN
  NRS_CHECK(N >= 1 && Nelements >= 0, "element lengths: bad sizes");
  NRS_CHECK(Nelements == 0 || (x && y && z && globalIds && lengths), "element lengths: null array");
  const auto links = nrs::face_links(N, Nelements, globalIds);
  const auto len = nrs::element_lengths(N, Nelements, x, y, z, links);
  std::memcpy(lengths, len.data(), len.size() * sizeof(double));
  NRS_API_END
}

int nrs_coarse_global_ids(int Nf, int Nelements, const long long *idsFine, int Nc, long long *idsCoarse)
{
  NRS_API_BEGIN
  NRS_CHECK(Nf >= 1 && Nc >= 1 && Nc <= Nf, "coarse ids: need 1 <= Nc <= Nf");
  NRS_CHECK(Nelements >= 0, "coarse ids: negative element count");
  NRS_CHECK(Nelements == 0 || (idsFine && idsCoarse), "coarse ids: null array");
  const auto ids = nrs::coarse_global_ids(Nf, Nelements, idsFine, Nc);
  std::memcpy(idsCoarse, ids.data(), ids.size() * sizeof(long long));
  NRS_API_END
}

int nrs_elliptic_profile_enable(nrs_elliptic *el, int on)
{
  NRS_API_BEGIN
  NRS_CHECK(el, "profile: null handle");
  el->profiling = on != 0;
  el->spans.clear();
  el->eventNext = 0;
  NRS_API_END
}

int nrs_elliptic_profile_read(nrs_elliptic *el, double ms[NRS_PHASE_COUNT], int64_t launches[NRS_PHASE_COUNT])
{
  NRS_API_BEGIN
  NRS_CHECK(el && ms && launches, "profile: null argument");
  for (int p = 0; p < NRS_PHASE_COUNT; ++p) { ms[p] = 0.0; launches[p] = 0; }
  for (const auto &sp : el->spans) {
    NRS_HIP(hipEventSynchronize(sp.b));
    float t = 0.0f;
    NRS_HIP(hipEventElapsedTime(&t, sp.a, sp.b));
    ms[sp.phase] += t;
    launches[sp.phase] += 1;
  }
  NRS_API_END
}

int nrs_elliptic_solve(nrs_elliptic *el, double *rhs, double *x, nrs_solve_stats *stats, void *stream)
{
  try {
    NRS_CHECK(el && rhs && x, "solve: null argument");
    const int rc = nrs::elliptic_solve(el, rhs, x, stats, as_stream(stream));
    if (rc != NRS_OK) nrs::set_last_error("solve: NaN residual norm");
    return rc;
  } catch (const nrs::Error &e) {
    nrs::set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    nrs::set_last_error(e.what());
    return NRS_ERR_INVALID;
  }
}

int nrs_elliptic_mg_level(nrs_elliptic *el, int level, int op, void *in, const void *x, void *out,
                          int xIsZero, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(el && el->mg, "mg level: handle without a multigrid preconditioner");
  el->mg->level_op(level, op, in, x, out, xIsZero != 0, as_stream(stream));
  NRS_API_END
}

int nrs_parse_mg_schedule(const char *schedule, int N, int defaultDegree, nrs_elliptic_opts *opts)
{
  NRS_API_BEGIN
  NRS_CHECK(schedule && opts, "schedule: null argument");
  nrs::parse_mg_schedule(schedule, N, defaultDegree, opts);
  NRS_API_END
}

int nrs_elliptic_set_user_preconditioner(nrs_elliptic *el, nrs_user_precon_fn fn, void *user)
{
  NRS_API_BEGIN
  NRS_CHECK(el != nullptr, "user preconditioner: null handle");
  el->userPrecon = fn;
  el->userCtx = fn ? user : nullptr;
  NRS_API_END
}

int nrs_elliptic_set_user_coarse_solver(nrs_elliptic *el, nrs_user_coarse_fn fn, void *user)
{
  NRS_API_BEGIN
  NRS_CHECK(el != nullptr, "user coarse solver: null handle");
  el->userCoarse = fn;
  el->userCoarseCtx = fn ? user : nullptr;
  NRS_API_END
}

int nrs_elliptic_krylov(nrs_elliptic *el, double *r, double *x, double tol, int maxIter,
                        double *rdotr, int *iterations, void *stream)
{
  try {
    NRS_CHECK(el && r && x && rdotr, "krylov: null argument");
    NRS_CHECK(maxIter >= 1, "krylov: maxIter must be >= 1");
    int iter = 0;
    double res = *rdotr, precon_s = 0.0;
    NRS_CHECK(std::isfinite(res) && res >= 0.0, "krylov: the initial norm must be finite and >= 0");
    const int rc = nrs::elliptic_krylov(el, r, x, res, tol, maxIter, false, iter, res, precon_s,
                                        as_stream(stream));
    *rdotr = res;
    if (iterations) *iterations = iter;
    if (rc != NRS_OK) nrs::set_last_error("krylov: NaN residual norm");
    return rc;
  } catch (const nrs::Error &e) {
    nrs::set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    nrs::set_last_error(e.what());
    return NRS_ERR_INVALID;
  }
}

}  // extern "C"

int nrs_tombo_setup(int N, int Nelements, const double *x, const double *y, const double *z,
                    const long long *globalIds, const int *bcType, int flags, nrs_tombo **out)
{
  NRS_API_BEGIN
  NRS_CHECK(out != nullptr, "tombo setup: null output handle");
  *out = nullptr;
  *out = nrs::tombo_setup(N, Nelements, x, y, z, globalIds, bcType, flags);
  NRS_API_END
}

int nrs_tombo_free(nrs_tombo *h)
{
  NRS_API_BEGIN
  delete h;
  NRS_API_END
}

int nrs_tombo_pressure_rhs(nrs_tombo *h, int64_t fieldOffset, const double *Ue, const double *U,
                           const double *BF, const double *mue, const double *irho,
                           const double *div, double g0idt, int viscContribution, double *rhs,
                           void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(h, "tombo: null handle");
  NRS_CHECK(fieldOffset >= h->Nlocal, "tombo: fieldOffset smaller than E*Np");
  NRS_CHECK(U && BF && div && rhs, "tombo: null array");
  NRS_CHECK(!viscContribution || (Ue && mue && irho), "tombo: null array for the viscous term");
  nrs::tombo_pressure_rhs(h, fieldOffset, Ue, U, BF, mue, irho, div, g0idt, viscContribution != 0,
                          rhs, as_stream(stream));
  NRS_API_END
}

int nrs_tombo_curl(nrs_tombo *h, int64_t fieldOffset, const double *U, double *W, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(h && U && W, "tombo curl: null argument");
  NRS_CHECK(fieldOffset >= h->Nlocal, "tombo curl: fieldOffset smaller than E*Np");
  NRS_CHECK(U != W, "tombo curl: in place is not supported");
  nrs::tombo_curl(h, fieldOffset, U, W, as_stream(stream));
  NRS_API_END
}

int nrs_tombo_buffer(nrs_tombo *h, int which, void **ptr, int64_t *count)
{
  NRS_API_BEGIN
  NRS_CHECK(h && ptr && count, "tombo buffer: null argument");
  switch (which) {
    case 0: *ptr = h->vgeo.get(); *count = (int64_t)h->vgeo.n; break;
    case 1: *ptr = h->invLMM.get(); *count = (int64_t)h->invLMM.n; break;
    case 2: *ptr = h->corners.get(); *count = (int64_t)h->corners.n; break;
    default: throw nrs::Error(NRS_ERR_INVALID, "tombo buffer: which must be 0, 1 or 2");
  }
  NRS_API_END
}

int nrs_tombo_copy_buffer(nrs_tombo *h, int which, double *dst, int64_t count, void *stream)
{
  NRS_API_BEGIN
  NRS_CHECK(h && dst, "tombo copy buffer: null argument");
  const nrs::DevBuf<double> *b = which == 0 ? &h->vgeo : which == 1 ? &h->invLMM : nullptr;
  NRS_CHECK(b != nullptr, "tombo copy buffer: which must be 0 or 1");
  NRS_CHECK(count >= 0 && (size_t)count <= b->n, "tombo copy buffer: count exceeds the buffer");
  NRS_HIP(hipMemcpyAsync(dst, b->get(), count * sizeof(double), hipMemcpyDeviceToDevice,
                         as_stream(stream)));
  NRS_HIP(hipStreamSynchronize(as_stream(stream)));
  NRS_API_END
}

int nrs_re2_info(const char *path, int *nelgt, int *nelgv, int *ncurve, int *nbc)
{
  NRS_API_BEGIN
  const nrs::Re2Header h = nrs::re2_header(path);
  if (nelgt) *nelgt = h.nelgt;
  if (nelgv) *nelgv = h.nelgv;
  if (ncurve) *ncurve = h.ncurve;
  if (nbc) *nbc = h.nbc;
  NRS_API_END
}

int nrs_re2_read(const char *path, double *vx, double *vy, double *vz, int *group, char *cbc)
{
  NRS_API_BEGIN
  NRS_CHECK(vx && vy && vz, "re2 read: null corner arrays");
  nrs::re2_read(path, vx, vy, vz, group, cbc);
  NRS_API_END
}

int nrs_re2_curves(const char *path, int *elem, int *side, double *params, char *type)
{
  NRS_API_BEGIN
  NRS_CHECK(path && elem && side && params && type, "re2 curves: null argument");
  nrs::re2_curves(path, elem, side, params, type);
  NRS_API_END
}

int nrs_mesh_midside(int N, int Nelements, const double *vx, const double *vy, const double *vz,
                     int ncurve, const int *elem, const int *side, const double *params,
                     const char *type, double *x, double *y, double *z)
{
  NRS_API_BEGIN
  NRS_CHECK(Nelements > 0 && ncurve >= 0, "midside: bad sizes");
  NRS_CHECK(vx && vy && vz && x && y && z, "midside: null array");
  NRS_CHECK(ncurve == 0 || (elem && side && params && type), "midside: null curve array");
  nrs::midside_map(N, Nelements, vx, vy, vz, ncurve, elem, side, params, type, x, y, z);
  NRS_API_END
}

int nrs_re2_periodic(const char *path, int *peer)
{
  NRS_API_BEGIN
  NRS_CHECK(path && peer, "re2: null argument");
  const nrs::Re2Header h = nrs::re2_header(path);
  std::vector<double> vx((size_t)h.nelgt * 8), vy(vx.size()), vz(vx.size());
  nrs::re2_read(path, vx.data(), vy.data(), vz.data(), nullptr, nullptr, peer);
  NRS_API_END
}

int nrs_mesh_from_vertices_periodic(int N, int Nelements, const double *vx, const double *vy,
                                    const double *vz, double tol, const int *peer, double *x, double *y,
                                    double *z, long long *globalIds)
{
  NRS_API_BEGIN
  NRS_CHECK(vx && vy && vz && peer && x && y && z && globalIds, "mesh: null argument");
  nrs::mesh_from_vertices(N, Nelements, vx, vy, vz, tol, x, y, z, globalIds, peer);
  NRS_API_END
}

int nrs_mesh_from_vertices(int N, int Nelements, const double *vx, const double *vy,
                           const double *vz, double tol, double *x, double *y, double *z,
                           long long *globalIds)
{
  NRS_API_BEGIN
  NRS_CHECK(vx && vy && vz && x && y && z && globalIds, "mesh from vertices: null array");
  nrs::mesh_from_vertices(N, Nelements, vx, vy, vz, tol, x, y, z, globalIds);
  NRS_API_END
}

int nrs_elliptic_fusion_info(nrs_elliptic *el, int64_t *fusedPairs)
{
  NRS_API_BEGIN
  NRS_CHECK(el && fusedPairs, "fusion info: null argument");
  *fusedPairs = el->fusedPairs;
  NRS_API_END
}

int nrs_tombo_info(nrs_tombo *h, int64_t *fusedPairs, int64_t *fluxElements)
{
  NRS_API_BEGIN
  NRS_CHECK(h, "tombo info: null handle");
  if (fusedPairs) *fusedPairs = h->fusedPairs;
  if (fluxElements) *fluxElements = h->Nb;
  NRS_API_END
}
