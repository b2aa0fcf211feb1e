// batch.hip -- many independent elliptic solves of uneven size and order,
// advanced together by captured HIP graphs.
//
// Each problem keeps its own handle (operator, gather-scatter, multigrid with a
// smoothed coarse level, device scalars), and `k` of its iterations are
// captured once into a graph.  A solve runs each problem's start-up work, then
// rounds: the graphs of the problems still active are launched over a pool of
// streams (the small problems' kernels overlap), a kernel gathers every
// problem's "active" flag, one 8-byte-per-problem readback decides the next
// round; at the end the initial guesses are added back.  Iterations of a
// converged problem leave its state unchanged (its update kernels test the
// device flag), so each result equals the problem's stand-alone solve bit for
// bit.  Alternatively (mode 0) all problems are captured into ONE graph as
// parallel branches (fork/join through events) and a converged problem's
// kernel nodes are switched off (hipGraphNodeSetEnabled); on the ROCm graph
// executor that single graph costs more per node than the per-problem graphs.
//
// The reference runs one elliptic_t per field and rank (src/solvers/elliptic/
// ellipticSolve.cpp:32-190); batching independent problems has no counterpart there.
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_set>

#include "batch.hpp"
#include "elliptic.hpp"
#include "kernels.hpp"
#include "linalg.hpp"
#include "mg.hpp"

namespace nrs {

namespace {

__global__ void gather_flags_kernel(int n, double *const *scal, double *flags)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) flags[p] = scal[p][S_ACTIVE];
}

}  // namespace

namespace {

constexpr int kPool = 8;   // streams the per-problem graphs are spread over

// Mode 0: one graph, itersPerLaunch iterations of every problem with the
// problems as parallel branches, and each problem's kernel nodes recorded so
// they can be switched off once it converges.
void capture_one_graph(nrs_batch *b)
{
  const int n = (int)b->probs.size();
  const int itersPerLaunch = b->itersPerLaunch;
  NRS_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  NRS_HIP(hipEventRecord(b->fork, b->cap));
  // the graph under construction, to attribute its nodes to problems
  hipGraph_t building = nullptr;
  {
    hipStreamCaptureStatus cs;
    b->canDisable = hipStreamGetCaptureInfo_v2(b->cap, &cs, nullptr, &building, nullptr, nullptr) ==
                        hipSuccess && building != nullptr;
    (void)hipGetLastError();
  }
  std::vector<hipGraphNode_t> all;
  std::unordered_set<hipGraphNode_t> seen;
  auto new_nodes = [&](std::vector<hipGraphNode_t> &mine) {
    if (!b->canDisable) return;
    size_t cnt = 0;
    if (hipGraphGetNodes(building, nullptr, &cnt) != hipSuccess) { b->canDisable = false; (void)hipGetLastError(); return; }
    all.resize(cnt);
    if (hipGraphGetNodes(building, all.data(), &cnt) != hipSuccess) { b->canDisable = false; (void)hipGetLastError(); return; }
    for (size_t i = 0; i < cnt; ++i) {
      if (!seen.insert(all[i]).second) continue;
      hipGraphNodeType t;
      if (hipGraphNodeGetType(all[i], &t) != hipSuccess) { b->canDisable = false; (void)hipGetLastError(); return; }
      if (t == hipGraphNodeTypeKernel || t == hipGraphNodeTypeMemcpy || t == hipGraphNodeTypeMemset)
        mine.push_back(all[i]);
    }
  };
  b->nodes.resize(n);
  b->disabled.assign(n, 0);
  for (int p = 0; p < n; ++p) {
    hipStream_t st = b->streams[p];
    NRS_HIP(hipStreamWaitEvent(st, b->fork, 0));
    for (int it = 0; it < itersPerLaunch; ++it)
      elliptic_iterate(b->probs[p], b->r[p].get(), b->x[p].get(), st);
    new_nodes(b->nodes[p]);   // the nodes not seen before are p's
    NRS_HIP(hipEventRecord(b->joins[p], st));
    NRS_HIP(hipStreamWaitEvent(b->cap, b->joins[p], 0));
  }
  gather_flags_kernel<<<(n + 63) / 64, 64, 0, b->cap>>>(n, b->scalPtrs.get(), b->flags.get());
  NRS_LAUNCH_CHECK();
  NRS_HIP(hipStreamEndCapture(b->cap, &b->graph));
  NRS_HIP(hipGraphGetNodes(b->graph, nullptr, &b->graphNodes));
  NRS_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}


void capture_per_problem(nrs_batch *b)
{
  const int n = (int)b->probs.size();
  b->pgraph.assign(n, nullptr);
  b->pexec.assign(n, nullptr);
  for (int p = 0; p < n; ++p) {
    hipStream_t st = b->streams[p];
    NRS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < b->itersPerLaunch; ++it)
      elliptic_iterate(b->probs[p], b->r[p].get(), b->x[p].get(), st);
    NRS_HIP(hipStreamEndCapture(st, &b->pgraph[p]));
    NRS_HIP(hipGraphInstantiate(&b->pexec[p], b->pgraph[p], nullptr, nullptr, 0));
  }
  b->poolDone.resize(std::min(n, kPool));
  for (auto &e : b->poolDone) NRS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  b->perProblemNodes = 0;
  for (auto g : b->pgraph) {
    size_t c = 0;
    NRS_HIP(hipGraphGetNodes(g, nullptr, &c));
    b->perProblemNodes += c;
  }
}

// one round of mode 1: the active problems' graphs on the stream pool, then
// the flags gathered on s
void launch_round(nrs_batch *b, const std::vector<int> &active, hipStream_t s)
{
  const int P = (int)b->poolDone.size();
  NRS_HIP(hipEventRecord(b->begun, s));
  for (int q = 0; q < P && q < (int)active.size(); ++q) NRS_HIP(hipStreamWaitEvent(b->streams[q], b->begun, 0));
  for (size_t i = 0; i < active.size(); ++i)
    NRS_HIP(hipGraphLaunch(b->pexec[active[i]], b->streams[i % P]));
  for (int q = 0; q < P && q < (int)active.size(); ++q) {
    NRS_HIP(hipEventRecord(b->poolDone[q], b->streams[q]));
    NRS_HIP(hipStreamWaitEvent(s, b->poolDone[q], 0));
  }
  const int n = (int)b->probs.size();
  gather_flags_kernel<<<(n + 63) / 64, 64, 0, s>>>(n, b->scalPtrs.get(), b->flags.get());
  NRS_LAUNCH_CHECK();
}

}  // namespace

void batch_set_mode(nrs_batch *b, int mode)
{
  NRS_CHECK(mode == 0 || mode == 1, "batch: mode 0 (one graph) or 1 (a graph per problem)");
  if (mode == 0 && !b->exec) capture_one_graph(b);
  if (mode == 1 && b->pexec.empty()) capture_per_problem(b);
  b->mode = mode;
}

nrs_batch *batch_setup(int n, const nrs_elliptic_desc *descs, const nrs_elliptic_opts *opts,
                       int itersPerLaunch)
{
  NRS_CHECK(n >= 1, "batch: at least one problem");
  NRS_CHECK(itersPerLaunch >= 1 && itersPerLaunch <= 64, "batch: iterations per launch 1..64");
  auto b = std::make_unique<nrs_batch>();
  b->itersPerLaunch = itersPerLaunch;
  for (int p = 0; p < n; ++p) {
    NRS_CHECK(descs[p].Nfields <= 1, "batch: one field per problem (a block solve is not device-driven)");
    NRS_CHECK(opts[p].initialGuess == NRS_GUESS_PREVIOUS,
              "batch: the projection initial guess needs host readbacks (not device-driven)");
    NRS_CHECK(opts[p].preconditioner != NRS_PRECON_USER,
              "batch: a user preconditioner is a host callback (not device-driven)");
    NRS_CHECK(opts[p].solver == NRS_SOLVER_PCG || opts[p].solver == NRS_SOLVER_FCG,
              "batch: PCG or flexible PCG (PGMRES needs the host for its Hessenberg system)");
    b->probs.push_back(elliptic_setup(&descs[p], &opts[p]));
    NRS_CHECK(!b->probs.back()->mg || b->probs.back()->mg->device_only(),
              "batch: multigrid problems need a coarse level without host round trips (a smoothed "
              "coarse level, or the AMG coarse solve with coarseTol <= 0)");
    const int64_t L = b->probs.back()->Nlocal;
    b->r.emplace_back(L);
    b->x.emplace_back(L);
  }
  std::vector<double *> sp;
  for (auto *e : b->probs) sp.push_back(e->scal.get());
  b->scalPtrs.upload(sp);
  b->flags.alloc(n);
  NRS_HIP(hipHostMalloc((void **)&b->flagsHost, n * sizeof(double), hipHostMallocDefault));
  b->streams.resize(n);
  b->joins.resize(n);
  for (int p = 0; p < n; ++p) {
    NRS_HIP(hipStreamCreateWithFlags(&b->streams[p], hipStreamNonBlocking));
    NRS_HIP(hipEventCreateWithFlags(&b->joins[p], hipEventDisableTiming));
  }
  NRS_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  NRS_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
  NRS_HIP(hipEventCreateWithFlags(&b->begun, hipEventDisableTiming));
  b->nodes.resize(n);
  b->disabled.assign(n, 0);
  batch_set_mode(b.get(), 1);
  return b.release();
}

int batch_solve(nrs_batch *b, double *const *rhs, double *const *x, nrs_solve_stats *st, hipStream_t s)
{
  const int n = (int)b->probs.size();
  auto t0 = std::chrono::steady_clock::now();
  int maxit = 0;
  for (int p = 0; p < n; ++p) {
    nrs_elliptic *e = b->probs[p];
    const int64_t L = e->Nlocal;
    NRS_HIP(hipMemcpyAsync(b->r[p].get(), rhs[p], L * sizeof(double), hipMemcpyDeviceToDevice, s));
    NRS_HIP(hipMemcpyAsync(b->x[p].get(), x[p], L * sizeof(double), hipMemcpyDeviceToDevice, s));
    elliptic_solve_begin(e, b->r[p].get(), b->x[p].get(), s);
    maxit = std::max(maxit, e->opts.maxIter);
  }
  const int maxLaunches = (maxit + b->itersPerLaunch - 1) / b->itersPerLaunch;
  // every branch runs again (a converged problem's iterations are no-ops on
  // its state; its nodes are switched off only to save the device time)
  auto set_enabled = [&](int p, bool on) {
    for (hipGraphNode_t nd : b->nodes[p]) NRS_HIP(hipGraphNodeSetEnabled(b->exec, nd, on ? 1u : 0u));
    b->disabled[p] = on ? 0 : 1;
  };
  if (b->mode == 0)
    for (int p = 0; p < n; ++p)
      if (b->disabled[p]) set_enabled(p, true);
  std::vector<int> active(n);
  for (int p = 0; p < n; ++p) active[p] = p;
  int launches = 0;
  for (; launches < maxLaunches; ++launches) {
    if (b->mode == 1) {
      launch_round(b, active, s);
      NRS_HIP(hipMemcpyAsync(b->flagsHost, b->flags.get(), n * sizeof(double), hipMemcpyDeviceToHost, s));
      NRS_HIP(hipStreamSynchronize(s));
      std::vector<int> still;
      for (int p : active)
        if (b->flagsHost[p] != 0.0) still.push_back(p);
      active.swap(still);
      if (active.empty()) {
        ++launches;
        break;
      }
      continue;
    }
    NRS_HIP(hipGraphLaunch(b->exec, s));
    NRS_HIP(hipMemcpyAsync(b->flagsHost, b->flags.get(), n * sizeof(double), hipMemcpyDeviceToHost, s));
    NRS_HIP(hipStreamSynchronize(s));
    bool any = false;
    for (int p = 0; p < n; ++p) {
      const bool active = b->flagsHost[p] != 0.0;
      any |= active;
      if (!active && !b->disabled[p] && b->canDisable) set_enabled(p, false);
    }
    if (!any) {
      ++launches;
      break;
    }
  }
  int status = NRS_OK;
  for (int p = 0; p < n; ++p) {
    nrs_elliptic *e = b->probs[p];
    elliptic_solve_end(e, b->x[p].get(), s);
    NRS_HIP(hipMemcpyAsync(x[p], b->x[p].get(), e->Nlocal * sizeof(double), hipMemcpyDeviceToDevice, s));
    NRS_HIP(hipMemcpyAsync(rhs[p], b->r[p].get(), e->Nlocal * sizeof(double), hipMemcpyDeviceToDevice, s));
    NRS_HIP(hipMemcpyAsync(e->scalHost, e->scal.get(), S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  NRS_HIP(hipStreamSynchronize(s));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::string nanProblems;
  for (int p = 0; p < n; ++p) {
    const double *h = b->probs[p]->scalHost;
    if (std::isnan(h[S_RDOTR]) || std::isnan(h[S_TMP3])) {
      status = NRS_ERR_NAN;
      nanProblems += (nanProblems.empty() ? "" : ", ") + std::to_string(p);
    }
    if (st) {
      st[p].iterations = (int)h[S_ITER];
      st[p].res0 = h[S_TMP3];
      st[p].res = h[S_RDOTR];
      st[p].converged = (h[S_RDOTR] <= h[S_TOL]) ? 1 : 0;
      st[p].seconds = secs;
      st[p].precon_seconds = 0.0;
      st[p].res00 = h[S_TMP3];
      st[p].projVectors = 0;
    }
  }
  b->lastLaunches = launches;
  if (status != NRS_OK) set_last_error("batch solve: NaN residual norm in problem(s) " + nanProblems);
  return status;
}

}  // namespace nrs
