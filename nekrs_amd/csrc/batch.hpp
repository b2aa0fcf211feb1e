// batch.hpp -- independent solves advanced by one captured graph (internal layout of nrs_batch).
#pragma once
#include <vector>

#include "elliptic.hpp"

struct nrs_batch {
  std::vector<nrs_elliptic *> probs;
  std::vector<nrs::DevBuf<double>> r, x;     // internal rhs / solution per problem (graph-fixed)
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t graphNodes = 0;
  int itersPerLaunch = 1;
  int lastLaunches = 0;
  nrs::DevBuf<double *> scalPtrs;
  nrs::DevBuf<double> flags;
  double *flagsHost = nullptr;
  // kernel/copy nodes of each problem's branch, switched off once it converges
  std::vector<std::vector<hipGraphNode_t>> nodes;
  std::vector<char> disabled;
  bool canDisable = false;
  // mode 1: one graph per problem, launched for the active problems only on a
  // pool of streams (captured at the first solve in that mode)
  int mode = 0;
  std::vector<hipGraph_t> pgraph;
  std::vector<hipGraphExec_t> pexec;
  std::vector<hipEvent_t> poolDone;
  hipEvent_t begun = nullptr;
  size_t perProblemNodes = 0;
  ~nrs_batch()
  {
    for (auto g : pexec) if (g) (void)hipGraphExecDestroy(g);
    for (auto g : pgraph) if (g) (void)hipGraphDestroy(g);
    for (auto e : poolDone) (void)hipEventDestroy(e);
    if (begun) (void)hipEventDestroy(begun);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : joins) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    for (auto st : streams) (void)hipStreamDestroy(st);
    if (cap) (void)hipStreamDestroy(cap);
    if (flagsHost) (void)hipHostFree(flagsHost);
    for (auto *p : probs) delete p;
  }
};

