// tune.hpp -- kernel launch variants measured at setup (the autotuner).
//
// A variant changes only how work is grouped onto workgroups and wavefronts
// (elements per block for Ax, elements per wavefront for the matrix-core FDM
// solve); the per-element arithmetic is the same in every variant.  Like the
// reference (src/bench/axHelm/benchmarkAx.cpp:268-383, src/bench/fdm/
// benchmarkFDM.cpp:180-210, enabled by KERNEL AUTOTUNING in
// src/core/registerEllipticKernels.cpp:170-183) each candidate is timed on
// the operator's own data and dropped when its result differs from the
// default variant's by more than 400 eps (Ax) or 100 eps (FDM) relative.
//
// Choices live in one process-wide table keyed by (kernel, order N, data
// type, element-count class = floor(log2 E)); launches look their key up and
// fall back to the default variant when nothing has been tuned.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nrs {

enum TuneKernel { TUNE_AX = 0, TUNE_FDM = 1, TUNE_OP = 2, TUNE_TOMBO = 3, TUNE_KERNELS = 4 };
constexpr int TUNE_TYPES = 8;      // data-type slots (nrs_dtype / nrs_fdm_precision codes)
constexpr int TUNE_ECLASS = 32;    // floor(log2 E)

int tune_eclass(int64_t E);
// tuned variant of (kernel, N, type, E), or dflt when none is recorded
int tuned(int kernel, int N, int type, int64_t E, int dflt);
// record a variant (0 clears the entry)
void set_tuned(int kernel, int N, int type, int64_t E, int variant);
// number of variants of a kernel and their parameter values
int tune_variants(int kernel, int *values, int max);

// deterministic pseudo-random values in [-0.5, 0.5) (hash of the index)
template <typename T> void fill_pattern(int64_t n, T *x, hipStream_t s);

// max_i |a_i - b_i| / |b_i| over b_i != 0 (and |a_i| where b_i == 0)
template <typename T> double max_rel_diff(const T *a, const T *b, int64_t n)
{
  double err = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double d = (double)a[i] - (double)b[i];
    const double r = (b[i] != T(0)) ? d / (double)b[i] : d;
    if (!(r == r)) return 1e300;
    err = r < 0 ? (-r > err ? -r : err) : (r > err ? r : err);
  }
  return err;
}

// Milliseconds per call of f (enqueues one launch sequence on s): two
// untimed calls, then the best of three timed groups of five.
template <typename F> double time_launches(F &&f, hipStream_t s)
{
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  f();
  f();
  float best = 1e30f;
  for (int rep = 0; rep < 3; ++rep) {
    (void)hipEventRecord(a, s);
    for (int k = 0; k < 5; ++k) f();
    (void)hipEventRecord(b, s);
    (void)hipEventSynchronize(b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    if (ms < best) best = ms;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best / 5.0;
}

// Time every Ax variant on this operator's data (element count E, order N),
// check that each gives the default variant's result bit for bit, record the
// fastest.  Returns the chosen elements-per-block multiplier.
template <typename T>
int autotune_ax(int N, int E, const T *ggeo, const T *D, const T *l0, const T *l1, bool poisson,
                hipStream_t s);

}  // namespace nrs
