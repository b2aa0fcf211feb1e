// tune.hip -- the variant table and the Ax autotuner.
#include <atomic>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "tune.hpp"

namespace nrs {

namespace {

// 0 = not tuned; N is 1..9
std::atomic<int> g_table[TUNE_KERNELS][10][TUNE_TYPES][TUNE_ECLASS];

const int kAxVariants[] = {1, 2, 4};         // elements-per-block multipliers
const int kFdmVariants[] = {1, 2, 4, 8};     // elements per wavefront
const int kOpVariants[] = {1, 2, 4, 8, 16};  // operator chunks (Ax overlapped with gs)
const int kTomboVariants[] = {1, 2};        // pressure-rhs volume kernels: inputs loaded at use / one plane ahead

bool key_ok(int kernel, int N, int type)
{
  return kernel >= 0 && kernel < TUNE_KERNELS && N >= 1 && N <= 9 && type >= 0 && type < TUNE_TYPES;
}

template <typename T> __global__ void pattern_kernel(int64_t n, T *x)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    x[i] = T((double)(h >> 8) * (1.0 / 16777216.0) - 0.5);
  }
}

}  // namespace

int tune_eclass(int64_t E)
{
  int c = 0;
  while (E > 1 && c < TUNE_ECLASS - 1) { E >>= 1; ++c; }
  return c;
}

int tuned(int kernel, int N, int type, int64_t E, int dflt)
{
  if (!key_ok(kernel, N, type)) return dflt;
  const int v = g_table[kernel][N][type][tune_eclass(E)].load(std::memory_order_relaxed);
  return v ? v : dflt;
}

void set_tuned(int kernel, int N, int type, int64_t E, int variant)
{
  NRS_CHECK(key_ok(kernel, N, type), "tune: bad key");
  if (variant) {
    int vals[8];
    const int nv = tune_variants(kernel, vals, 8);
    bool ok = false;
    for (int i = 0; i < nv; ++i) ok |= (vals[i] == variant);
    NRS_CHECK(ok, "tune: not a variant of this kernel");
  }
  g_table[kernel][N][type][tune_eclass(E)].store(variant, std::memory_order_relaxed);
}

int tune_variants(int kernel, int *values, int max)
{
  const int *v = kernel == TUNE_AX ? kAxVariants : kernel == TUNE_FDM ? kFdmVariants
               : kernel == TUNE_OP ? kOpVariants : kTomboVariants;
  const int n = kernel == TUNE_AX ? 3 : kernel == TUNE_FDM ? 4 : kernel == TUNE_OP ? 5 : 2;
  for (int i = 0; i < n && i < max; ++i) values[i] = v[i];
  return n;
}

template <typename T> void fill_pattern(int64_t n, T *x, hipStream_t s)
{
  if (n <= 0) return;
  const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  pattern_kernel<T><<<grid, 256, 0, s>>>(n, x);
  NRS_LAUNCH_CHECK();
}
template void fill_pattern<double>(int64_t, double *, hipStream_t);
template void fill_pattern<float>(int64_t, float *, hipStream_t);

template <typename T>
int autotune_ax(int N, int E, const T *ggeo, const T *D, const T *l0, const T *l1, bool poisson,
                hipStream_t s)
{
  const int Np = (N + 1) * (N + 1) * (N + 1);
  const int64_t L = (int64_t)E * Np;
  DevBuf<T> q(L), ref(L), out(L);
  fill_pattern<T>(L, q.get(), s);
  auto run = [&](int mult, T *o) {
    launch_ax_variant<T>(N, mult, E, nullptr, ggeo, D, l0, l1, poisson, false, q.get(), o, nullptr, s);
  };
  run(1, ref.get());
  std::vector<T> hr(L), ho(L);
  NRS_HIP(hipMemcpyAsync(hr.data(), ref.get(), L * sizeof(T), hipMemcpyDeviceToHost, s));
  NRS_HIP(hipStreamSynchronize(s));
  const double eps = std::is_same<T, double>::value ? 2.220446049250313e-16 : 1.1920928955078125e-07;
  int best = 1;
  double bestMs = 1e300;
  for (int mult : kAxVariants) {
    run(mult, out.get());
    NRS_HIP(hipMemcpyAsync(ho.data(), out.get(), L * sizeof(T), hipMemcpyDeviceToHost, s));
    NRS_HIP(hipStreamSynchronize(s));
    if (max_rel_diff(ho.data(), hr.data(), L) > 400.0 * eps) continue;   // benchmarkAx.cpp:288-304
    const double ms = time_launches([&] { run(mult, out.get()); }, s);
    if (ms < bestMs) { bestMs = ms; best = mult; }
  }
  set_tuned(TUNE_AX, N, dtype_of<T>::value, E, best);
  return best;
}
template int autotune_ax<double>(int, int, const double *, const double *, const double *,
                                 const double *, bool, hipStream_t);
template int autotune_ax<float>(int, int, const float *, const float *, const float *,
                                const float *, bool, hipStream_t);

}  // namespace nrs
