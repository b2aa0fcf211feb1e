// linalg.hip -- BLAS-1 on E-vectors and the fused PCG vector kernels.
//
// Reductions replace weightedInnerProdMany / ellipticBlockUpdatePCG + host sums
// (src/linAlg/linAlg.cpp:986-1033, linearSolver/PCG.cpp:33-83): a fixed grid of
// 1024 blocks writes one partial per block and the last block to arrive (atomic
// ticket, acq_rel at agent scope) adds the partials in a fixed tree order, so the
// result is deterministic and stays in device memory.  alpha and beta are formed
// on the device by that last block.
#include "linalg.hpp"

namespace nrs {

Reducer::Reducer() : partials(kBlocks * 4), ticket(1)
{
  NRS_HIP(hipMemset(ticket.get(), 0, sizeof(unsigned)));
}

namespace {

constexpr int RB = Reducer::kBlocks;
constexpr int RT = Reducer::kThreads;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-reduce NV values; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
  __shared__ double s[NV][RT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    const double w = wave_sum(v[c]);
    if (lane == 0) s[c][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      double a = 0.0;
#pragma unroll
      for (int k = 0; k < RT / 64; ++k) a += s[c][k];
      v[c] = a;
    }
  }
}

// Publish this block's partials; returns true (in every thread) for the last
// block, whose thread 0 then holds the grid totals in tot[].
template <int NV>
__device__ bool grid_finish(double (&v)[NV], double *partials, unsigned *ticket, double (&tot)[NV])
{
  __shared__ int s_last;
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NV; ++c) partials[blockIdx.x * NV + c] = v[c];
    const unsigned t =
        __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  double a[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    a[c] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += RT)
      a[c] += __hip_atomic_load(&partials[b * NV + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum<NV>(a);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NV; ++c) tot[c] = a[c];
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

template <typename T>
__global__ void __launch_bounds__(RT)
wdot_kernel(int64_t n, const T *__restrict__ w, const T *__restrict__ a, const T *__restrict__ b,
            double *partials, unsigned *ticket, double *scal, int slot)
{
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT)
    v[0] += (double)w[i] * (double)a[i] * (double)b[i];
  double tot[1];
  if (grid_finish<1>(v, partials, ticket, tot) && threadIdx.x == 0) scal[slot] = tot[0];
}

__global__ void __launch_bounds__(RT)
wnorm_kernel(int64_t n, const double *__restrict__ w, const double *__restrict__ a, double factor,
             double *partials, unsigned *ticket, double *scal, int slot)
{
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT)
    v[0] += w[i] * a[i] * a[i];
  double tot[1];
  if (grid_finish<1>(v, partials, ticket, tot) && threadIdx.x == 0) scal[slot] = sqrt(tot[0] * factor);
}

__global__ void __launch_bounds__(RT)
plane_sum_kernel(int64_t E, int Np, int64_t estride, int64_t off, const double *__restrict__ x,
                 double *partials, unsigned *ticket, double *scal, int slot)
{
  double v[1] = {0.0};
  const int64_t n = E * Np;
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    const int64_t e = i / Np;
    v[0] += x[e * estride + off + (i - e * Np)];
  }
  double tot[1];
  if (grid_finish<1>(v, partials, ticket, tot) && threadIdx.x == 0) scal[slot] = tot[0];
}

template <typename T> __global__ void fill_kernel(int64_t n, T v, T *x)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}
template <typename T> __global__ void axpby_kernel(int64_t n, T a, const T *__restrict__ x, T b, T *__restrict__ y)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
template <typename T> __global__ void axmyz_kernel(int64_t n, const T *__restrict__ x, const T *__restrict__ y, T *__restrict__ z)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    z[i] = x[i] * y[i];
}
template <typename T> __global__ void reciprocal_kernel(int64_t n, T *y)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = T(1) / y[i];
}
template <typename A, typename B> __global__ void convert_kernel(int64_t n, const A *__restrict__ x, B *__restrict__ y)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (B)x[i];
}

enum { PM_JACOBI = 0, PM_GIVEN = 1, PM_NONE = 2 };

// Preconditioned dots and beta (PCG.cpp:115-159).
template <int MODE, bool FLEX>
__global__ void __launch_bounds__(RT)
precon_dots_kernel(int64_t n, const double *__restrict__ w, const double *__restrict__ invDiag,
                   const double *__restrict__ r, double *__restrict__ z,
                   const double *__restrict__ Ap, int iter, double *partials, unsigned *ticket,
                   double *scal)
{
  double v[2] = {0.0, 0.0};
  const bool needZAp = FLEX && iter > 1;
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    const double ri = r[i];
    double zi;
    if (MODE == PM_JACOBI) { zi = invDiag[i] * ri; z[i] = zi; }
    else if (MODE == PM_GIVEN) zi = z[i];
    else zi = ri;
    const double wi = w[i];
    if (MODE != PM_NONE) v[0] += wi * ri * zi;
    if (needZAp) v[1] += wi * zi * Ap[i];
  }
  double tot[2];
  if (grid_finish<2>(v, partials, ticket, tot) && threadIdx.x == 0) {
    const double rdotz2 = scal[S_RDOTZ1];
    // PCG.cpp:128-130: with no preconditioner rdotz1 is the running residual norm
    const double rdotz1 = (MODE == PM_NONE) ? scal[S_RDOTR] : tot[0];
    double beta = 0.0;
    if (iter > 1) {
      beta = rdotz1 / rdotz2;
      if (FLEX) beta = -scal[S_ALPHA] * tot[1] / rdotz2;
    }
    scal[S_RDOTZ2] = rdotz2;
    scal[S_RDOTZ1] = rdotz1;
    scal[S_ZDOTAP] = tot[1];
    scal[S_BETA] = beta;
  }
}

__global__ void update_p_kernel(int64_t n, const double *__restrict__ z, double *__restrict__ p,
                                const double *__restrict__ scal)
{
  const double beta = scal[S_BETA];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

__global__ void __launch_bounds__(RT)
alpha_kernel(int64_t n, const double *__restrict__ w, const double *__restrict__ p,
             const double *__restrict__ Ap, double *partials, unsigned *ticket, double *scal)
{
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT)
    v[0] += w[i] * p[i] * Ap[i];
  double tot[1];
  if (grid_finish<1>(v, partials, ticket, tot) && threadIdx.x == 0) {
    scal[S_PAP] = tot[0];
    scal[S_ALPHA] = scal[S_RDOTZ1] / (tot[0] + 1e-300);
  }
}

// ellipticBlockUpdatePCG (kernels/elliptic/ellipticBlockUpdatePCG.c:27-50) fused
// with x += alpha p (PCG.cpp:67-74).
__global__ void __launch_bounds__(RT)
update_rx_kernel(int64_t n, const double *__restrict__ w, const double *__restrict__ Ap,
                 const double *__restrict__ p, double *__restrict__ r, double *__restrict__ x,
                 double resNormFactor, double *partials, unsigned *ticket, double *scal)
{
  const double alpha = scal[S_ALPHA];
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    const double rn = r[i] - alpha * Ap[i];
    v[0] += rn * rn * w[i];
    r[i] = rn;
    x[i] += alpha * p[i];
  }
  double tot[1];
  if (grid_finish<1>(v, partials, ticket, tot) && threadIdx.x == 0)
    scal[S_RDOTR] = sqrt(tot[0] * resNormFactor);
}

constexpr int EB = 256;
inline unsigned egrid(int64_t n) { return grid_for(n, EB, 8192); }

}  // namespace

void wdot(Reducer &R, int64_t n, const double *w, const double *a, const double *b, double *scal,
          int slot, hipStream_t s)
{
  wdot_kernel<double><<<RB, RT, 0, s>>>(n, w, a, b, R.partials.get(), R.ticket.get(), scal, slot);
  NRS_LAUNCH_CHECK();
}
void wdot_f32(Reducer &R, int64_t n, const float *w, const float *a, const float *b, double *scal,
              int slot, hipStream_t s)
{
  wdot_kernel<float><<<RB, RT, 0, s>>>(n, w, a, b, R.partials.get(), R.ticket.get(), scal, slot);
  NRS_LAUNCH_CHECK();
}
void wnorm(Reducer &R, int64_t n, const double *w, const double *a, double factor, double *scal,
           int slot, hipStream_t s)
{
  wnorm_kernel<<<RB, RT, 0, s>>>(n, w, a, factor, R.partials.get(), R.ticket.get(), scal, slot);
  NRS_LAUNCH_CHECK();
}

void plane_sum(Reducer &R, int64_t E, int Np, int64_t estride, int64_t off, const double *x,
               double *scal, int slot, hipStream_t s)
{
  plane_sum_kernel<<<RB, RT, 0, s>>>(E, Np, estride, off, x, R.partials.get(), R.ticket.get(), scal, slot);
  NRS_LAUNCH_CHECK();
}

template <typename T> void fill(int64_t n, T v, T *x, hipStream_t s)
{
  if (n <= 0) return;
  fill_kernel<T><<<egrid(n), EB, 0, s>>>(n, v, x);
  NRS_LAUNCH_CHECK();
}
template <typename T> void axpby(int64_t n, T a, const T *x, T b, T *y, hipStream_t s)
{
  if (n <= 0) return;
  axpby_kernel<T><<<egrid(n), EB, 0, s>>>(n, a, x, b, y);
  NRS_LAUNCH_CHECK();
}
template <typename T> void axmyz(int64_t n, const T *x, const T *y, T *z, hipStream_t s)
{
  if (n <= 0) return;
  axmyz_kernel<T><<<egrid(n), EB, 0, s>>>(n, x, y, z);
  NRS_LAUNCH_CHECK();
}
template <typename T> void reciprocal(int64_t n, T *y, hipStream_t s)
{
  if (n <= 0) return;
  reciprocal_kernel<T><<<egrid(n), EB, 0, s>>>(n, y);
  NRS_LAUNCH_CHECK();
}
template <typename A, typename B> void convert(int64_t n, const A *x, B *y, hipStream_t s)
{
  if (n <= 0) return;
  convert_kernel<A, B><<<egrid(n), EB, 0, s>>>(n, x, y);
  NRS_LAUNCH_CHECK();
}

template void fill<double>(int64_t, double, double *, hipStream_t);
template void fill<float>(int64_t, float, float *, hipStream_t);
template void fill<int>(int64_t, int, int *, hipStream_t);
template void axpby<double>(int64_t, double, const double *, double, double *, hipStream_t);
template void axpby<float>(int64_t, float, const float *, float, float *, hipStream_t);
template void axmyz<double>(int64_t, const double *, const double *, double *, hipStream_t);
template void axmyz<float>(int64_t, const float *, const float *, float *, hipStream_t);
template void reciprocal<double>(int64_t, double *, hipStream_t);
template void reciprocal<float>(int64_t, float *, hipStream_t);
template void convert<double, float>(int64_t, const double *, float *, hipStream_t);
template void convert<float, double>(int64_t, const float *, double *, hipStream_t);

void pcg_precon_dots(Reducer &R, int64_t n, const double *w, const double *invDiag, const double *r,
                     double *z, const double *Ap, bool jacobi, bool flexible, bool none_precon,
                     int iter, double *scal, hipStream_t s)
{
#define PD(MODE, FLEX) \
  precon_dots_kernel<MODE, FLEX><<<RB, RT, 0, s>>>(n, w, invDiag, r, z, Ap, iter, R.partials.get(), R.ticket.get(), scal)
  const int mode = none_precon ? PM_NONE : (jacobi ? PM_JACOBI : PM_GIVEN);
  if (mode == PM_JACOBI) { if (flexible) PD(PM_JACOBI, true); else PD(PM_JACOBI, false); }
  else if (mode == PM_GIVEN) { if (flexible) PD(PM_GIVEN, true); else PD(PM_GIVEN, false); }
  else { if (flexible) PD(PM_NONE, true); else PD(PM_NONE, false); }
#undef PD
  NRS_LAUNCH_CHECK();
}

void pcg_update_p(int64_t n, const double *z, double *p, const double *scal, hipStream_t s)
{
  update_p_kernel<<<egrid(n), EB, 0, s>>>(n, z, p, scal);
  NRS_LAUNCH_CHECK();
}

void pcg_alpha(Reducer &R, int64_t n, const double *w, const double *p, const double *Ap,
               double *scal, hipStream_t s)
{
  alpha_kernel<<<RB, RT, 0, s>>>(n, w, p, Ap, R.partials.get(), R.ticket.get(), scal);
  NRS_LAUNCH_CHECK();
}

void pcg_update_rx(Reducer &R, int64_t n, const double *w, const double *Ap, const double *p,
                   double *r, double *x, double resNormFactor, double *scal, hipStream_t s)
{
  update_rx_kernel<<<RB, RT, 0, s>>>(n, w, Ap, p, r, x, resNormFactor, R.partials.get(),
                                     R.ticket.get(), scal);
  NRS_LAUNCH_CHECK();
}

}  // namespace nrs
