// linalg.hip -- BLAS-1 on E-vectors and the fused PCG vector kernels.
//
// Reductions replace weightedInnerProdMany / ellipticBlockUpdatePCG + host sums
// (src/linAlg/linAlg.cpp:986-1033, linearSolver/PCG.cpp:33-83).  Each reduction
// is two launches: the streaming kernel writes one partial per block (grid
// sized to the vector, at most 1024 blocks), a one-block kernel adds them in a
// fixed order and performs the scalar step of the algorithm (alpha, beta, the
// residual norm, the convergence flag), so the result is deterministic and
// stays in device memory.
#include "linalg.hpp"

namespace nrs {

Reducer::Reducer() : partials(kBlocks * 2) {}

namespace {

constexpr int RB = Reducer::kBlocks;
constexpr int RT = Reducer::kThreads;

inline int rgrid(int64_t n)
{
  const int64_t per = (int64_t)RT * 8;    // at least 8 entries per thread
  int64_t g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > RB) g = RB;
  return (int)g;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-reduce NV values in a fixed order; the result is valid in thread 0.
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

// partials laid out [c][block]
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double *partials)
{
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NV; ++c) partials[c * gridDim.x + blockIdx.x] = v[c];
  }
}

// sum of nb partials per component, valid in thread 0
template <int NV>
__device__ __forceinline__ void sum_partials(int nb, const double *partials, double (&tot)[NV])
{
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    tot[c] = 0.0;
    for (int b = threadIdx.x; b < nb; b += RT) tot[c] += partials[c * nb + b];
  }
  block_sum<NV>(tot);
}

// ------------------------------------------------------------------ stage-1 kernels

template <typename T>
__global__ void __launch_bounds__(RT)
wdot_kernel(int64_t n, const T *__restrict__ w, const T *__restrict__ a, const T *__restrict__ b,
            double *partials)
{
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT)
    v[0] += (double)w[i] * (double)a[i] * (double)b[i];
  write_partials<1>(v, partials);
}

__global__ void __launch_bounds__(RT)
plane_sum_kernel(int64_t E, int Np, int64_t estride, int64_t off, const double *__restrict__ x,
                 double *partials)
{
  double v[1] = {0.0};
  const int64_t n = E * Np;
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    const int64_t e = i / Np;
    v[0] += x[e * estride + off + (i - e * Np)];
  }
  write_partials<1>(v, partials);
}

enum { PM_JACOBI = 0, PM_GIVEN = 1, PM_NONE = 2 };

// Preconditioned dots (PCG.cpp:115-153): rdotz = <r,z>_w, zdotAp = <z,Ap>_w.
template <int MODE, bool FLEX>
__global__ void __launch_bounds__(RT)
precon_dots_kernel(int64_t n, const double *__restrict__ w, const double *__restrict__ invDiag,
                   const double *__restrict__ r, double *__restrict__ z,
                   const double *__restrict__ Ap, const double *__restrict__ scal, double *partials)
{
  if (scal[S_ACTIVE] == 0.0) return;
  const bool needZAp = FLEX && scal[S_ITER] >= 1.0;
  double v[2] = {0.0, 0.0};
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
  write_partials<2>(v, partials);
}

__global__ void update_p_kernel(int64_t n, const double *__restrict__ z, double *__restrict__ p,
                                const double *__restrict__ scal)
{
  if (scal[S_ACTIVE] == 0.0) return;
  const double beta = scal[S_BETA];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

__global__ void __launch_bounds__(RT)
pap_kernel(int64_t n, const double *__restrict__ w, const double *__restrict__ p,
           const double *__restrict__ Ap, const double *__restrict__ scal, double *partials)
{
  if (scal[S_ACTIVE] == 0.0) return;
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT)
    v[0] += w[i] * p[i] * Ap[i];
  write_partials<1>(v, partials);
}

// ellipticBlockUpdatePCG (kernels/elliptic/ellipticBlockUpdatePCG.c:27-50) fused
// with x += alpha p (PCG.cpp:67-74).
__global__ void __launch_bounds__(RT)
update_rx_kernel(int64_t n, const double *__restrict__ w, const double *__restrict__ Ap,
                 const double *__restrict__ p, double *__restrict__ r, double *__restrict__ x,
                 const double *__restrict__ scal, double *partials)
{
  if (scal[S_ACTIVE] == 0.0) return;
  const double alpha = scal[S_ALPHA];
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    const double rn = r[i] - alpha * Ap[i];
    v[0] += rn * rn * w[i];
    r[i] = rn;
    x[i] += alpha * p[i];
  }
  write_partials<1>(v, partials);
}

// ------------------------------------------------------------------ stage-2 kernels

enum { FIN_STORE = 0, FIN_SQRT = 1, FIN_DOTS = 2, FIN_ALPHA = 3, FIN_RDOTR = 4 };

template <int NV, int KIND, int MODE = 0, bool FLEX = false>
__global__ void __launch_bounds__(RT)
finish_kernel(int nb, const double *__restrict__ partials, double *scal, int slot, double factor)
{
  if (KIND >= FIN_DOTS && scal[S_ACTIVE] == 0.0) return;
  double tot[NV];
  sum_partials<NV>(nb, partials, tot);
  if (threadIdx.x != 0) return;
  if (KIND == FIN_STORE) {
    scal[slot] = tot[0];
  } else if (KIND == FIN_SQRT) {
    scal[slot] = sqrt(tot[0] * factor);
  } else if (KIND == FIN_DOTS) {
    const int iter = (int)scal[S_ITER] + 1;
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
    scal[S_ZDOTAP] = NV > 1 ? tot[NV - 1] : 0.0;
    scal[S_BETA] = beta;
  } else if (KIND == FIN_ALPHA) {
    scal[S_PAP] = tot[0];
    scal[S_ALPHA] = scal[S_RDOTZ1] / (tot[0] + 1e-300);
  } else {   // FIN_RDOTR
    const double rdotr = sqrt(tot[0] * factor);
    const double iter = scal[S_ITER] + 1.0;
    scal[S_RDOTR] = rdotr;
    scal[S_ITER] = iter;
    scal[S_ACTIVE] = (rdotr > scal[S_TOL] && iter < scal[S_MAXIT]) ? 1.0 : 0.0;   // NaN stops too
  }
}

__global__ void pcg_init_kernel(double *scal, double rdotr0, double tol, int maxit)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    scal[S_RDOTR] = rdotr0;
    scal[S_RDOTZ1] = 0.0;
    scal[S_ALPHA] = 0.0;
    scal[S_TOL] = tol;
    scal[S_ITER] = 0.0;
    scal[S_MAXIT] = (double)maxit;
    scal[S_ACTIVE] = 1.0;
  }
}

// The same initialisation from the device-resident initial norm in S_RDOTR
// (no host round trip): S_TMP3 keeps res0, the tolerance is relative to it
// when asked; a NaN norm leaves the iteration inactive.
__global__ void pcg_init_dev_kernel(double *scal, double tol, int relative, int maxit)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double r0 = scal[S_RDOTR];
    scal[S_TMP3] = r0;
    scal[S_RDOTZ1] = 0.0;
    scal[S_ALPHA] = 0.0;
    scal[S_TOL] = relative ? tol * r0 : tol;
    scal[S_ITER] = 0.0;
    scal[S_MAXIT] = (double)maxit;
    scal[S_ACTIVE] = (r0 == r0) ? 1.0 : 0.0;
  }
}

// ------------------------------------------------------------------ element-wise kernels

template <typename T> __global__ void fill_kernel(int64_t n, T v, T *x)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}
// y = a x + b y; with b == 0 the old y is not read, so a freshly allocated
// (uninitialised, possibly non-finite) y is simply overwritten.
template <typename T> __global__ void axpby_kernel(int64_t n, T a, const T *__restrict__ x, T b, T *__restrict__ y)
{
  if (b == T(0)) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
      y[i] = a * x[i];
    return;
  }
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

constexpr int EB = 256;
inline unsigned egrid(int64_t n) { return grid_for(n, EB, 8192); }

// ------------------------------------------------------------------ linAlg_t reductions
// One streaming kernel over N entries of Nfields fields spaced fieldOffset
// apart, one value per thread accumulated in a fixed order, one partial per
// block; a one-block kernel combines the partials in a fixed tree order.
template <int OP> struct RedTraits {
  static constexpr bool kMax = OP == NRS_RED_MAX || OP == NRS_RED_AMAX;
  static constexpr bool kMin = OP == NRS_RED_MIN;
  static __device__ __forceinline__ double identity()
  {
    return kMax ? (OP == NRS_RED_AMAX ? 0.0 : -HUGE_VAL) : (kMin ? HUGE_VAL : 0.0);
  }
  static __device__ __forceinline__ double combine(double a, double b)
  {
    return kMax ? fmax(a, b) : (kMin ? fmin(a, b) : a + b);
  }
};

template <int OP>
__device__ __forceinline__ double red_block(double v, double *s)
{
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = RedTraits<OP>::combine(v, __shfl_down(v, off, 64));
  if (lane == 0) s[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = s[0];
#pragma unroll
    for (int k = 1; k < RT / 64; ++k) v = RedTraits<OP>::combine(v, s[k]);
  }
  return v;
}

template <int OP>
__global__ void __launch_bounds__(RT)
reduce_many_kernel(int64_t n, int nf, int64_t off, const double *__restrict__ w,
                   const double *__restrict__ x, const double *__restrict__ y, double *partials)
{
  __shared__ double s[RT / 64];
  double v = RedTraits<OP>::identity();
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    double t = 0.0;
    for (int f = 0; f < nf; ++f) {
      const double a = x[i + f * off];
      double c;
      switch (OP) {
        case NRS_RED_SUM: case NRS_RED_MAX: case NRS_RED_MIN: c = a; break;
        case NRS_RED_NORM2: case NRS_RED_WNORM2: c = a * a; break;
        case NRS_RED_NORM1: case NRS_RED_AMAX: c = fabs(a); break;
        case RED_WSUM: c = a; break;
        default: c = a * y[i + f * off]; break;   // NRS_RED_DOT, NRS_RED_WDOT
      }
      t = f == 0 ? c : RedTraits<OP>::combine(t, c);
    }
    if (OP == NRS_RED_WNORM2 || OP == NRS_RED_WDOT || OP == RED_WSUM) t *= w[i];
    v = RedTraits<OP>::combine(v, t);
  }
  v = red_block<OP>(v, s);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

template <int OP>
__global__ void __launch_bounds__(RT) reduce_finish_kernel(int nb, const double *__restrict__ partials, double *out)
{
  __shared__ double s[RT / 64];
  double v = RedTraits<OP>::identity();
  for (int b = threadIdx.x; b < nb; b += RT) v = RedTraits<OP>::combine(v, partials[b]);
  v = red_block<OP>(v, s);
  if (threadIdx.x == 0) out[0] = (OP == NRS_RED_NORM2 || OP == NRS_RED_WNORM2) ? sqrt(v) : v;
}

// y += factor * scal[slot] (the scalar read on the device)
__global__ void add_scalar_kernel(int64_t n, double factor, const double *__restrict__ scal, int slot,
                                  double *__restrict__ y)
{
  const double a = factor * scal[slot];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a;
}

// ellipticBlockUpdatePCG (kernels/elliptic/ellipticBlockUpdatePCG.c:27-50):
// r -= alpha Ap on every field, partial sums of invDegree r^2.
__global__ void __launch_bounds__(RT)
block_update_pcg_kernel(int64_t n, int nf, int64_t off, const double *__restrict__ w,
                        const double *__restrict__ Ap, double alpha, double *__restrict__ r,
                        double *partials)
{
  __shared__ double s[RT / 64];
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    const double wi = w[i];
    for (int f = 0; f < nf; ++f) {
      const int64_t k = i + f * off;
      const double rn = r[k] - alpha * Ap[k];
      v += rn * rn * wi;
      r[k] = rn;
    }
  }
  v = red_block<NRS_RED_SUM>(v, s);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

}  // namespace

int reduce_partials_needed(int64_t n) { return rgrid(n); }

void linalg_reduce(int op, int64_t n, int nf, int64_t off, const double *w, const double *x,
                   const double *y, double *partials, double *out, hipStream_t s)
{
  const int g = rgrid(n);
#define NRS_RED_CASE(OP_)                                                                         case OP_:                                                                                         reduce_many_kernel<OP_><<<g, RT, 0, s>>>(n, nf, off, w, x, y, partials);                        reduce_finish_kernel<OP_><<<1, RT, 0, s>>>(g, partials, out);                                   break;
  switch (op) {
    NRS_RED_CASE(NRS_RED_SUM) NRS_RED_CASE(NRS_RED_MAX) NRS_RED_CASE(NRS_RED_MIN)
    NRS_RED_CASE(NRS_RED_NORM2) NRS_RED_CASE(NRS_RED_NORM1) NRS_RED_CASE(NRS_RED_AMAX)
    NRS_RED_CASE(NRS_RED_WNORM2) NRS_RED_CASE(NRS_RED_WDOT) NRS_RED_CASE(NRS_RED_DOT)
    NRS_RED_CASE(RED_WSUM)
    default: throw Error(NRS_ERR_INVALID, "linalg reduce: unknown reduction");
  }
#undef NRS_RED_CASE
  NRS_LAUNCH_CHECK();
}

void add_device_scalar(int64_t n, double factor, const double *scal, int slot, double *y, hipStream_t s)
{
  if (n <= 0) return;
  add_scalar_kernel<<<egrid(n), EB, 0, s>>>(n, factor, scal, slot, y);
  NRS_LAUNCH_CHECK();
}

void remove_mean(Reducer &R, int64_t n, const double *w, double wsum, double *q, double *scal, int slot,
                 hipStream_t s)
{
  if (n <= 0) return;
  linalg_reduce(w ? RED_WSUM : NRS_RED_SUM, n, 1, n, w, q, nullptr, R.partials.get(), scal + slot, s);
  add_device_scalar(n, -1.0 / wsum, scal, slot, q, s);
}

void block_update_pcg(int64_t n, int nf, int64_t off, const double *w, const double *Ap, double alpha,
                      double *r, double *partials, double *out, hipStream_t s)
{
  const int g = rgrid(n);
  block_update_pcg_kernel<<<g, RT, 0, s>>>(n, nf, off, w, Ap, alpha, r, partials);
  reduce_finish_kernel<NRS_RED_SUM><<<1, RT, 0, s>>>(g, partials, out);
  NRS_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ host side

void wdot(Reducer &R, int64_t n, const double *w, const double *a, const double *b, double *scal,
          int slot, hipStream_t s)
{
  const int g = rgrid(n);
  wdot_kernel<double><<<g, RT, 0, s>>>(n, w, a, b, R.partials.get());
  finish_kernel<1, FIN_STORE><<<1, RT, 0, s>>>(g, R.partials.get(), scal, slot, 1.0);
  NRS_LAUNCH_CHECK();
}
void wdot_f32(Reducer &R, int64_t n, const float *w, const float *a, const float *b, double *scal,
              int slot, hipStream_t s)
{
  const int g = rgrid(n);
  wdot_kernel<float><<<g, RT, 0, s>>>(n, w, a, b, R.partials.get());
  finish_kernel<1, FIN_STORE><<<1, RT, 0, s>>>(g, R.partials.get(), scal, slot, 1.0);
  NRS_LAUNCH_CHECK();
}
void wnorm(Reducer &R, int64_t n, const double *w, const double *a, double factor, double *scal,
           int slot, hipStream_t s)
{
  const int g = rgrid(n);
  wdot_kernel<double><<<g, RT, 0, s>>>(n, w, a, a, R.partials.get());
  finish_kernel<1, FIN_SQRT><<<1, RT, 0, s>>>(g, R.partials.get(), scal, slot, factor);
  NRS_LAUNCH_CHECK();
}
void plane_sum(Reducer &R, int64_t E, int Np, int64_t estride, int64_t off, const double *x,
               double *scal, int slot, hipStream_t s)
{
  const int g = rgrid(E * Np);
  plane_sum_kernel<<<g, RT, 0, s>>>(E, Np, estride, off, x, R.partials.get());
  finish_kernel<1, FIN_STORE><<<1, RT, 0, s>>>(g, R.partials.get(), scal, slot, 1.0);
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

void pcg_init(double *scal, double rdotr0, double tol, int maxit, hipStream_t s)
{
  pcg_init_kernel<<<1, 64, 0, s>>>(scal, rdotr0, tol, maxit);
  NRS_LAUNCH_CHECK();
}

void pcg_init_device(double *scal, double tol, bool relative, int maxit, hipStream_t s)
{
  pcg_init_dev_kernel<<<1, 64, 0, s>>>(scal, tol, relative ? 1 : 0, maxit);
  NRS_LAUNCH_CHECK();
}

void pcg_precon_dots(Reducer &R, int64_t n, const double *w, const double *invDiag, const double *r,
                     double *z, const double *Ap, bool jacobi, bool flexible, bool none_precon,
                     double *scal, hipStream_t s)
{
  const int g = rgrid(n);
  double *P = R.partials.get();
#define PD(MODE, FLEX)                                                                            \
  do {                                                                                            \
    precon_dots_kernel<MODE, FLEX><<<g, RT, 0, s>>>(n, w, invDiag, r, z, Ap, scal, P);            \
    finish_kernel<2, FIN_DOTS, MODE, FLEX><<<1, RT, 0, s>>>(g, P, scal, 0, 1.0);                  \
  } while (0)
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
  const int g = rgrid(n);
  pap_kernel<<<g, RT, 0, s>>>(n, w, p, Ap, scal, R.partials.get());
  finish_kernel<1, FIN_ALPHA><<<1, RT, 0, s>>>(g, R.partials.get(), scal, 0, 1.0);
  NRS_LAUNCH_CHECK();
}

void pcg_update_rx(Reducer &R, int64_t n, const double *w, const double *Ap, const double *p,
                   double *r, double *x, double resNormFactor, double *scal, hipStream_t s)
{
  const int g = rgrid(n);
  update_rx_kernel<<<g, RT, 0, s>>>(n, w, Ap, p, r, x, scal, R.partials.get());
  finish_kernel<1, FIN_RDOTR><<<1, RT, 0, s>>>(g, R.partials.get(), scal, 0, resNormFactor);
  NRS_LAUNCH_CHECK();
}

}  // namespace nrs
