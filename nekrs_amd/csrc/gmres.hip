// gmres.hip -- the vector kernels of (flexible) restarted PGMRES.
//
// Restate, in one pass over the vectors each, the reference's
//   weightedInnerProdMulti          y_j = <V_j, w>_W, j < k   (src/linAlg/linAlg.cpp)
//   gramSchmidtOrthogonalization    w -= sum_j y_j V_j, then <w, w>_W
//                                   (kernels/elliptic/gramSchmidtOrthogonalization.okl)
//   updatePGMRESSolution            x += sum_j y_j Z_j (updatePGMRESSolution.okl)
//   fusedResidualAndNorm            r = b - Ax, <r, r>_W (fusedResidualAndNorm.okl)
// The k sums of a pass are held in registers (k is a template parameter up to
// 32), each block writes its partials, and a one-block finisher adds them in
// a fixed order into device memory: no atomics, deterministic results.
#include "gmres.hpp"

namespace nrs {

namespace {

constexpr int GT = 256;      // threads per block
constexpr int GB = 4096;     // max blocks of a partial-sum pass

unsigned ggrid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + GT - 1) / GT, GB)); }

// block sum of K values; result valid in thread 0
template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K])
{
  __shared__ double red[K][GT / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    double x = v[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[j][wv] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      double s = 0.0;
      for (int q = 0; q < GT / 64; ++q) s += red[j][q];
      v[j] = s;
    }
  }
}

template <int K>
__global__ void __launch_bounds__(GT)
multi_dot_kernel(int64_t n, int k, int64_t ld, const double *__restrict__ w, const double *__restrict__ V,
                 const double *__restrict__ x, double *__restrict__ partials)
{
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)GT + threadIdx.x; i < n; i += (int64_t)gridDim.x * GT) {
    const double wx = w[i] * x[i];
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (j < k) acc[j] += wx * V[j * ld + i];
  }
  block_sums<K>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (j < k) partials[j * gridDim.x + blockIdx.x] = acc[j];
  }
}

template <int K>
__global__ void __launch_bounds__(GT)
gram_schmidt_kernel(int64_t n, int k, int64_t ld, const double *__restrict__ w, const double *__restrict__ y,
                    const double *__restrict__ V, double *__restrict__ x, double *__restrict__ partials)
{
  double yj[K];
#pragma unroll
  for (int j = 0; j < K; ++j) yj[j] = (j < k) ? y[j] : 0.0;
  double acc[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)GT + threadIdx.x; i < n; i += (int64_t)gridDim.x * GT) {
    double xi = x[i];
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (j < k) xi -= yj[j] * V[j * ld + i];
    x[i] = xi;
    acc[0] += xi * xi * w[i];
  }
  block_sums<1>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

template <int K>
__global__ void update_kernel(int64_t n, int k, int64_t ld, const double *__restrict__ y,
                              const double *__restrict__ Z, double *__restrict__ x)
{
  double yj[K];
#pragma unroll
  for (int j = 0; j < K; ++j) yj[j] = (j < k) ? y[j] : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)GT + threadIdx.x; i < n; i += (int64_t)gridDim.x * GT) {
    double xi = x[i];
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (j < k) xi += yj[j] * Z[j * ld + i];
    x[i] = xi;
  }
}

__global__ void __launch_bounds__(GT)
residual_kernel(int64_t n, const double *__restrict__ w, const double *__restrict__ b,
                const double *__restrict__ Ax, double *__restrict__ r, double *__restrict__ partials)
{
  double acc[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)GT + threadIdx.x; i < n; i += (int64_t)gridDim.x * GT) {
    const double ri = b[i] - Ax[i];
    r[i] = ri;
    acc[0] += ri * ri * w[i];
  }
  block_sums<1>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

// out[j] = sum_b partials[j * nb + b], j < k
__global__ void __launch_bounds__(GT)
finish_multi_kernel(int nb, const double *__restrict__ partials, double *__restrict__ out)
{
  // one block per sum (blockIdx.x = j), partials added in a fixed order
  const int j = blockIdx.x;
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += GT) v[0] += partials[j * nb + b];
  block_sums<1>(v);
  if (threadIdx.x == 0) out[j] = v[0];
}

template <template <int> class F, typename... A>
void dispatch_k(int k, A... a)
{
  switch ((k + 3) / 4) {
    case 1: F<4>::run(a...); break;
    case 2: F<8>::run(a...); break;
    case 3: F<12>::run(a...); break;
    case 4: F<16>::run(a...); break;
    case 5: F<20>::run(a...); break;
    case 6: F<24>::run(a...); break;
    case 7: F<28>::run(a...); break;
    case 8: F<32>::run(a...); break;
    default: throw Error(NRS_ERR_INVALID, "gmres: at most 32 vectors");
  }
}

template <int K> struct MultiDot {
  static void run(int64_t n, int k, int64_t ld, const double *w, const double *V, const double *x,
                  double *partials, double *y, hipStream_t s)
  {
    const unsigned g = ggrid(n);
    multi_dot_kernel<K><<<g, GT, 0, s>>>(n, k, ld, w, V, x, partials);
    finish_multi_kernel<<<k, GT, 0, s>>>((int)g, partials, y);
  }
};
template <int K> struct GramSchmidt {
  static void run(int64_t n, int k, int64_t ld, const double *w, const double *y, const double *V,
                  double *x, double *partials, double *nw, hipStream_t s)
  {
    const unsigned g = ggrid(n);
    gram_schmidt_kernel<K><<<g, GT, 0, s>>>(n, k, ld, w, y, V, x, partials);
    finish_multi_kernel<<<1, GT, 0, s>>>((int)g, partials, nw);
  }
};
template <int K> struct Update {
  static void run(int64_t n, int k, int64_t ld, const double *y, const double *Z, double *x, hipStream_t s)
  {
    update_kernel<K><<<ggrid(n), GT, 0, s>>>(n, k, ld, y, Z, x);
  }
};

}  // namespace

void gmres_multi_dot(int64_t n, int k, int64_t ld, const double *w, const double *V, const double *x,
                     double *partials, double *y, hipStream_t s)
{
  dispatch_k<MultiDot>(k, n, k, ld, w, V, x, partials, y, s);
  NRS_LAUNCH_CHECK();
}

void gmres_gram_schmidt(int64_t n, int k, int64_t ld, const double *w, const double *y, const double *V,
                        double *x, double *partials, double *nw, hipStream_t s)
{
  dispatch_k<GramSchmidt>(k, n, k, ld, w, y, V, x, partials, nw, s);
  NRS_LAUNCH_CHECK();
}

void gmres_update(int64_t n, int k, int64_t ld, const double *y, const double *Z, double *x, hipStream_t s)
{
  dispatch_k<Update>(k, n, k, ld, y, Z, x, s);
  NRS_LAUNCH_CHECK();
}

void gmres_residual(int64_t n, const double *w, const double *b, const double *Ax, double *r,
                    double *partials, double *nr, hipStream_t s)
{
  const unsigned g = ggrid(n);
  residual_kernel<<<g, GT, 0, s>>>(n, w, b, Ax, r, partials);
  finish_multi_kernel<<<1, GT, 0, s>>>((int)g, partials, nr);
  NRS_LAUNCH_CHECK();
}

}  // namespace nrs
