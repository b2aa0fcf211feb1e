// transfer.hip -- p-multigrid intergrid transfers as sum-factorised tensor contractions.
//
// coarsen    qc  = (R x R x R) qf           kernels/elliptic/ellipticPreconCoarsenHex3D.c:26-100
// prolongate qf += (R^T x R^T x R^T) qc     kernels/elliptic/ellipticPreconProlongateHex3D.c:27-114
// R is [NqC][NqF].  One block per element: the element is staged in LDS and the
// three 1D contractions (i, then j, then k) run out of LDS; accumulation in the
// element type (FP32 on the multigrid levels, as the reference's pfloat kernels).
#include "common.hpp"
#include "kernels.hpp"

namespace nrs {

namespace {
constexpr int kMaxNq = 10;
constexpr int kTB = 256;

// out = (A x A x A) (w .* (sub - in)), A is [NO][NI] given through
// (a_stride_o, a_stride_i) into M; w = null: no weight; sub = null: in itself.
// The difference and the weighted input are rounded on their own, as separate
// passes would leave them.
template <typename T, bool ACCUM>
__global__ void __launch_bounds__(kTB)
tensor3_kernel(int E, int NI, int NO, const T *__restrict__ M, int so, int si,
               const T *__restrict__ in, T *__restrict__ out, const T *__restrict__ w,
               const T *__restrict__ sub)
{
  __shared__ T s_A[kMaxNq * kMaxNq];
  __shared__ T s_in[kMaxNq * kMaxNq * kMaxNq];
  __shared__ T s_t1[kMaxNq * kMaxNq * kMaxNq];
  __shared__ T s_t2[kMaxNq * kMaxNq * kMaxNq];
  const int64_t e = blockIdx.x;
  const int NPI = NI * NI * NI, NPO = NO * NO * NO;
  for (int t = threadIdx.x; t < NO * NI; t += kTB) {
    const int o = t / NI, i = t % NI;
    s_A[o * NI + i] = M[o * so + i * si];
  }
  if (w && sub) {
    for (int t = threadIdx.x; t < NPI; t += kTB) {
      const T r = sub[e * NPI + t] - in[e * NPI + t];
      s_in[t] = w[e * NPI + t] * r;
    }
  } else if (w) {
    for (int t = threadIdx.x; t < NPI; t += kTB) s_in[t] = w[e * NPI + t] * in[e * NPI + t];
  } else {
    for (int t = threadIdx.x; t < NPI; t += kTB) s_in[t] = in[e * NPI + t];
  }
  __syncthreads();
  // contract i: t1[k][j][io] over NI*NI*NO
  for (int t = threadIdx.x; t < NI * NI * NO; t += kTB) {
    const int io = t % NO, j = (t / NO) % NI, k = t / (NO * NI);
    T acc = 0;
    for (int i = 0; i < NI; ++i) acc += s_A[io * NI + i] * s_in[i + NI * (j + NI * k)];
    s_t1[io + NO * (j + NI * k)] = acc;
  }
  __syncthreads();
  // contract j: t2[k][jo][io] over NI*NO*NO
  for (int t = threadIdx.x; t < NI * NO * NO; t += kTB) {
    const int io = t % NO, jo = (t / NO) % NO, k = t / (NO * NO);
    T acc = 0;
    for (int j = 0; j < NI; ++j) acc += s_A[jo * NI + j] * s_t1[io + NO * (j + NI * k)];
    s_t2[io + NO * (jo + NO * k)] = acc;
  }
  __syncthreads();
  // contract k
  for (int t = threadIdx.x; t < NPO; t += kTB) {
    const int io = t % NO, jo = (t / NO) % NO, ko = t / (NO * NO);
    T acc = 0;
    for (int k = 0; k < NI; ++k) acc += s_A[ko * NI + k] * s_t2[io + NO * (jo + NO * k)];
    if (ACCUM) out[e * NPO + t] += acc;
    else out[e * NPO + t] = acc;
  }
}
}  // namespace

template <typename T>
void launch_coarsen(int Nf, int Nc, int E, const T *R, const T *qf, T *qc, hipStream_t s, const T *w,
                    const T *sub)
{
  NRS_CHECK(!sub || w, "coarsen: a subtrahend needs the weight");
  NRS_CHECK(Nf + 1 <= kMaxNq && Nc >= 1 && Nc < Nf, "coarsen: need 1 <= Nc < Nf <= 9");
  if (E <= 0) return;
  const int NqF = Nf + 1, NqC = Nc + 1;
  // A[o][i] = R[o][i] = R[o*NqF + i]
  tensor3_kernel<T, false><<<E, kTB, 0, s>>>(E, NqF, NqC, R, NqF, 1, qf, qc, w, sub);
  NRS_LAUNCH_CHECK();
}

template <typename T>
void launch_prolongate(int Nf, int Nc, int E, const T *R, const T *qc, T *qf, hipStream_t s)
{
  NRS_CHECK(Nf + 1 <= kMaxNq && Nc >= 1 && Nc < Nf, "prolongate: need 1 <= Nc < Nf <= 9");
  if (E <= 0) return;
  const int NqF = Nf + 1, NqC = Nc + 1;
  // A[o][i] = R^T[o][i] = R[i][o] = R[i*NqF + o]
  tensor3_kernel<T, true><<<E, kTB, 0, s>>>(E, NqC, NqF, R, 1, NqF, qc, qf, nullptr, nullptr);
  NRS_LAUNCH_CHECK();
}

template void launch_coarsen<double>(int, int, int, const double *, const double *, double *, hipStream_t,
                                     const double *, const double *);
template void launch_coarsen<float>(int, int, int, const float *, const float *, float *, hipStream_t,
                                    const float *, const float *);
template void launch_prolongate<double>(int, int, int, const double *, const double *, double *, hipStream_t);
template void launch_prolongate<float>(int, int, int, const float *, const float *, float *, hipStream_t);

}  // namespace nrs
