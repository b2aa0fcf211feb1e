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
// passes would leave them.  omask ([E][ceil(NO^3/32)] bits, optional): output
// nodes stored as zero, the Dirichlet mask of the coarse level.  Its masked
// nodes are in no row of that level's gather-scatter, so zeroing them here
// equals masking after the gather-scatter.
template <typename T, bool ACCUM>
__global__ void __launch_bounds__(kTB)
tensor3_kernel(int E, int NI, int NO, const T *__restrict__ M, int so, int si,
               const T *__restrict__ in, T *__restrict__ out, const T *__restrict__ w,
               const T *__restrict__ sub, const uint32_t *__restrict__ omask)
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
    if (omask && ((omask[e * ((NPO + 31) / 32) + (t >> 5)] >> (t & 31)) & 1u)) acc = T(0);
    if (ACCUM) out[e * NPO + t] += acc;
    else out[e * NPO + t] = acc;
  }
}
// Order 7 <-> order 3 (NqF = 8, NqC = 4), one element per 64-lane wave, four
// per block; lane (i, j) of the fine plane.  The contractions run in the order
// and with the accumulation of tensor3_kernel (i, then j, then k, each sum in
// increasing index), so the results are the same bit for bit.
constexpr int kTW = 4;   // elements (waves) per block

// qf += (A x A x A) qc, A[o][c] = R[c][o] (R is [NqC][NqF])
template <typename T>
__global__ void __launch_bounds__(64 * kTW)
prolong_8_4_kernel(int E, const T *__restrict__ R, const T *__restrict__ qc, T *__restrict__ qf)
{
  __shared__ T s_c[kTW][64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int i = l & 7, j = l >> 3;
  const int64_t e = (int64_t)blockIdx.x * kTW + w;
  if (e >= E) return;   // whole waves only: the wave-level barrier below needs no other wave
  s_c[w][l] = qc[e * 64 + l];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  T Ai[4], Aj[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    Ai[c] = R[c * 8 + i];
    Aj[c] = R[c * 8 + j];
  }
  // t1[kc][jc] = sum_ic A[i][ic] qc[kc][jc][ic]
  T t1[4][4];
#pragma unroll
  for (int kc = 0; kc < 4; ++kc)
#pragma unroll
    for (int jc = 0; jc < 4; ++jc) {
      T acc = 0;
#pragma unroll
      for (int ic = 0; ic < 4; ++ic) acc += Ai[ic] * s_c[w][ic + 4 * (jc + 4 * kc)];
      t1[kc][jc] = acc;
    }
  // t2[kc] = sum_jc A[j][jc] t1[kc][jc]
  T t2[4];
#pragma unroll
  for (int kc = 0; kc < 4; ++kc) {
    T acc = 0;
#pragma unroll
    for (int jc = 0; jc < 4; ++jc) acc += Aj[jc] * t1[kc][jc];
    t2[kc] = acc;
  }
  // qf[k][j][i] += sum_kc A[k][kc] t2[kc]
  T *o = qf + e * 512 + l;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    T acc = 0;
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) acc += R[kc * 8 + k] * t2[kc];
    o[k * 64] += acc;
  }
}

// qc = (A x A x A) (w .* (sub - qf)) with A = R ([NqC][NqF]); w, sub and the
// coarse mask bits cmask ([E][2]) optional as for tensor3_kernel
template <typename T>
__global__ void __launch_bounds__(64 * kTW)
coarsen_8_4_kernel(int E, const T *__restrict__ R, const T *__restrict__ qf, T *__restrict__ qc,
                   const T *__restrict__ w, const T *__restrict__ sub, const uint32_t *__restrict__ cmask)
{
  __shared__ T s_f[kTW][512];    // the weighted fine element, [k][j][i]
  __shared__ T s_t1[kTW][256];   // [k][j][io]
  __shared__ T s_t2[kTW][128];   // [k][jo][io]
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * kTW + wv;
  if (e >= E) return;   // whole waves only
  const int64_t b = e * 512 + l;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    T v;
    if (w && sub) {
      const T r = sub[b + k * 64] - qf[b + k * 64];
      v = w[b + k * 64] * r;
    } else if (w) {
      v = w[b + k * 64] * qf[b + k * 64];
    } else {
      v = qf[b + k * 64];
    }
    s_f[wv][k * 64 + l] = v;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // t1[k][j][io] = sum_i R[io][i] f[k][j][i]: lane (j, k) = (l & 7, l >> 3), io = 0..3
  {
    const int jj = l & 7, kk = l >> 3;
    T f[8];
#pragma unroll
    for (int ii = 0; ii < 8; ++ii) f[ii] = s_f[wv][ii + 8 * (jj + 8 * kk)];
#pragma unroll
    for (int io = 0; io < 4; ++io) {
      T acc = 0;
#pragma unroll
      for (int ii = 0; ii < 8; ++ii) acc += R[io * 8 + ii] * f[ii];
      s_t1[wv][io + 4 * (jj + 8 * kk)] = acc;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // t2[k][jo][io] = sum_j R[jo][j] t1[k][j][io]: 128 values, two per lane
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = l + 64 * h;
    const int io = t & 3, jo = (t >> 2) & 3, kk = t >> 4;
    T acc = 0;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) acc += R[jo * 8 + jj] * s_t1[wv][io + 4 * (jj + 8 * kk)];
    s_t2[wv][io + 4 * (jo + 4 * kk)] = acc;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // qc[ko][jo][io] = sum_k R[ko][k] t2[k][jo][io]: one value per lane
  {
    const int io = l & 3, jo = (l >> 2) & 3, ko = l >> 4;
    T acc = 0;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) acc += R[ko * 8 + kk] * s_t2[wv][io + 4 * (jo + 4 * kk)];
    if (cmask && ((cmask[e * 2 + (l >> 5)] >> (l & 31)) & 1u)) acc = T(0);
    qc[e * 64 + l] = acc;
  }
}
}  // namespace

template <typename T>
void launch_coarsen(int Nf, int Nc, int E, const T *R, const T *qf, T *qc, hipStream_t s, const T *w,
                    const T *sub, const uint32_t *cmask)
{
  NRS_CHECK(!sub || w, "coarsen: a subtrahend needs the weight");
  NRS_CHECK(Nf + 1 <= kMaxNq && Nc >= 1 && Nc < Nf, "coarsen: need 1 <= Nc < Nf <= 9");
  if (E <= 0) return;
  const int NqF = Nf + 1, NqC = Nc + 1;
  if (Nf == 7 && Nc == 3) {
    coarsen_8_4_kernel<T><<<(unsigned)((E + kTW - 1) / kTW), 64 * kTW, 0, s>>>(E, R, qf, qc, w, sub, cmask);
    NRS_LAUNCH_CHECK();
    return;
  }
  // A[o][i] = R[o][i] = R[o*NqF + i]
  tensor3_kernel<T, false><<<E, kTB, 0, s>>>(E, NqF, NqC, R, NqF, 1, qf, qc, w, sub, cmask);
  NRS_LAUNCH_CHECK();
}

template <typename T>
void launch_prolongate(int Nf, int Nc, int E, const T *R, const T *qc, T *qf, hipStream_t s)
{
  NRS_CHECK(Nf + 1 <= kMaxNq && Nc >= 1 && Nc < Nf, "prolongate: need 1 <= Nc < Nf <= 9");
  if (E <= 0) return;
  const int NqF = Nf + 1, NqC = Nc + 1;
  if (Nf == 7 && Nc == 3) {
    prolong_8_4_kernel<T><<<(unsigned)((E + kTW - 1) / kTW), 64 * kTW, 0, s>>>(E, R, qc, qf);
    NRS_LAUNCH_CHECK();
    return;
  }
  // A[o][i] = R^T[o][i] = R[i][o] = R[i*NqF + o]
  tensor3_kernel<T, true><<<E, kTB, 0, s>>>(E, NqC, NqF, R, 1, NqF, qc, qf, nullptr, nullptr, nullptr);
  NRS_LAUNCH_CHECK();
}

template void launch_coarsen<double>(int, int, int, const double *, const double *, double *, hipStream_t,
                                     const double *, const double *, const uint32_t *);
template void launch_coarsen<float>(int, int, int, const float *, const float *, float *, hipStream_t,
                                    const float *, const float *, const uint32_t *);
template void launch_prolongate<double>(int, int, int, const double *, const double *, double *, hipStream_t);
template void launch_prolongate<float>(int, int, int, const float *, const float *, float *, hipStream_t);

}  // namespace nrs
