// ax.hip -- element stiffness apply (Poisson / Helmholtz), its diagonal, the
// three-field block apply, and the Dirichlet mask, for gfx950.
//
// Design (one element = one (N+1)^2 slab of lanes, N=7: exactly one wave64):
//   * lane (i,j) keeps the k-pencil q(i,j,0..N) and the output pencil Aq(i,j,:)
//     in registers, so the t-direction contractions never touch LDS;
//   * the r/s contractions of plane k read the plane from LDS (one 8x8 slab
//     of doubles = 512 B), D rows/columns live in registers;
//   * ggeo streams as 6 coalesced runs of (N+1)^2 values per plane;
//   * the Dirichlet mask is fused into the store epilogue via a per-element
//     bit mask (Np bits), so the operator needs no separate mask launch.
// Small orders pack several elements per block so that a block is >= 64 lanes.
#include "common.hpp"
#include "kernels.hpp"
#include "trilin.hpp"
#include "tune.hpp"

namespace nrs {

namespace {
constexpr int kG00 = 0, kG01 = 1, kG11 = 2, kG12 = 3, kG02 = 4, kG22 = 5, kGWJ = 6;
}

#ifndef AX_WAVES_PER_EU
#define AX_WAVES_PER_EU 3
#endif

// Geometric factors are read exactly once per apply: load them non-temporally so
// they do not evict q/Aq lines that the gather-scatter reuses.
template <typename T> __device__ __forceinline__ T stream_load(const T *p)
{
  return __builtin_nontemporal_load(p);
}

// Synchronisation between the stages of one plane.  When an element is exactly
// one wave64 (NQ = 8) its LDS slabs are private to that wave, and a wave's LDS
// operations complete in issue order, so a wave-level barrier (compiler fence,
// as HIP's __syncwarp) is enough; other orders share or split waves and use
// the workgroup barrier.
template <int NQ> __device__ __forceinline__ void plane_sync()
{
  if constexpr (NQ * NQ == 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <int NQ, int MULT = 1> struct AxShape {
  // elements per block: fill at least one wave64, times the tuned multiplier
  static constexpr int NE = ((NQ * NQ >= 64) ? 1 : (64 + NQ * NQ - 1) / (NQ * NQ)) * MULT;
  static constexpr int THREADS = NQ * NQ * NE;
};

// -------------------------------------------------------------------------------------
// Single-field apply.  Restates ellipticPartialAxCoeffHex3D
// (kernels/elliptic/ellipticPartialAxCoeffHex3D.c:2-97) with the same arithmetic:
//   qr = sum_m D[i][m] q(m,j,k), qs = sum_m D[j][m] q(i,m,k), qt = sum_m D[k][m] q(i,j,m)
//   [Gqr,Gqs,Gqt] = lambda0 * G * [qr,qs,qt]
//   Aq = sum_m D[m][i] Gqr(m,j,k) + D[m][j] Gqs(i,m,k) + D[m][k] Gqt(i,j,m) (+ GwJ lambda1 q)
// -------------------------------------------------------------------------------------
//
// FUSEX: the block's consecutive elements e, e+1 with fx[e] = 1 share their
// face i = N / i = 0 node for node; the interior nodes of that face are summed
// here through LDS (both copies get v_e + v_e+1, the gather-scatter's pair sum,
// bit for bit) and the caller's gather-scatter handle leaves them out.
//
// TRI: trilinear elements.  The geometric factors are not read but formed at
// each node from the element's eight corners (corners [E][3][8]) and the GLL
// nodes and weights (zw [2][NQ]) by TriGeo (trilin.hpp).
template <typename T, int NQ, int MULT, bool POISSON, bool LFIELD, bool FUSEX = false, bool TRI = false>
__global__ void __launch_bounds__((AxShape<NQ, MULT>::THREADS), AX_WAVES_PER_EU)
ax_kernel(int E, const int *__restrict__ elist, const T *__restrict__ ggeo,
          const T *__restrict__ Dg, const T *__restrict__ lam0, const T *__restrict__ lam1,
          const T *__restrict__ q, T *__restrict__ Aq, const uint32_t *__restrict__ maskBits,
          const uint8_t *__restrict__ fx = nullptr, const T *__restrict__ corners = nullptr,
          const T *__restrict__ zw = nullptr)
{
  constexpr int NE = AxShape<NQ, MULT>::NE;
  constexpr int NP = NQ * NQ * NQ;
  constexpr int NW = (NP + 31) / 32;
  __shared__ T s_D[NQ][NQ];
  __shared__ T s_q[NE][NQ][NQ];
  __shared__ T s_Gr[NE][NQ][NQ];
  __shared__ T s_Gs[NE][NQ][NQ];

  const int t = threadIdx.x;
  const int i = t % NQ;
  const int j = (t / NQ) % NQ;
  const int le = (NE == 1) ? 0 : t / (NQ * NQ);
  const int eidx = blockIdx.x * NE + le;
  const bool active = eidx < E;
  int64_t e = active ? (elist ? elist[eidx] : eidx) : 0;
  if (NE == 1 || NQ * NQ == 64) e = __builtin_amdgcn_readfirstlane((int)e);   // wave-uniform element

  if (t < NQ * NQ) s_D[t / NQ][t % NQ] = Dg[t];
  __syncthreads();

  T Di[NQ], Dj[NQ], DTi[NQ], DTj[NQ];
#pragma unroll
  for (int m = 0; m < NQ; ++m) {
    Di[m] = s_D[i][m];
    Dj[m] = s_D[j][m];
    DTi[m] = s_D[m][i];
    DTj[m] = s_D[m][j];
  }

  const int lid = j * NQ + i;          // lane offset inside a plane (32-bit)
  const T *qe = q + e * NP;
  const T *G = ggeo + e * 7 * NP;
  T r_q[NQ], r_Aq[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    r_q[k] = active ? qe[k * NQ * NQ + lid] : T(0);
    r_Aq[k] = T(0);
  }

  const T l0c = LFIELD ? T(0) : lam0[0];
  const T l1c = (LFIELD || POISSON) ? T(0) : lam1[0];

  // geometric factors of plane 0; later planes are prefetched one plane ahead
  // (TRI: the corner differences that the derivatives of the map take at this
  // lane's (r, s), formed once per element)
  T g[7];
  TriGeo<T> tri;
  if constexpr (TRI) {
    tri.init(corners + e * 24, zw[i], zw[j], zw[NQ + i], zw[NQ + j]);
  } else {
#pragma unroll
    for (int c = 0; c < 7; ++c) g[c] = (active && (c < 6 || !POISSON)) ? stream_load(G + c * NP + lid) : T(0);
  }

#pragma unroll 1
  for (int k = 0; k < NQ; ++k) {
    const int n = k * NQ * NQ;
    T qk = r_q[0];
#pragma unroll
    for (int m = 1; m < NQ; ++m) qk = (k == m) ? r_q[m] : qk;
    T gn[7];
    if constexpr (TRI) {
      T v[10];
      tri.at(zw[k], zw[NQ + k], v);
      const T rx = v[0], ry = v[1], rz = v[2], sx = v[3], sy = v[4], sz = v[5], tx = v[6], ty = v[7], tz = v[8];
      const T JW = v[9];
      g[kG00] = JW * (rx * rx + ry * ry + rz * rz);
      g[kG01] = JW * (rx * sx + ry * sy + rz * sz);
      g[kG02] = JW * (rx * tx + ry * ty + rz * tz);
      g[kG11] = JW * (sx * sx + sy * sy + sz * sz);
      g[kG12] = JW * (sx * tx + sy * ty + sz * tz);
      g[kG22] = JW * (tx * tx + ty * ty + tz * tz);
      g[kGWJ] = JW;
    } else {
      const bool more = active && (k + 1 < NQ);
#pragma unroll
      for (int c = 0; c < 7; ++c)
        gn[c] = (more && (c < 6 || !POISSON)) ? stream_load(G + c * NP + n + NQ * NQ + lid) : T(0);
    }
    T l0 = l0c, l1 = l1c;
    if (LFIELD && active) {
      l0 = lam0[e * NP + n + lid];
      if (!POISSON) l1 = lam1[e * NP + n + lid];
    }
    const T G00 = g[kG00], G01 = g[kG01], G02 = g[kG02], G11 = g[kG11], G12 = g[kG12], G22 = g[kG22];

    s_q[le][j][i] = qk;
    T qt = 0;
#pragma unroll
    for (int m = 0; m < NQ; ++m) qt += Dg[k * NQ + m] * r_q[m];
    plane_sync<NQ>();

    T qr = 0, qs = 0;
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      qr += Di[m] * s_q[le][j][m];
      qs += Dj[m] * s_q[le][m][i];
    }
    T Gqr = G00 * qr; Gqr += G01 * qs; Gqr += G02 * qt;
    T Gqs = G01 * qr; Gqs += G11 * qs; Gqs += G12 * qt;
    T Gqt = G02 * qr; Gqt += G12 * qs; Gqt += G22 * qt;
    s_Gr[le][j][i] = l0 * Gqr;
    s_Gs[le][j][i] = l0 * Gqs;
    Gqt *= l0;
    plane_sync<NQ>();

    T a = 0;
#pragma unroll
    for (int m = 0; m < NQ; ++m) a += DTi[m] * s_Gr[le][j][m];
    T b = 0;
#pragma unroll
    for (int m = 0; m < NQ; ++m) b += DTj[m] * s_Gs[le][m][i];
    T ab = a + b;
    if (!POISSON) ab += g[kGWJ] * l1 * qk;
#pragma unroll
    for (int m = 0; m < NQ; ++m) r_Aq[m] += Dg[k * NQ + m] * Gqt;
#pragma unroll
    for (int m = 0; m < NQ; ++m) r_Aq[m] += (k == m) ? ab : T(0);
    if constexpr (!TRI) {
#pragma unroll
      for (int c = 0; c < 7; ++c) g[c] = gn[c];
    }
  }

  if constexpr (FUSEX) {
    static_assert(NE >= 2, "face fusion needs at least two elements per block");
    __shared__ T s_x[NE][2][NQ][NQ];     // [element][side: i = 0, i = N][j][k]
    const bool face = (i == 0 || i == NQ - 1) && j > 0 && j < NQ - 1;
    if (face) {
      const int side = (i == 0) ? 0 : 1;
#pragma unroll
      for (int k = 1; k < NQ - 1; ++k) s_x[le][side][j][k] = r_Aq[k];
    }
    __syncthreads();
    if (active && face) {
      if (i == NQ - 1 && le + 1 < NE && eidx + 1 < E && fx[eidx]) {
#pragma unroll
        for (int k = 1; k < NQ - 1; ++k) r_Aq[k] = r_Aq[k] + s_x[le + 1][0][j][k];
      } else if (i == 0 && le >= 1 && fx[eidx - 1]) {
#pragma unroll
        for (int k = 1; k < NQ - 1; ++k) r_Aq[k] = s_x[le - 1][1][j][k] + r_Aq[k];
      }
    }
  }
  if (!active) return;
  T *Ae = Aq + e * NP + lid;
  if (maskBits) {
    const uint32_t *mb = maskBits + e * NW;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const int n = k * NQ * NQ + j * NQ + i;
      const bool masked = (mb[n >> 5] >> (n & 31)) & 1u;
      Ae[k * NQ * NQ] = masked ? T(0) : r_Aq[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < NQ; ++k) Ae[k * NQ * NQ] = r_Aq[k];
  }
}

template <typename T, int NQ, int MULT>
static void launch_ax_nqm(int E, const int *elist, const T *ggeo, const T *D, const T *l0,
                          const T *l1, bool poisson, bool lfield, const T *q, T *Aq,
                          const uint32_t *maskBits, hipStream_t s)
{
  using S = AxShape<NQ, MULT>;
  const unsigned grid = (unsigned)((E + S::NE - 1) / S::NE);
  if (grid == 0) return;
  if (poisson && !lfield)
    ax_kernel<T, NQ, MULT, true, false><<<grid, S::THREADS, 0, s>>>(E, elist, ggeo, D, l0, l1, q, Aq, maskBits);
  else if (poisson && lfield)
    ax_kernel<T, NQ, MULT, true, true><<<grid, S::THREADS, 0, s>>>(E, elist, ggeo, D, l0, l1, q, Aq, maskBits);
  else if (!poisson && !lfield)
    ax_kernel<T, NQ, MULT, false, false><<<grid, S::THREADS, 0, s>>>(E, elist, ggeo, D, l0, l1, q, Aq, maskBits);
  else
    ax_kernel<T, NQ, MULT, false, true><<<grid, S::THREADS, 0, s>>>(E, elist, ggeo, D, l0, l1, q, Aq, maskBits);
  NRS_LAUNCH_CHECK();
}

template <typename T, int NQ>
static void launch_ax_nq(int mult, int E, const int *elist, const T *ggeo, const T *D, const T *l0,
                         const T *l1, bool poisson, bool lfield, const T *q, T *Aq,
                         const uint32_t *maskBits, hipStream_t s)
{
  switch (mult) {
    case 2: launch_ax_nqm<T, NQ, 2>(E, elist, ggeo, D, l0, l1, poisson, lfield, q, Aq, maskBits, s); break;
    case 4: launch_ax_nqm<T, NQ, 4>(E, elist, ggeo, D, l0, l1, poisson, lfield, q, Aq, maskBits, s); break;
    default: launch_ax_nqm<T, NQ, 1>(E, elist, ggeo, D, l0, l1, poisson, lfield, q, Aq, maskBits, s); break;
  }
}

template <typename T>
void launch_ax_variant(int N, int mult, int E, const int *elist, const T *ggeo, const T *D,
                       const T *l0, const T *l1, bool poisson, bool lfield, const T *q, T *Aq,
                       const uint32_t *maskBits, hipStream_t s)
{
  switch (N + 1) {
#define NRS_AX_CASE(NQ_)                                                                          \
  case NQ_: launch_ax_nq<T, NQ_>(mult, E, elist, ggeo, D, l0, l1, poisson, lfield, q, Aq, maskBits, s); break;
    NRS_AX_CASE(2) NRS_AX_CASE(3) NRS_AX_CASE(4) NRS_AX_CASE(5) NRS_AX_CASE(6) NRS_AX_CASE(7)
    NRS_AX_CASE(8) NRS_AX_CASE(9) NRS_AX_CASE(10)
#undef NRS_AX_CASE
    default: throw Error(NRS_ERR_UNSUPPORTED, "Ax: order N must be 1..9");
  }
}

// Ax with the geometric factors formed from the element corners (trilinear
// elements); corners [E][24] relative to q, zw [2][N+1]
template <typename T, int NQ>
static void launch_ax_trilin_nq(int mult, int E, const T *corners, const T *zw, const T *D, const T *l0,
                                const T *l1, bool poisson, const T *q, T *Aq, const uint32_t *maskBits,
                                hipStream_t s)
{
#define NRS_AXT(M_)                                                                                   \
  {                                                                                                  \
    using S = AxShape<NQ, M_>;                                                                       \
    const unsigned grid = (unsigned)((E + S::NE - 1) / S::NE);                                       \
    if (grid == 0) return;                                                                           \
    if (poisson)                                                                                     \
      ax_kernel<T, NQ, M_, true, false, false, true><<<grid, S::THREADS, 0, s>>>(                     \
          E, nullptr, nullptr, D, l0, l1, q, Aq, maskBits, nullptr, corners, zw);                    \
    else                                                                                             \
      ax_kernel<T, NQ, M_, false, false, false, true><<<grid, S::THREADS, 0, s>>>(                    \
          E, nullptr, nullptr, D, l0, l1, q, Aq, maskBits, nullptr, corners, zw);                    \
  }
  switch (mult) {
    case 2: NRS_AXT(2) break;
    case 4: NRS_AXT(4) break;
    default: NRS_AXT(1) break;
  }
#undef NRS_AXT
  NRS_LAUNCH_CHECK();
}

template <typename T>
void launch_ax_trilin(int N, int E, const T *corners, const T *zw, const T *D, const T *l0, const T *l1,
                      bool poisson, const T *q, T *Aq, const uint32_t *maskBits, hipStream_t s)
{
  const int mult = tuned(TUNE_AX, N, dtype_of<T>::value, E, 1);
  switch (N + 1) {
#define NRS_AXT_CASE(NQ_)                                                                          \
  case NQ_: launch_ax_trilin_nq<T, NQ_>(mult, E, corners, zw, D, l0, l1, poisson, q, Aq, maskBits, s); break;
    NRS_AXT_CASE(2) NRS_AXT_CASE(3) NRS_AXT_CASE(4) NRS_AXT_CASE(5) NRS_AXT_CASE(6) NRS_AXT_CASE(7)
    NRS_AXT_CASE(8) NRS_AXT_CASE(9) NRS_AXT_CASE(10)
#undef NRS_AXT_CASE
    default: throw Error(NRS_ERR_UNSUPPORTED, "Ax: order N must be 1..9");
  }
}
template void launch_ax_trilin<double>(int, int, const double *, const double *, const double *,
                                       const double *, const double *, bool, const double *, double *,
                                       const uint32_t *, hipStream_t);
template void launch_ax_trilin<float>(int, int, const float *, const float *, const float *, const float *,
                                      const float *, bool, const float *, float *, const uint32_t *,
                                      hipStream_t);

// Ax with the face fusion (N = 7, M elements per block); fx [E] relative to q
template <typename T, int M>
static void launch_ax_fused_m(int E, const T *ggeo, const T *D, const T *l0, const T *l1, bool poisson,
                              const T *q, T *Aq, const uint32_t *maskBits, const uint8_t *fx, hipStream_t s)
{
  using S = AxShape<8, M>;
  static_assert(S::NE == M, "one element per wave at N = 7");
  const unsigned grid = (unsigned)((E + S::NE - 1) / S::NE);
  if (grid == 0) return;
  if (poisson)
    ax_kernel<T, 8, M, true, false, true><<<grid, S::THREADS, 0, s>>>(E, nullptr, ggeo, D, l0, l1, q, Aq,
                                                                      maskBits, fx);
  else
    ax_kernel<T, 8, M, false, false, true><<<grid, S::THREADS, 0, s>>>(E, nullptr, ggeo, D, l0, l1, q, Aq,
                                                                       maskBits, fx);
  NRS_LAUNCH_CHECK();
}

// nb: elements per block, the block the face-fusion flags were made for (fx[e] = 0 for e % nb = nb - 1)
template <typename T>
void launch_ax_fused(int N, int E, const T *ggeo, const T *D, const T *l0, const T *l1, bool poisson,
                     const T *q, T *Aq, const uint32_t *maskBits, const uint8_t *fx, hipStream_t s, int nb)
{
  NRS_CHECK(N == 7 && fx != nullptr, "Ax face fusion: order 7 only");
  if (nb == 2)
    launch_ax_fused_m<T, 2>(E, ggeo, D, l0, l1, poisson, q, Aq, maskBits, fx, s);
  else if (nb == 4)
    launch_ax_fused_m<T, 4>(E, ggeo, D, l0, l1, poisson, q, Aq, maskBits, fx, s);
  else
    throw Error(NRS_ERR_INVALID, "Ax face fusion: blocks of 2 or 4 elements");
}
template void launch_ax_fused<double>(int, int, const double *, const double *, const double *,
                                      const double *, bool, const double *, double *, const uint32_t *,
                                      const uint8_t *, hipStream_t, int);
template void launch_ax_fused<float>(int, int, const float *, const float *, const float *,
                                     const float *, bool, const float *, float *, const uint32_t *,
                                     const uint8_t *, hipStream_t, int);

// elements-per-block multiplier: the autotuner's choice for (N, type), else 1
template <typename T>
void launch_ax(int N, int E, const int *elist, const T *ggeo, const T *D, const T *l0, const T *l1,
               bool poisson, bool lfield, const T *q, T *Aq, const uint32_t *maskBits,
               hipStream_t s)
{
  const int mult = tuned(TUNE_AX, N, dtype_of<T>::value, E, 1);
  launch_ax_variant<T>(N, mult, E, elist, ggeo, D, l0, l1, poisson, lfield, q, Aq, maskBits, s);
}
template void launch_ax_variant<double>(int, int, int, const int *, const double *, const double *,
                                        const double *, const double *, bool, bool, const double *,
                                        double *, const uint32_t *, hipStream_t);
template void launch_ax_variant<float>(int, int, int, const int *, const float *, const float *,
                                       const float *, const float *, bool, bool, const float *,
                                       float *, const uint32_t *, hipStream_t);

template void launch_ax<double>(int, int, const int *, const double *, const double *, const double *,
                                const double *, bool, bool, const double *, double *,
                                const uint32_t *, hipStream_t);
template void launch_ax<float>(int, int, const int *, const float *, const float *, const float *,
                               const float *, bool, bool, const float *, float *, const uint32_t *,
                               hipStream_t);

// -------------------------------------------------------------------------------------
// Three-field Helmholtz apply with coefficient fields.  Restates
// ellipticBlockPartialAxCoeffHex3D (kernels/elliptic/ellipticBlockPartialAxCoeffHex3D.c:2-147);
// the geometric factors are read once for the three fields.
// -------------------------------------------------------------------------------------
// Same scheme as ax_kernel: one (i,j) column per lane, the k-pencils of q and
// of the result in registers with the k loop rolled (static register
// indexing through selects), r/s contractions through LDS slabs, geometric
// factors prefetched one plane ahead and shared by the three fields.
template <int NQ, bool LFIELD>
__global__ void __launch_bounds__(AxShape<NQ>::THREADS, LFIELD ? 1 : 2)
block_ax_kernel(int E, int64_t offset, int64_t loffset, const double *__restrict__ ggeo,
                const double *__restrict__ Dg, const double *__restrict__ lam0,
                const double *__restrict__ lam1, const double *__restrict__ q,
                double *__restrict__ Aq)
{
  constexpr int NE = AxShape<NQ>::NE;
  constexpr int NP = NQ * NQ * NQ;
  constexpr int NF = 3;
  __shared__ double s_D[NQ][NQ];
  __shared__ double s_q[NF][NE][NQ][NQ];
  __shared__ double s_Gr[NF][NE][NQ][NQ];
  __shared__ double s_Gs[NF][NE][NQ][NQ];
  const int t = threadIdx.x;
  const int i = t % NQ, j = (t / NQ) % NQ, le = t / (NQ * NQ);
  const int eidx = blockIdx.x * NE + le;
  const bool active = eidx < E;
  const int64_t e = active ? eidx : 0;
  if (t < NQ * NQ) s_D[t / NQ][t % NQ] = Dg[t];
  __syncthreads();
  const int lid = j * NQ + i;
  double r_q[NF][NQ], r_Aq[NF][NQ];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      r_q[f][k] = active ? q[f * offset + e * NP + k * NQ * NQ + lid] : 0.0;
      r_Aq[f][k] = 0.0;
    }
  const double *G = ggeo + e * 7 * NP + lid;
  double g[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) g[c] = active ? stream_load(G + c * NP) : 0.0;
  const double l0c = LFIELD ? 0.0 : lam0[0];
  const double l1c = LFIELD ? 0.0 : lam1[0];

#pragma unroll 1
  for (int k = 0; k < NQ; ++k) {
    const int n = k * NQ * NQ;
    const int64_t id = e * NP + n + lid;
    double gn[7];
    const bool more = active && (k + 1 < NQ);
#pragma unroll
    for (int c = 0; c < 7; ++c) gn[c] = more ? stream_load(G + c * NP + n + NQ * NQ) : 0.0;
    double qk[NF], qt[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      qk[f] = r_q[f][0];
#pragma unroll
      for (int m = 1; m < NQ; ++m) qk[f] = (k == m) ? r_q[f][m] : qk[f];
      s_q[f][le][j][i] = qk[f];
      double a = 0.0;
#pragma unroll
      for (int m = 0; m < NQ; ++m) a += s_D[k][m] * r_q[f][m];
      qt[f] = a;
    }
    __syncthreads();
    double Gqt[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double qr = 0.0, qs = 0.0;
#pragma unroll
      for (int m = 0; m < NQ; ++m) {
        qr += s_D[i][m] * s_q[f][le][j][m];
        qs += s_D[j][m] * s_q[f][le][m][i];
      }
      const double l0 = LFIELD ? (active ? lam0[f * loffset + id] : 0.0) : l0c;
      s_Gr[f][le][j][i] = l0 * (g[kG00] * qr + g[kG01] * qs + g[kG02] * qt[f]);
      s_Gs[f][le][j][i] = l0 * (g[kG01] * qr + g[kG11] * qs + g[kG12] * qt[f]);
      Gqt[f] = l0 * (g[kG02] * qr + g[kG12] * qs + g[kG22] * qt[f]);
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double l1 = LFIELD ? (active ? lam1[f * loffset + id] : 0.0) : l1c;
      double a = 0.0, b = 0.0;
#pragma unroll
      for (int m = 0; m < NQ; ++m) {
        a += s_D[m][i] * s_Gr[f][le][j][m];
        b += s_D[m][j] * s_Gs[f][le][m][i];
      }
      const double ab = a + b + g[kGWJ] * l1 * qk[f];
#pragma unroll
      for (int m = 0; m < NQ; ++m) r_Aq[f][m] += s_D[k][m] * Gqt[f];
#pragma unroll
      for (int m = 0; m < NQ; ++m) r_Aq[f][m] += (k == m) ? ab : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 7; ++c) g[c] = gn[c];
  }
  if (!active) return;
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int k = 0; k < NQ; ++k) Aq[f * offset + e * NP + k * NQ * NQ + lid] = r_Aq[f][k];
}

void launch_block_ax(int N, int E, int64_t offset, int64_t loffset, const double *ggeo,
                     const double *D, const double *l0, const double *l1, const double *q,
                     double *Aq, hipStream_t s, bool lfield)
{
#define BLK(NQ)                                                                              \
  case NQ: {                                                                                 \
    const unsigned g = (unsigned)((E + AxShape<NQ>::NE - 1) / AxShape<NQ>::NE);              \
    if (g && lfield)                                                                         \
      block_ax_kernel<NQ, true><<<g, AxShape<NQ>::THREADS, 0, s>>>(E, offset, loffset, ggeo, D, l0, l1, q, Aq); \
    else if (g)                                                                              \
      block_ax_kernel<NQ, false><<<g, AxShape<NQ>::THREADS, 0, s>>>(E, offset, loffset, ggeo, D, l0, l1, q, Aq); \
    break;                                                                                   \
  }
  switch (N + 1) {
    BLK(2) BLK(3) BLK(4) BLK(5) BLK(6) BLK(7) BLK(8) BLK(9) BLK(10)
    default: throw Error(NRS_ERR_UNSUPPORTED, "block Ax: order N must be 1..9");
  }
#undef BLK
  NRS_LAUNCH_CHECK();
}

// -------------------------------------------------------------------------------------
// Element diagonal.  Restates ellipticBlockBuildDiagonalHex3D
// (kernels/elliptic/ellipticBlockBuildDiagonalHex3D.okl:27-117) for one field:
//   diag = 2 l0 (G01 D_ii D_jj + G02 D_ii D_kk + G12 D_jj D_kk)
//        + sum_m [G00 l0](m,j,k) D_mi^2 + [G11 l0](i,m,k) D_mj^2 + [G22 l0](i,j,m) D_mk^2
//        (+ GwJ l1).  Setup-time kernel: one thread per node.
// -------------------------------------------------------------------------------------
template <typename T>
__global__ void build_diag_kernel(int N, int E, const T *__restrict__ ggeo, const T *__restrict__ D,
                                  const T *__restrict__ lam0, const T *__restrict__ lam1,
                                  int poisson, int lfield, T *__restrict__ diag)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gid >= (int64_t)E * Np) return;
  const int64_t e = gid / Np;
  const int n = (int)(gid - e * Np);
  const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
  const T *G = ggeo + e * 7 * Np;
  auto L0 = [&](int nn) { return lfield ? lam0[e * Np + nn] : lam0[0]; };
  const T l0 = L0(n);
  T r = 0;
  r += T(2) * G[kG01 * Np + n] * l0 * D[i * Nq + i] * D[j * Nq + j];
  r += T(2) * G[kG02 * Np + n] * l0 * D[i * Nq + i] * D[k * Nq + k];
  r += T(2) * G[kG12 * Np + n] * l0 * D[j * Nq + j] * D[k * Nq + k];
  for (int m = 0; m < Nq; ++m) {
    const int nr = m + Nq * (j + Nq * k), ns = i + Nq * (m + Nq * k), nt = i + Nq * (j + Nq * m);
    r += G[kG00 * Np + nr] * L0(nr) * D[m * Nq + i] * D[m * Nq + i];
    r += G[kG11 * Np + ns] * L0(ns) * D[m * Nq + j] * D[m * Nq + j];
    r += G[kG22 * Np + nt] * L0(nt) * D[m * Nq + k] * D[m * Nq + k];
  }
  if (!poisson) r += G[kGWJ * Np + n] * (lfield ? lam1[gid] : lam1[0]);
  diag[gid] = r;
}

template <typename T>
void launch_build_diag(int N, int E, const T *ggeo, const T *D, const T *l0, const T *l1,
                       bool poisson, bool lfield, T *diag, hipStream_t s)
{
  const int64_t n = (int64_t)E * (N + 1) * (N + 1) * (N + 1);
  if (n == 0) return;
  build_diag_kernel<T><<<grid_for(n, 256, 1ll << 30), 256, 0, s>>>(N, E, ggeo, D, l0, l1, poisson,
                                                                   lfield, diag);
  NRS_LAUNCH_CHECK();
}
template void launch_build_diag<double>(int, int, const double *, const double *, const double *,
                                        const double *, bool, bool, double *, hipStream_t);
template void launch_build_diag<float>(int, int, const float *, const float *, const float *,
                                       const float *, bool, bool, float *, hipStream_t);

// -------------------------------------------------------------------------------------
// mask (kernels/core/mask.okl:27-35)
// -------------------------------------------------------------------------------------
template <typename T>
__global__ void mask_kernel(int64_t Nm, const int *__restrict__ ids, T *__restrict__ v)
{
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < Nm;
       n += (int64_t)gridDim.x * blockDim.x)
    v[ids[n]] = T(0);
}

template <typename T>
void launch_mask(int64_t Nm, const int *ids, T *v, hipStream_t s)
{
  if (Nm <= 0) return;
  mask_kernel<T><<<grid_for(Nm, 256, 8192), 256, 0, s>>>(Nm, ids, v);
  NRS_LAUNCH_CHECK();
}
template void launch_mask<double>(int64_t, const int *, double *, hipStream_t);
template void launch_mask<float>(int64_t, const int *, float *, hipStream_t);

}  // namespace nrs
