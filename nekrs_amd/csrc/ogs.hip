// ogs.hip -- gather-scatter (QQ^T) on E-vectors as a bucketed segment reduction.
//
// Setup restates the local part of ogsSetup (3rd_party/gslib/ogs/src/ogsSetup.cpp:
// 111-246, 364-392): rows = groups of equal nonzero global ids, ordered by the
// local index of their first member, members ascending; invDegree = 1/count (1 at
// excluded nodes).  Unlike the reference, singleton rows are not stored at all
// and multi-member rows are bucketed by length (2, 4, 8, other) so that each
// thread reads its member indices with one aligned vector load and runs a
// fixed-trip loop.  The reduction order inside a row is the reference's
// (gatherScatterMany.okl:58-86: identity op m0 op m1 ...), so FP64 results match
// the reference kernel bit for bit.
#include <algorithm>
#include <numeric>

#include "ogs.hpp"

namespace nrs {

namespace {

template <typename T> struct OpId;
template <> struct OpId<double> {
  static __device__ double add() { return 0.0; }
  static __device__ double mul() { return 1.0; }
  static __device__ double mn() { return __builtin_huge_val(); }
  static __device__ double mx() { return -__builtin_huge_val(); }
};
template <> struct OpId<float> {
  static __device__ float add() { return 0.0f; }
  static __device__ float mul() { return 1.0f; }
  static __device__ float mn() { return __builtin_huge_valf(); }
  static __device__ float mx() { return -__builtin_huge_valf(); }
};
template <> struct OpId<int> {
  static __device__ int add() { return 0; }
  static __device__ int mul() { return 1; }
  static __device__ int mn() { return 2147483647; }
  static __device__ int mx() { return -2147483647 - 1; }
};

template <typename T, int OP> struct Red {
  static __device__ T id() {
    if constexpr (OP == NRS_ADD) return OpId<T>::add();
    else if constexpr (OP == NRS_MUL) return OpId<T>::mul();
    else if constexpr (OP == NRS_MIN) return OpId<T>::mn();
    else return OpId<T>::mx();
  }
  static __device__ T apply(T g, T a) {
    if constexpr (OP == NRS_ADD) return g + a;
    else if constexpr (OP == NRS_MUL) return g * a;
    else if constexpr (OP == NRS_MIN) return (a < g) ? a : g;
    else return (a > g) ? a : g;
  }
};

constexpr int kGsBlock = 256;

template <typename T, int OP, int W, bool WEIGHTED = false>
__device__ __forceinline__ void gs_fixed_row(const int *__restrict__ idx, T *__restrict__ v,
                                             int nvec, int64_t stride)
{
  int m[W];
  if constexpr (W == 2) {
    const int2 a = *reinterpret_cast<const int2 *>(idx);
    m[0] = a.x; m[1] = a.y;
  } else {
#pragma unroll
    for (int c = 0; c < W / 4; ++c) {
      const int4 a = *reinterpret_cast<const int4 *>(idx + 4 * c);
      m[4 * c] = a.x; m[4 * c + 1] = a.y; m[4 * c + 2] = a.z; m[4 * c + 3] = a.w;
    }
  }
  for (int k = 0; k < nvec; ++k) {
    T *vk = v + (int64_t)k * stride;
    T val[W];
#pragma unroll
    for (int c = 0; c < W; ++c) val[c] = vk[m[c]];
    if constexpr (WEIGHTED) {   // invDegree of a member of a W-row is 1/W, exact for W = 2, 4, 8
#pragma unroll
      for (int c = 0; c < W; ++c) val[c] = val[c] * (T(1) / T(W));
    }
    T g = Red<T, OP>::id();
#pragma unroll
    for (int c = 0; c < W; ++c) g = Red<T, OP>::apply(g, val[c]);
#pragma unroll
    for (int c = 0; c < W; ++c) vk[m[c]] = g;
  }
}

// One launch covers all buckets: blocks [0,b2) pairs, [b2,b4) quads, [b4,b8) octs,
// [b8,end) CSR rows; the bucket choice is uniform per block.
template <typename T, int OP>
__global__ void __launch_bounds__(kGsBlock)
gs_kernel(int n2, int n4, int n8, int nx, int b2, int b4, int b8, const int *__restrict__ pairs,
          const int *__restrict__ quads, const int *__restrict__ octs, const int *__restrict__ xoff,
          const int *__restrict__ xids, T *__restrict__ v, int nvec, int64_t stride)
{
  const int b = blockIdx.x;
  if (b < b2) {
    const int r = b * kGsBlock + threadIdx.x;
    if (r < n2) gs_fixed_row<T, OP, 2>(pairs + 2 * (int64_t)r, v, nvec, stride);
  } else if (b < b4) {
    const int r = (b - b2) * kGsBlock + threadIdx.x;
    if (r < n4) gs_fixed_row<T, OP, 4>(quads + 4 * (int64_t)r, v, nvec, stride);
  } else if (b < b8) {
    const int r = (b - b4) * kGsBlock + threadIdx.x;
    if (r < n8) gs_fixed_row<T, OP, 8>(octs + 8 * (int64_t)r, v, nvec, stride);
  } else {
    const int r = (b - b8) * kGsBlock + threadIdx.x;
    if (r >= nx) return;
    const int s = xoff[r], e = xoff[r + 1];
    for (int k = 0; k < nvec; ++k) {
      T *vk = v + (int64_t)k * stride;
      T g = Red<T, OP>::id();
      for (int m = s; m < e; ++m) g = Red<T, OP>::apply(g, vk[xids[m]]);
      for (int m = s; m < e; ++m) vk[xids[m]] = g;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kGsBlock)
gs_weighted_kernel(int n2, int n4, int n8, int nx, int b2, int b4, int b8, const int *__restrict__ pairs,
                   const int *__restrict__ quads, const int *__restrict__ octs,
                   const int *__restrict__ xoff, const int *__restrict__ xids,
                   const T *__restrict__ w, T *__restrict__ v)
{
  const int b = blockIdx.x;
  if (b < b2) {
    const int r = b * kGsBlock + threadIdx.x;
    if (r < n2) gs_fixed_row<T, NRS_ADD, 2, true>(pairs + 2 * (int64_t)r, v, 1, 0);
  } else if (b < b4) {
    const int r = (b - b2) * kGsBlock + threadIdx.x;
    if (r < n4) gs_fixed_row<T, NRS_ADD, 4, true>(quads + 4 * (int64_t)r, v, 1, 0);
  } else if (b < b8) {
    const int r = (b - b4) * kGsBlock + threadIdx.x;
    if (r < n8) gs_fixed_row<T, NRS_ADD, 8, true>(octs + 8 * (int64_t)r, v, 1, 0);
  } else {
    const int r = (b - b8) * kGsBlock + threadIdx.x;
    if (r >= nx) return;
    const int s0 = xoff[r], e0 = xoff[r + 1];
    T g = T(0);
    for (int m = s0; m < e0; ++m) g = g + v[xids[m]] * w[xids[m]];
    for (int m = s0; m < e0; ++m) v[xids[m]] = g;
  }
}

}  // namespace

template <typename T>
void ogs_gather_scatter_weighted(const nrs_ogs *o, T *v, const T *w, hipStream_t s)
{
  NRS_CHECK(o != nullptr, "ogs: null handle");
  if (o->Nrows == 0) return;
  const int b2 = (int)((o->n2 + kGsBlock - 1) / kGsBlock);
  const int b4 = b2 + (int)((o->n4 + kGsBlock - 1) / kGsBlock);
  const int b8 = b4 + (int)((o->n8 + kGsBlock - 1) / kGsBlock);
  const int bt = b8 + (int)((o->nx + kGsBlock - 1) / kGsBlock);
  gs_weighted_kernel<T><<<bt, kGsBlock, 0, s>>>((int)o->n2, (int)o->n4, (int)o->n8, (int)o->nx, b2, b4,
                                               b8, o->pairs.get(), o->quads.get(), o->octs.get(),
                                               o->xoff.get(), o->xids.get(), w, v);
  NRS_LAUNCH_CHECK();
}
template void ogs_gather_scatter_weighted<double>(const nrs_ogs *, double *, const double *, hipStream_t);
template void ogs_gather_scatter_weighted<float>(const nrs_ogs *, float *, const float *, hipStream_t);

template <typename T>
void ogs_gather_scatter(const nrs_ogs *o, T *v, int nvec, int64_t stride, int op, hipStream_t s)
{
  NRS_CHECK(o != nullptr, "ogs: null handle");
  NRS_CHECK(nvec >= 1, "ogs: nvec must be >= 1");
  NRS_CHECK(nvec == 1 || stride >= o->N, "ogs: stride smaller than the vector length");
  if (o->Nrows == 0) return;
  const int b2 = (int)((o->n2 + kGsBlock - 1) / kGsBlock);
  const int b4 = b2 + (int)((o->n4 + kGsBlock - 1) / kGsBlock);
  const int b8 = b4 + (int)((o->n8 + kGsBlock - 1) / kGsBlock);
  const int bt = b8 + (int)((o->nx + kGsBlock - 1) / kGsBlock);
#define GS_LAUNCH(OPV)                                                                             \
  gs_kernel<T, OPV><<<bt, kGsBlock, 0, s>>>((int)o->n2, (int)o->n4, (int)o->n8, (int)o->nx, b2, b4, \
                                            b8, o->pairs.get(), o->quads.get(), o->octs.get(),     \
                                            o->xoff.get(), o->xids.get(), v, nvec, stride)
  switch (op) {
    case NRS_ADD: GS_LAUNCH(NRS_ADD); break;
    case NRS_MUL: GS_LAUNCH(NRS_MUL); break;
    case NRS_MIN: GS_LAUNCH(NRS_MIN); break;
    case NRS_MAX: GS_LAUNCH(NRS_MAX); break;
    default: throw Error(NRS_ERR_INVALID, "ogs: unknown op");
  }
#undef GS_LAUNCH
  NRS_LAUNCH_CHECK();
}
template void ogs_gather_scatter<double>(const nrs_ogs *, double *, int, int64_t, int, hipStream_t);
template void ogs_gather_scatter<float>(const nrs_ogs *, float *, int, int64_t, int, hipStream_t);
template void ogs_gather_scatter<int>(const nrs_ogs *, int *, int, int64_t, int, hipStream_t);

template <typename T>
void ogs_gather_scatter_chunks(const nrs_ogs *o, T *v, int c0, int c1, hipStream_t s)
{
  NRS_CHECK(o != nullptr, "ogs: null handle");
  NRS_CHECK(o->nChunks > 0 && c0 >= 0 && c0 <= c1 && c1 <= o->nChunks, "ogs: bad chunk range");
  const int64_t r2 = o->cOff[0][c0], r4 = o->cOff[1][c0], r8 = o->cOff[2][c0], rx = o->cOff[3][c0];
  const int n2 = (int)(o->cOff[0][c1] - r2), n4 = (int)(o->cOff[1][c1] - r4);
  const int n8 = (int)(o->cOff[2][c1] - r8), nx = (int)(o->cOff[3][c1] - rx);
  const int b2 = (n2 + kGsBlock - 1) / kGsBlock;
  const int b4 = b2 + (n4 + kGsBlock - 1) / kGsBlock;
  const int b8 = b4 + (n8 + kGsBlock - 1) / kGsBlock;
  const int bt = b8 + (nx + kGsBlock - 1) / kGsBlock;
  if (bt == 0) return;
  gs_kernel<T, NRS_ADD><<<bt, kGsBlock, 0, s>>>(n2, n4, n8, nx, b2, b4, b8, o->pairs.get() + 2 * r2,
                                                o->quads.get() + 4 * r4, o->octs.get() + 8 * r8,
                                                o->xoff.get() + rx, o->xids.get(), v, 1, 0);
  NRS_LAUNCH_CHECK();
}
template void ogs_gather_scatter_chunks<double>(const nrs_ogs *, double *, int, int, hipStream_t);
template void ogs_gather_scatter_chunks<float>(const nrs_ogs *, float *, int, int, hipStream_t);

void ogs_set_chunks(nrs_ogs *o, int64_t nodesPerChunk, int nChunks)
{
  NRS_CHECK(o != nullptr && nodesPerChunk >= 1 && nChunks >= 1, "ogs: bad chunking");
  auto chunk_of = [&](int last) { return (int)std::min<int64_t>(last / nodesPerChunk, nChunks - 1); };
  // fixed-width buckets: stable sort of rows by chunk
  auto regroup = [&](std::vector<int> &h, int W, std::vector<int64_t> &off) {
    const int64_t n = (int64_t)h.size() / W;
    off.assign(nChunks + 1, 0);
    for (int64_t r = 0; r < n; ++r) off[chunk_of(h[r * W + W - 1]) + 1]++;
    for (int c = 0; c < nChunks; ++c) off[c + 1] += off[c];
    std::vector<int> out(h.size());
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    for (int64_t r = 0; r < n; ++r) {
      const int64_t d = pos[chunk_of(h[r * W + W - 1])]++;
      std::copy(h.begin() + r * W, h.begin() + (r + 1) * W, out.begin() + d * W);
    }
    h.swap(out);
  };
  regroup(o->h2, 2, o->cOff[0]);
  regroup(o->h4, 4, o->cOff[1]);
  regroup(o->h8, 8, o->cOff[2]);
  {   // CSR rows
    const int64_t n = (int64_t)o->hxo.size() - 1;
    auto &off = o->cOff[3];
    off.assign(nChunks + 1, 0);
    for (int64_t r = 0; r < n; ++r) off[chunk_of(o->hxi[o->hxo[r + 1] - 1]) + 1]++;
    for (int c = 0; c < nChunks; ++c) off[c + 1] += off[c];
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    std::vector<int64_t> dst(n);
    for (int64_t r = 0; r < n; ++r) dst[r] = pos[chunk_of(o->hxi[o->hxo[r + 1] - 1])]++;
    std::vector<int> len(n), xo(n + 1, 0), xi(o->hxi.size());
    for (int64_t r = 0; r < n; ++r) len[dst[r]] = o->hxo[r + 1] - o->hxo[r];
    for (int64_t r = 0; r < n; ++r) xo[r + 1] = xo[r] + len[r];
    for (int64_t r = 0; r < n; ++r)
      std::copy(o->hxi.begin() + o->hxo[r], o->hxi.begin() + o->hxo[r + 1], xi.begin() + xo[dst[r]]);
    o->hxo.swap(xo);
    o->hxi.swap(xi);
  }
  o->pairs.upload(o->h2);
  o->quads.upload(o->h4);
  o->octs.upload(o->h8);
  o->xoff.upload(o->hxo);
  o->xids.upload(o->hxi);
  o->nChunks = nChunks;
  o->nodesPerChunk = nodesPerChunk;
}

nrs_ogs *ogs_setup(int64_t N, const long long *ids)
{
  NRS_CHECK(N >= 0 && N < (int64_t)INT32_MAX, "ogs: N must fit in int32 (dlong)");
  NRS_CHECK(N == 0 || ids != nullptr, "ogs: null id array");
  auto *o = new nrs_ogs();
  o->N = N;
  // group members by id: counting sort when ids are dense, comparison sort otherwise
  long long maxId = 0, minId = 0;
  int64_t n = 0;
  for (int64_t i = 0; i < N; ++i) {
    if (ids[i] == 0) continue;
    maxId = std::max(maxId, ids[i]);
    minId = (n == 0) ? ids[i] : std::min(minId, ids[i]);
    ++n;
  }
  std::vector<int64_t> grpOf(N, -1);  // local index -> group number (first-appearance order)
  std::vector<int64_t> count;
  if (n > 0 && minId > 0 && maxId <= 8 * N + 1024) {
    std::vector<int64_t> firstRow(maxId + 1, -1);
    for (int64_t i = 0; i < N; ++i) {
      const long long g = ids[i];
      if (g == 0) continue;
      if (firstRow[g] < 0) { firstRow[g] = (int64_t)count.size(); count.push_back(0); }
      grpOf[i] = firstRow[g];
      count[firstRow[g]]++;
    }
  } else if (n > 0) {
    std::vector<std::pair<long long, int64_t>> nodes;
    nodes.reserve(n);
    for (int64_t i = 0; i < N; ++i)
      if (ids[i] != 0) nodes.emplace_back(ids[i], i);
    std::sort(nodes.begin(), nodes.end());
    // group index in sorted order, then renumber by first appearance
    std::vector<int64_t> sortedGrp(n);
    int64_t ng = 0;
    for (int64_t s = 0; s < n; ++s) {
      if (s == 0 || nodes[s].first != nodes[s - 1].first) ++ng;
      sortedGrp[s] = ng - 1;
      grpOf[nodes[s].second] = ng - 1;
    }
    std::vector<int64_t> rowOf(ng, -1);
    count.assign(ng, 0);
    int64_t r = 0;
    for (int64_t i = 0; i < N; ++i) {
      if (grpOf[i] < 0) continue;
      int64_t &ro = rowOf[grpOf[i]];
      if (ro < 0) ro = r++;
      grpOf[i] = ro;
      count[ro]++;
    }
  }
  const int64_t ng = (int64_t)count.size();
  // CSR of all rows (members ascending because locals are scanned in order)
  std::vector<int64_t> off(ng + 1, 0);
  for (int64_t g = 0; g < ng; ++g) off[g + 1] = off[g] + count[g];
  std::vector<int> mem(off[ng]);
  {
    std::vector<int64_t> fill(ng, 0);
    for (int64_t i = 0; i < N; ++i)
      if (grpOf[i] >= 0) mem[off[grpOf[i]] + fill[grpOf[i]]++] = (int)i;
  }
  o->invDegreeHost.assign(N, 1.0);
  std::vector<int> h2, h4, h8, hxo(1, 0), hxi;
  for (int64_t g = 0; g < ng; ++g) {
    const int64_t c = count[g];
    for (int64_t m = off[g]; m < off[g + 1]; ++m) o->invDegreeHost[mem[m]] = 1.0 / (double)c;
    if (c == 1) { o->Nsingletons++; continue; }
    o->Nrows++;
    o->Nmembers += c;
    std::vector<int> *dst = (c == 2) ? &h2 : (c == 4) ? &h4 : (c == 8) ? &h8 : nullptr;
    if (dst) {
      dst->insert(dst->end(), mem.begin() + off[g], mem.begin() + off[g + 1]);
    } else {
      hxi.insert(hxi.end(), mem.begin() + off[g], mem.begin() + off[g + 1]);
      hxo.push_back((int)hxi.size());
    }
  }
  o->n2 = (int64_t)h2.size() / 2;
  o->n4 = (int64_t)h4.size() / 4;
  o->n8 = (int64_t)h8.size() / 8;
  o->nx = (int64_t)hxo.size() - 1;
  o->pairs.upload(h2);
  o->quads.upload(h4);
  o->octs.upload(h8);
  o->xoff.upload(hxo);
  o->xids.upload(hxi);
  o->invDegree.upload(o->invDegreeHost);
  o->h2.swap(h2); o->h4.swap(h4); o->h8.swap(h8); o->hxo.swap(hxo); o->hxi.swap(hxi);
  return o;
}

}  // namespace nrs
