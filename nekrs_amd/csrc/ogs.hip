// ogs.hip -- gather-scatter (QQ^T) on E-vectors as a bucketed segment reduction.
//
// Setup restates the local part of ogsSetup (3rd_party/gslib/ogs/src/ogsSetup.cpp:
// 111-246, 364-392): rows = groups of equal nonzero global ids, ordered by the
// local index of their first member, members ascending; invDegree = 1/count (1 at
// excluded nodes).  Unlike the reference, singleton rows are not stored at all
// and multi-member rows are bucketed by length (2, 4, 8, other) so that each
// thread reads its member indices with one aligned vector load and runs a
// fixed-trip loop.  On the device the rows are grouped into tiles by their last
// member (one block per tile, XCD-contiguous), which keeps the face lines a
// tile's rows share in L2.  The reduction order inside a row is the reference's
// (gatherScatterMany.okl:58-86: identity op m0 op m1 ...), so FP64 results match
// the reference kernel bit for bit.
#include <algorithm>
#include <memory>
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

// NV > 0: the vector count known at compile time; every vector's members are
// loaded before any is stored (the vectors are stride >= N apart, so a row's
// stores never touch another vector's members), which keeps NV x W loads in
// flight instead of W.  NV = 0: a loop over nvec.
template <typename T, int OP, int W, bool WEIGHTED, int NV>
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
  if constexpr (NV > 0) {
    T val[NV][W];
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int c = 0; c < W; ++c) val[k][c] = v[(int64_t)k * stride + m[c]];
    if constexpr (WEIGHTED) {   // invDegree of a member of a W-row is 1/W, exact for W = 2, 4, 8
#pragma unroll
      for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int c = 0; c < W; ++c) val[k][c] = val[k][c] * (T(1) / T(W));
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      T g = Red<T, OP>::id();
#pragma unroll
      for (int c = 0; c < W; ++c) g = Red<T, OP>::apply(g, val[k][c]);
#pragma unroll
      for (int c = 0; c < W; ++c) v[(int64_t)k * stride + m[c]] = g;
    }
  } else {
    for (int k = 0; k < nvec; ++k) {
      T *vk = v + (int64_t)k * stride;
      T val[W];
#pragma unroll
      for (int c = 0; c < W; ++c) val[c] = vk[m[c]];
      if constexpr (WEIGHTED) {
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
}

// One block per tile; the tile's pairs, quads, octs and CSR rows run one after
// the other, so the lines they share (a face line holds face, edge and vertex
// nodes) stay in one XCD's L2 between them (dealing all of a tile's rows to the
// threads as one list measured 3-6 % slower).  Blocks are dealt round-robin to
// the 8 XCDs, so block b runs tile (b mod 8) * per + b / 8: each XCD walks a
// contiguous run of tiles and the lines that neighbouring tiles share stay in
// its L2 as well.  WEIGHTED: v <- gs_add(w .* v) (nvec = 1, NV = 1), and the nz
// nodes zids (in no row) stored as zero, dealt evenly over the blocks.  NV: the
// vector count when fixed at compile time (1 or 3), 0 for any nvec.
template <typename T, int OP, bool WEIGHTED, int NV>
__global__ void __launch_bounds__(kGsBlock)
gs_kernel(int t0, int nt, int per, const int *__restrict__ toff, int tstride,
          const int *__restrict__ pairs, const int *__restrict__ quads, const int *__restrict__ octs,
          const int *__restrict__ xoff, const int *__restrict__ xids, const T *__restrict__ w,
          T *__restrict__ v, int nvec, int64_t stride, const int *__restrict__ zids, int64_t nz)
{
  if constexpr (WEIGHTED) {
    if (nz > 0) {
      const int64_t b = blockIdx.x, nb = gridDim.x, z1 = nz * (b + 1) / nb;
      for (int64_t i = nz * b / nb + threadIdx.x; i < z1; i += kGsBlock) v[zids[i]] = T(0);
    }
  }
  const int t = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (t >= nt) return;
  const int *tb = toff + t0 + t;
  const int s2 = tb[0], e2 = tb[1];
  const int s4 = tb[tstride], e4 = tb[tstride + 1];
  const int s8 = tb[2 * tstride], e8 = tb[2 * tstride + 1];
  const int sx = tb[3 * tstride], ex = tb[3 * tstride + 1];
  for (int r = s2 + (int)threadIdx.x; r < e2; r += kGsBlock)
    gs_fixed_row<T, OP, 2, WEIGHTED, NV>(pairs + 2 * (int64_t)r, v, nvec, stride);
  for (int r = s4 + (int)threadIdx.x; r < e4; r += kGsBlock)
    gs_fixed_row<T, OP, 4, WEIGHTED, NV>(quads + 4 * (int64_t)r, v, nvec, stride);
  for (int r = s8 + (int)threadIdx.x; r < e8; r += kGsBlock)
    gs_fixed_row<T, OP, 8, WEIGHTED, NV>(octs + 8 * (int64_t)r, v, nvec, stride);
  for (int r = sx + (int)threadIdx.x; r < ex; r += kGsBlock) {
    const int s0 = xoff[r], e0 = xoff[r + 1];
    if constexpr (WEIGHTED) {
      T g = T(0);
      for (int m = s0; m < e0; ++m) g = g + v[xids[m]] * w[xids[m]];
      for (int m = s0; m < e0; ++m) v[xids[m]] = g;
    } else {
      for (int k = 0; k < nvec; ++k) {
        T *vk = v + (int64_t)k * stride;
        T g = Red<T, OP>::id();
        for (int m = s0; m < e0; ++m) g = Red<T, OP>::apply(g, vk[xids[m]]);
        for (int m = s0; m < e0; ++m) vk[xids[m]] = g;
      }
    }
  }
}

// Launches tiles [t0, t1) of the handle.
template <typename T, int OP, bool WEIGHTED>
void launch_gs(const nrs_ogs *o, int t0, int t1, const T *w, T *v, int nvec, int64_t stride,
               hipStream_t s, const int *zids = nullptr, int64_t nz = 0)
{
  const int nt = t1 - t0;
  if (nt <= 0) return;
  const int per = (nt + 7) / 8;
  auto go = [&](auto kernel) {
    kernel<<<8 * per, kGsBlock, 0, s>>>(t0, nt, per, o->tOff.get(), o->nTiles + 1, o->pairs.get(),
                                        o->quads.get(), o->octs.get(), o->xoff.get(), o->xids.get(), w, v,
                                        nvec, stride, zids, nz);
  };
  if (nvec == 1) go(gs_kernel<T, OP, WEIGHTED, 1>);
  else if (nvec == 3 && !WEIGHTED) go(gs_kernel<T, OP, WEIGHTED, 3>);
  else go(gs_kernel<T, OP, WEIGHTED, 0>);
  NRS_LAUNCH_CHECK();
}

// nodes per tile: a few hundred rows per block (N = 7: two elements, ~250 rows;
// 512 / 1024 / 2048 / 4096 timed at 32^3 and 48^3, 1024 the fastest by 5-8 %)
constexpr int64_t kTileNodes = 1024;

// Device layout of the buckets grouped by tile (last member / npt); tiles
// never straddle a chunk when npt divides the chunk size.
void ogs_layout(nrs_ogs *o, int64_t npt)
{
  const int64_t nt64 = std::max<int64_t>(1, (o->N + npt - 1) / npt);
  NRS_CHECK(nt64 < (int64_t)INT32_MAX / 4, "ogs: too many tiles");
  const int nt = (int)nt64;
  std::vector<int> toff(4 * (size_t)(nt + 1), 0);
  auto tile_of = [&](int last) { return (int)std::min<int64_t>(last / npt, nt - 1); };
  auto group = [&](const std::vector<int> &h, int W, int b, DevBuf<int> &dev) {
    const int64_t n = (int64_t)h.size() / W;
    int *off = toff.data() + (size_t)b * (nt + 1);
    for (int64_t r = 0; r < n; ++r) off[tile_of(h[r * W + W - 1]) + 1]++;
    for (int t = 0; t < nt; ++t) off[t + 1] += off[t];
    std::vector<int> out(h.size());
    std::vector<int> pos(off, off + nt);
    for (int64_t r = 0; r < n; ++r) {
      const int64_t d = pos[tile_of(h[r * W + W - 1])]++;
      std::copy(h.begin() + r * W, h.begin() + (r + 1) * W, out.begin() + d * W);
    }
    dev.upload(out);
  };
  group(o->h2, 2, 0, o->pairs);
  group(o->h4, 4, 1, o->quads);
  group(o->h8, 8, 2, o->octs);
  {   // CSR rows
    const int64_t n = (int64_t)o->hxo.size() - 1;
    int *off = toff.data() + (size_t)3 * (nt + 1);
    for (int64_t r = 0; r < n; ++r) off[tile_of(o->hxi[o->hxo[r + 1] - 1]) + 1]++;
    for (int t = 0; t < nt; ++t) off[t + 1] += off[t];
    std::vector<int> pos(off, off + nt);
    std::vector<int64_t> dst(n);
    for (int64_t r = 0; r < n; ++r) dst[r] = pos[tile_of(o->hxi[o->hxo[r + 1] - 1])]++;
    std::vector<int> len(n), xo(n + 1, 0), xi(o->hxi.size());
    for (int64_t r = 0; r < n; ++r) len[dst[r]] = o->hxo[r + 1] - o->hxo[r];
    for (int64_t r = 0; r < n; ++r) xo[r + 1] = xo[r] + len[r];
    for (int64_t r = 0; r < n; ++r)
      std::copy(o->hxi.begin() + o->hxo[r], o->hxi.begin() + o->hxo[r + 1], xi.begin() + xo[dst[r]]);
    o->xoff.upload(xo);
    o->xids.upload(xi);
  }
  o->tOff.upload(toff);
  o->nTiles = nt;
  o->nodesPerTile = npt;
}

}  // namespace

template <typename T>
void ogs_gather_scatter_weighted(const nrs_ogs *o, T *v, const T *w, hipStream_t s, const int *zeroIds,
                                 int64_t nZero)
{
  NRS_CHECK(o != nullptr, "ogs: null handle");
  NRS_CHECK(nZero == 0 || (zeroIds && o->Nrows > 0), "ogs: ids to zero need a handle with rows");
  if (o->Nrows == 0) return;
  launch_gs<T, NRS_ADD, true>(o, 0, o->nTiles, w, v, 1, 0, s, zeroIds, nZero);
}
template void ogs_gather_scatter_weighted<double>(const nrs_ogs *, double *, const double *, hipStream_t,
                                                  const int *, int64_t);
template void ogs_gather_scatter_weighted<float>(const nrs_ogs *, float *, const float *, hipStream_t,
                                                 const int *, int64_t);

template <typename T>
void ogs_gather_scatter(const nrs_ogs *o, T *v, int nvec, int64_t stride, int op, hipStream_t s)
{
  NRS_CHECK(o != nullptr, "ogs: null handle");
  NRS_CHECK(nvec >= 1, "ogs: nvec must be >= 1");
  NRS_CHECK(nvec == 1 || stride >= o->N, "ogs: stride smaller than the vector length");
  if (o->Nrows == 0) return;
  switch (op) {
    case NRS_ADD: launch_gs<T, NRS_ADD, false>(o, 0, o->nTiles, nullptr, v, nvec, stride, s); break;
    case NRS_MUL: launch_gs<T, NRS_MUL, false>(o, 0, o->nTiles, nullptr, v, nvec, stride, s); break;
    case NRS_MIN: launch_gs<T, NRS_MIN, false>(o, 0, o->nTiles, nullptr, v, nvec, stride, s); break;
    case NRS_MAX: launch_gs<T, NRS_MAX, false>(o, 0, o->nTiles, nullptr, v, nvec, stride, s); break;
    default: throw Error(NRS_ERR_INVALID, "ogs: unknown op");
  }
}
template void ogs_gather_scatter<double>(const nrs_ogs *, double *, int, int64_t, int, hipStream_t);
template void ogs_gather_scatter<float>(const nrs_ogs *, float *, int, int64_t, int, hipStream_t);
template void ogs_gather_scatter<int>(const nrs_ogs *, int *, int, int64_t, int, hipStream_t);

template <typename T>
void ogs_gather_scatter_chunks(const nrs_ogs *o, T *v, int c0, int c1, hipStream_t s)
{
  NRS_CHECK(o != nullptr, "ogs: null handle");
  NRS_CHECK(o->nChunks > 0 && c0 >= 0 && c0 <= c1 && c1 <= o->nChunks, "ogs: bad chunk range");
  if (o->Nrows == 0) return;
  launch_gs<T, NRS_ADD, false>(o, o->cTile[c0], o->cTile[c1], nullptr, v, 1, 0, s);
}
template void ogs_gather_scatter_chunks<double>(const nrs_ogs *, double *, int, int, hipStream_t);
template void ogs_gather_scatter_chunks<float>(const nrs_ogs *, float *, int, int, hipStream_t);

void ogs_set_chunks(nrs_ogs *o, int64_t nodesPerChunk, int nChunks)
{
  NRS_CHECK(o != nullptr && nodesPerChunk >= 1 && nChunks >= 1, "ogs: bad chunking");
  // tiles of npt = nodesPerChunk / k nodes, the largest such divisor not above the default
  int64_t k = (nodesPerChunk + kTileNodes - 1) / kTileNodes;
  while (nodesPerChunk % k) ++k;
  const int64_t npt = nodesPerChunk / k;
  ogs_layout(o, npt);
  // chunk c = min(last / nodesPerChunk, nChunks - 1) covers tiles [c k, (c + 1) k), the
  // last chunk every tile after it
  o->cTile.assign(nChunks + 1, 0);
  for (int c = 0; c <= nChunks; ++c)
    o->cTile[c] = (int)std::min<int64_t>((int64_t)c * k, o->nTiles);
  o->cTile[nChunks] = o->nTiles;
  o->nChunks = nChunks;
  o->nodesPerChunk = nodesPerChunk;
}

nrs_ogs *ogs_setup(int64_t N, const long long *ids)
{
  NRS_CHECK(N >= 0 && N < (int64_t)INT32_MAX, "ogs: N must fit in int32 (dlong)");
  NRS_CHECK(N == 0 || ids != nullptr, "ogs: null id array");
  std::unique_ptr<nrs_ogs> o(new nrs_ogs());
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
  o->h2.swap(h2); o->h4.swap(h4); o->h8.swap(h8); o->hxo.swap(hxo); o->hxi.swap(hxi);
  ogs_layout(o.get(), kTileNodes);
  o->invDegree.upload(o->invDegreeHost);
  return o.release();
}

}  // namespace nrs
