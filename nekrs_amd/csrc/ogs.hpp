// ogs.hpp -- gather-scatter handle (internal layout of nrs_ogs).
#pragma once
#include "common.hpp"

struct nrs_ogs {
  int64_t N = 0;              // local E-vector length
  int64_t Nrows = 0;          // rows with > 1 member
  int64_t Nmembers = 0;       // members of those rows
  int64_t Nsingletons = 0;    // rows with exactly one member (never touched)
  // Rows bucketed by multiplicity; inside a bucket rows keep the reference's
  // first-appearance order and members ascend, so sums are formed in the same
  // order as gatherScatterMany (left to right from the op identity).
  int64_t n2 = 0, n4 = 0, n8 = 0, nx = 0;   // rows per bucket (x = any other size)
  nrs::DevBuf<int> pairs;     // [n2][2]
  nrs::DevBuf<int> quads;     // [n4][4]
  nrs::DevBuf<int> octs;      // [n8][8]
  nrs::DevBuf<int> xoff;      // [nx+1] CSR offsets into xids
  nrs::DevBuf<int> xids;
  nrs::DevBuf<double> invDegree;  // [N]
  std::vector<double> invDegreeHost;
  // Tiles: on the device the rows of each bucket are grouped by the tile of
  // their last member (local index / nodesPerTile), first-appearance order
  // inside a tile; tile t's rows of bucket b are [tOff[b][t], tOff[b][t+1]).
  // One block runs one tile, all four buckets, so the lines that a tile's
  // pairs, quads and octs share are fetched once into one XCD's L2.
  int nTiles = 0;
  int64_t nodesPerTile = 0;
  nrs::DevBuf<int> tOff;      // [4][nTiles + 1]
  // Chunks (optional, nrs::ogs_set_chunks): chunk c (last member / nodesPerChunk,
  // at most nChunks - 1) is the tile range [cTile[c], cTile[c+1]).  A row of
  // chunk c only touches nodes below the end of chunk c.
  int nChunks = 0;
  int64_t nodesPerChunk = 0;
  std::vector<int> cTile;
  // host copies of the buckets in first-appearance order
  std::vector<int> h2, h4, h8, hxo, hxi;
  // the call between nrs_ogs_start and nrs_ogs_finish (oogs::start / finish)
  struct Pending { const void *v; int nvec; int64_t stride; int dtype, op; };
  bool pending = false;
  Pending pend{};
};

namespace nrs {
// In-place gather-scatter of nvec vectors (stride apart) of type T with op.
template <typename T>
void ogs_gather_scatter(const nrs_ogs *o, T *v, int nvec, int64_t stride, int op, hipStream_t s);
nrs_ogs *ogs_setup(int64_t N, const long long *ids);
// v <- gs_add(w .* v) for one vector, w = the handle's invDegree in T: each
// member is scaled before its row sum; nodes outside any row are left as they
// are (their weight is 1), except the nZero nodes zeroIds, stored as zero in
// the same launch (a mask; they must be in no row of the handle, which needs
// rows when nZero > 0).
template <typename T>
void ogs_gather_scatter_weighted(const nrs_ogs *o, T *v, const T *w, hipStream_t s,
                                 const int *zeroIds = nullptr, int64_t nZero = 0);
// Group rows by chunk (chunk = last member / nodesPerChunk, at most nChunks - 1),
// re-tiling so that tiles nest in chunks.  Reorders rows inside buckets only:
// every gather-scatter result is unchanged.
void ogs_set_chunks(nrs_ogs *o, int64_t nodesPerChunk, int nChunks);
// v <- gs_add(v) on the rows of chunks [c0, c1) only (one vector).  After the
// nodes of those chunks are final, this completes their rows; rows of later
// chunks are untouched.
template <typename T>
void ogs_gather_scatter_chunks(const nrs_ogs *o, T *v, int c0, int c1, hipStream_t s);
}  // namespace nrs
