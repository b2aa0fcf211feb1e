// re2.hip -- the .re2 binary mesh format and the numbering of an unstructured
// hexahedral mesh (host code).
//
// Format (src/io/re2Reader.cpp:3-41; 3rd_party/nek5000/core/reader_re2.f):
//   80-byte ASCII header "#v002 nelgt ndim nelgv ..." (#v002/#v003/#v004),
//   4-byte float 6.54321 (byte-order test), then in 8-byte words:
//   nelgt element records  [group, x[8], y[8], z[8]]            (buf_to_xyz)
//   ncurve, then ncurve    [element, face, curve[5], type]       (buf_to_curve)
//   per field: nbc, then   [element, face, bc[5], "cbc"]         (buf_to_bc)
//   vertices in the preprocessor's cyclic order (bottom face counter-clockwise,
//   then top), faces 1..6 = s-, r+, s+, r-, t-, t+ (1-based element ids).
// The numbering: vertices that coincide within a tolerance share an id
// (nek's set_glo_num from coordinates), then every GLL node is numbered by the
// vertex, edge, face or interior it belongs to (the same topology rule as the
// multigrid levels' ids).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <unordered_map>

#include "mesh.hpp"
#include "mg.hpp"
#include "re2.hpp"

namespace nrs {

namespace {

struct File {
  FILE *f = nullptr;
  explicit File(const char *p) : f(std::fopen(p, "rb")) {}
  ~File() { if (f) std::fclose(f); }
};

void read_exact(FILE *f, void *dst, size_t n, const char *what)
{
  if (std::fread(dst, 1, n, f) != n) throw Error(NRS_ERR_INVALID, std::string("re2: truncated file (") + what + ")");
}

double swap8(double v)
{
  unsigned char b[8];
  std::memcpy(b, &v, 8);
  for (int i = 0; i < 4; ++i) std::swap(b[i], b[7 - i]);
  std::memcpy(&v, b, 8);
  return v;
}

// cyclic (file) vertex j -> lexicographic corner i + 2 j + 4 k (meshNekReader.cpp:30)
constexpr int kVtxMap[8] = {0, 1, 3, 2, 4, 5, 7, 6};
// file face 1..6 -> solver face (0 t-, 1 s-, 2 r+, 3 s+, 4 r-, 5 t+) (meshNekReader.cpp:31)
constexpr int kFaceMap[6] = {1, 2, 3, 4, 0, 5};

}  // namespace

Re2Header re2_header(const char *path)
{
  NRS_CHECK(path != nullptr, "re2: null path");
  File fh(path);
  NRS_CHECK(fh.f != nullptr, std::string("re2: cannot open ") + path);
  char hdr[81] = {0};
  read_exact(fh.f, hdr, 80, "header");
  Re2Header h;
  char ver[6] = {0};
  std::memcpy(ver, hdr, 5);
  int ndim = 0;
  if (std::strcmp(ver, "#v004") == 0) {
    NRS_CHECK(std::sscanf(hdr, "%5s %d %d %d", ver, &h.nelgt, &ndim, &h.nelgv) == 4, "re2: bad #v004 header");
  } else if (std::strcmp(ver, "#v002") == 0 || std::strcmp(ver, "#v003") == 0) {
    NRS_CHECK(std::sscanf(hdr, "%5s %9d %1d %9d", ver, &h.nelgt, &ndim, &h.nelgv) == 4, "re2: bad header");
  } else {
    throw Error(NRS_ERR_UNSUPPORTED, std::string("re2: unsupported version ") + ver);
  }
  NRS_CHECK(ndim == 3, "re2: only 3D meshes");
  NRS_CHECK(h.nelgt > 0 && h.nelgv > 0 && h.nelgv <= h.nelgt, "re2: invalid element counts");
  float test = 0.0f;
  read_exact(fh.f, &test, 4, "byte-order word");
  if (std::fabs(test - 6.54321f) < 1e-5f) {
    h.swap = false;
  } else {
    unsigned char b[4];
    std::memcpy(b, &test, 4);
    std::swap(b[0], b[3]); std::swap(b[1], b[2]);
    std::memcpy(&test, b, 4);
    NRS_CHECK(std::fabs(test - 6.54321f) < 1e-5f, "re2: bad byte-order word");
    h.swap = true;
  }
  // the counts must fit the file: no record or array is sized beyond what it holds
  NRS_CHECK(std::fseek(fh.f, 0, SEEK_END) == 0, "re2: seek");
  const long long bytes = std::ftell(fh.f);
  NRS_CHECK(bytes >= 84 + (long long)h.nelgt * 200 + 16, "re2: file shorter than its element count");
  // skip the elements, read the curve and boundary counts
  NRS_CHECK(std::fseek(fh.f, 84 + (long)h.nelgt * 25 * 8, SEEK_SET) == 0, "re2: seek");
  double c = 0.0;
  read_exact(fh.f, &c, 8, "curve count");
  c = h.swap ? swap8(c) : c;
  NRS_CHECK(c >= 0.0 && c <= (double)(bytes - 84 - (long long)h.nelgt * 200 - 16) / 64.0,
            "re2: bad curve count");
  h.ncurve = (int)c;
  NRS_CHECK(std::fseek(fh.f, (long)h.ncurve * 8 * 8, SEEK_CUR) == 0, "re2: seek");
  double nb = 0.0;
  read_exact(fh.f, &nb, 8, "boundary count");
  nb = h.swap ? swap8(nb) : nb;
  NRS_CHECK(nb >= 0.0 && nb <= (double)(bytes - 84 - (long long)h.nelgt * 200 - 16 -
                                         (long long)h.ncurve * 64) / 64.0,
            "re2: bad boundary count");
  h.nbc = (int)nb;
  return h;
}

void re2_read(const char *path, double *vx, double *vy, double *vz, int *group, char *cbc, int *peer)
{
  const Re2Header h = re2_header(path);
  File fh(path);
  NRS_CHECK(fh.f != nullptr, std::string("re2: cannot open ") + path);
  NRS_CHECK(std::fseek(fh.f, 84, SEEK_SET) == 0, "re2: seek");
  std::vector<double> rec(25);
  for (int e = 0; e < h.nelgt; ++e) {
    read_exact(fh.f, rec.data(), 25 * 8, "element record");
    if (h.swap) for (auto &v : rec) v = swap8(v);
    if (group) group[e] = (int)rec[0];
    for (int j = 0; j < 8; ++j) {
      const int p = kVtxMap[j];
      vx[e * 8 + p] = rec[1 + j];
      vy[e * 8 + p] = rec[9 + j];
      vz[e * 8 + p] = rec[17 + j];
    }
  }
  NRS_CHECK(std::fseek(fh.f, 8 + (long)h.ncurve * 8 * 8, SEEK_CUR) == 0, "re2: seek");
  double nb = 0.0;
  read_exact(fh.f, &nb, 8, "boundary count");
  if (cbc || peer) {
    if (cbc) std::memset(cbc, ' ', (size_t)h.nelgt * 6 * 3);
    if (peer) std::fill(peer, peer + (size_t)h.nelgt * 12, -1);
    std::vector<double> b(8);
    for (int r = 0; r < h.nbc; ++r) {
      read_exact(fh.f, b.data(), 64, "boundary record");
      double eg = b[0], f = b[1], c0 = b[2], c1 = b[3];
      char code[8];
      std::memcpy(code, &b[7], 8);
      if (h.swap) { eg = swap8(eg); f = swap8(f); c0 = swap8(c0); c1 = swap8(c1); }
      const int e = (int)eg - 1, face = (int)f - 1;
      NRS_CHECK(e >= 0 && e < h.nelgt && face >= 0 && face < 6, "re2: boundary record out of range");
      const size_t sf = (size_t)e * 6 + kFaceMap[face];
      if (cbc) std::memcpy(cbc + sf * 3, code, 3);
      // periodic: bc(1) = the connected element, bc(2) = its face (buf_to_bc, reader_re2.f:503-541)
      if (peer && code[0] == 'P' && code[1] == ' ') {
        const int e2 = (int)c0 - 1, f2 = (int)c1 - 1;
        NRS_CHECK(e2 >= 0 && e2 < h.nelgt && f2 >= 0 && f2 < 6, "re2: periodic record out of range");
        peer[sf * 2] = e2;
        peer[sf * 2 + 1] = kFaceMap[f2];
      }
    }
  }
}

std::vector<long long> vertex_ids(int E, const double *vx, const double *vy, const double *vz, double tol)
{
  // tolerance: a fraction of the shortest element edge unless given
  if (!(tol > 0.0)) {
    double hmin = 1e300;
    const int edges[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                              {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    for (int e = 0; e < E; ++e)
      for (auto &ed : edges) {
        const int a = e * 8 + ed[0], b = e * 8 + ed[1];
        const double d = std::sqrt((vx[a] - vx[b]) * (vx[a] - vx[b]) + (vy[a] - vy[b]) * (vy[a] - vy[b]) +
                                   (vz[a] - vz[b]) * (vz[a] - vz[b]));
        hmin = std::min(hmin, d);
      }
    NRS_CHECK(hmin > 0.0, "mesh: degenerate element (an edge of length zero)");
    tol = 1e-3 * hmin;
  }
  // spatial hash with cells of size 4 tol; search the 27 neighbouring cells
  const double cell = 4.0 * tol;
  double vmax = 0.0;
  for (int64_t n = 0; n < (int64_t)E * 8; ++n) {
    NRS_CHECK(std::isfinite(vx[n]) && std::isfinite(vy[n]) && std::isfinite(vz[n]),
              "mesh: corner coordinates must be finite");
    vmax = std::max(vmax, std::max(std::fabs(vx[n]), std::max(std::fabs(vy[n]), std::fabs(vz[n]))));
  }
  NRS_CHECK(vmax / cell < 1e15, "mesh: tolerance too small for the coordinate range");
  struct KeyHash {
    size_t operator()(const std::array<long long, 3> &k) const
    {
      return (size_t)((uint64_t)k[0] * 73856093ULL ^ (uint64_t)k[1] * 19349663ULL ^
                      (uint64_t)k[2] * 83492791ULL);
    }
  };
  std::unordered_map<std::array<long long, 3>, std::vector<int>, KeyHash> grid;
  std::vector<double> px, py, pz;
  std::vector<long long> ids((size_t)E * 8);
  for (int64_t n = 0; n < (int64_t)E * 8; ++n) {
    const long long cx = (long long)std::floor(vx[n] / cell), cy = (long long)std::floor(vy[n] / cell),
                    cz = (long long)std::floor(vz[n] / cell);
    int found = -1;
    for (int dx = -1; dx <= 1 && found < 0; ++dx)
      for (int dy = -1; dy <= 1 && found < 0; ++dy)
        for (int dz = -1; dz <= 1 && found < 0; ++dz) {
          auto it = grid.find({cx + dx, cy + dy, cz + dz});
          if (it == grid.end()) continue;
          for (int u : it->second) {
            const double d2 = (px[u] - vx[n]) * (px[u] - vx[n]) + (py[u] - vy[n]) * (py[u] - vy[n]) +
                              (pz[u] - vz[n]) * (pz[u] - vz[n]);
            if (d2 <= tol * tol) { found = u; break; }
          }
        }
    if (found < 0) {
      found = (int)px.size();
      px.push_back(vx[n]); py.push_back(vy[n]); pz.push_back(vz[n]);
      grid[{cx, cy, cz}].push_back(found);
    }
    ids[n] = found + 1;
  }
  return ids;
}

namespace {

// corners (lexicographic i + 2j + 4k) of the solver's faces 0..5
const int kSolverFaceCorners[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {1, 3, 5, 7},
                                      {2, 3, 6, 7}, {0, 2, 4, 6}, {4, 5, 6, 7}};

long long find_root(std::vector<long long> &par, long long a)
{
  while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; }
  return a;
}

// Vertex ids with the corners of periodic face pairs identified: each corner
// of face (e, f) is matched with the corner of its peer face that lies closest
// after the translation between the two face centres.  Refuses meshes whose
// periodic identification makes two corners of one element, or two different
// edges, the same (fewer than three elements across a periodic direction).
void connect_periodic(int E, const double *vx, const double *vy, const double *vz, const int *peer,
                      std::vector<long long> &vid)
{
  long long nv = 0;
  for (long long v : vid) nv = std::max(nv, v);
  std::vector<long long> par(nv + 1);
  for (long long i = 0; i <= nv; ++i) par[i] = i;
  auto P = [&](int64_t c) { return std::array<double, 3>{vx[c], vy[c], vz[c]}; };
  for (int e = 0; e < E; ++e)
    for (int f = 0; f < 6; ++f) {
      const int e2 = peer[((size_t)e * 6 + f) * 2], f2 = peer[((size_t)e * 6 + f) * 2 + 1];
      if (e2 < 0) continue;
      NRS_CHECK(e2 < E && f2 >= 0 && f2 < 6, "mesh: periodic peer out of range");
      std::array<double, 3> c1{0, 0, 0}, c2{0, 0, 0};
      for (int q = 0; q < 4; ++q)
        for (int d = 0; d < 3; ++d) {
          c1[d] += 0.25 * P((int64_t)e * 8 + kSolverFaceCorners[f][q])[d];
          c2[d] += 0.25 * P((int64_t)e2 * 8 + kSolverFaceCorners[f2][q])[d];
        }
      for (int q = 0; q < 4; ++q) {
        const int64_t a = (int64_t)e * 8 + kSolverFaceCorners[f][q];
        int64_t best = -1;
        double bd = 1e300;
        for (int q2 = 0; q2 < 4; ++q2) {
          const int64_t b = (int64_t)e2 * 8 + kSolverFaceCorners[f2][q2];
          double d2 = 0;
          for (int d = 0; d < 3; ++d) {
            const double t = (P(b)[d] - c2[d]) - (P(a)[d] - c1[d]);
            d2 += t * t;
          }
          if (d2 < bd) { bd = d2; best = b; }
        }
        const long long ra = find_root(par, vid[a]), rb = find_root(par, vid[best]);
        if (ra != rb) par[std::max(ra, rb)] = std::min(ra, rb);
      }
    }
  // compact 1-based ids in order of first appearance of the roots
  std::vector<long long> label(nv + 1, 0);
  long long next = 0;
  for (auto &v : vid) {
    const long long r = find_root(par, v);
    if (!label[r]) label[r] = ++next;
    v = label[r];
  }
  const int edges[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                            {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  std::map<std::pair<long long, long long>, std::array<double, 3>> dir;
  for (int e = 0; e < E; ++e) {
    for (int a = 0; a < 8; ++a)
      for (int b = a + 1; b < 8; ++b)
        NRS_CHECK(vid[(size_t)e * 8 + a] != vid[(size_t)e * 8 + b],
                  "mesh: periodic connection joins two corners of one element (fewer than three "
                  "elements across a periodic direction)");
    for (auto &ed : edges) {
      int64_t a = (int64_t)e * 8 + ed[0], b = (int64_t)e * 8 + ed[1];
      if (vid[a] > vid[b]) std::swap(a, b);
      const std::array<double, 3> v{vx[b] - vx[a], vy[b] - vy[a], vz[b] - vz[a]};
      auto it = dir.find({vid[a], vid[b]});
      if (it == dir.end()) { dir[{vid[a], vid[b]}] = v; continue; }
      const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      double d2 = 0;
      for (int d = 0; d < 3; ++d) d2 += (it->second[d] - v[d]) * (it->second[d] - v[d]);
      NRS_CHECK(std::sqrt(d2) <= 1e-6 * len, "mesh: periodic connection makes two different edges one "
                                             "(fewer than three elements across a periodic direction)");
    }
  }
}

}  // namespace

void mesh_from_vertices(int N, int E, const double *vx, const double *vy, const double *vz, double tol,
                        double *x, double *y, double *z, long long *gids, const int *peer)
{
  NRS_CHECK(N >= 1 && N <= 9, "mesh: N must be 1..9");
  NRS_CHECK(E > 0, "mesh: no elements");
  std::vector<long long> vid = vertex_ids(E, vx, vy, vz, tol);
  if (peer) connect_periodic(E, vx, vy, vz, peer, vid);
  const std::vector<long long> g = coarse_global_ids(1, E, vid.data(), N);
  std::vector<double> r, w;
  gll_nodes(N, r, w);
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  for (int64_t e = 0; e < E; ++e)
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const double a[2] = {0.5 * (1 - r[i]), 0.5 * (1 + r[i])};
          const double b[2] = {0.5 * (1 - r[j]), 0.5 * (1 + r[j])};
          const double c[2] = {0.5 * (1 - r[k]), 0.5 * (1 + r[k])};
          double X = 0, Y = 0, Z = 0;
          for (int cc = 0; cc < 2; ++cc)
            for (int bb = 0; bb < 2; ++bb)
              for (int aa = 0; aa < 2; ++aa) {
                const double wgt = a[aa] * b[bb] * c[cc];
                const int64_t v = e * 8 + aa + 2 * bb + 4 * cc;
                X += wgt * vx[v]; Y += wgt * vy[v]; Z += wgt * vz[v];
              }
          const int64_t n = e * Np + i + Nq * (j + Nq * k);
          x[n] = X; y[n] = Y; z[n] = Z;
          gids[n] = g[n];
        }
}

// Curved-side records (readp_re2_curve / buf_to_curve, reader_re2.f:159-284,
// 473-498): [element, side, curve[5], type] in 8-byte words, the type a
// character in the last word.
void re2_curves(const char *path, int *elem, int *side, double *params, char *type)
{
  const Re2Header h = re2_header(path);
  File fh(path);
  NRS_CHECK(fh.f != nullptr, std::string("re2: cannot open ") + path);
  NRS_CHECK(std::fseek(fh.f, 84 + (long)h.nelgt * 25 * 8 + 8, SEEK_SET) == 0, "re2: seek");
  double b[8];
  for (int r = 0; r < h.ncurve; ++r) {
    read_exact(fh.f, b, 64, "curve record");
    char t[8];
    std::memcpy(t, &b[7], 8);
    if (h.swap) for (int k = 0; k < 7; ++k) b[k] = swap8(b[k]);
    const int e = (int)b[0] - 1, sd = (int)b[1];
    NRS_CHECK(e >= 0 && e < h.nelgt && sd >= 1 && sd <= 12, "re2: curve record out of range");
    elem[r] = e;
    side[r] = sd;
    for (int k = 0; k < 5; ++k) params[(int64_t)r * 5 + k] = b[2 + k];
    type[r] = t[0];
  }
}

// Midside-node ('m') curved edges: the element becomes the triquadratic
// Gordon-Hall map through its corners and edge midpoints, as Nek5000's
// xyzquad (3rd_party/nek5000/core/genxyz.f:1334-1410, called from linquad
// :1235-1267): the trilinear map sampled on the 3 x 3 x 3 points r, s, t in
// {-1, 0, 1}; each 'm' edge replaces its midpoint; vertex + edge transfinite
// interpolation (gh_face_extend_3d, gh_type = 2, navier5.f:2501-2590) fills
// the face and interior points; quadratic Lagrange interpolation onto the GLL
// nodes.  Other curve types (circle arcs 'C', spheres 's', 'e') are refused.
void midside_map(int N, int E, const double *vx, const double *vy, const double *vz, int ncurve,
                 const int *elem, const int *side, const double *params, const char *type, double *x,
                 double *y, double *z)
{
  NRS_CHECK(N >= 1 && N <= 9, "midside: N must be 1..9");
  // 3 x 3 x 3 index (i + 3 j + 9 k) of the midpoint of Nek's edges 1..12 (xyzquad's eindx)
  static const int kEdgeMid[12] = {1, 5, 7, 3, 19, 23, 25, 21, 9, 11, 17, 15};
  std::vector<std::vector<int>> curved(E);
  for (int r = 0; r < ncurve; ++r) {
    NRS_CHECK(elem[r] >= 0 && elem[r] < E && side[r] >= 1 && side[r] <= 12,
              "midside: curve record out of range");
    const char t = type[r];
    if (t == ' ' || t == '\0') continue;
    if (t != 'm')
      throw Error(NRS_ERR_UNSUPPORTED, std::string("midside: curved side type '") + t +
                                           "' (only midside nodes 'm' are applied)");
    curved[elem[r]].push_back(r);
  }
  if (N == 1) return;            // linquad: order 1 keeps the trilinear map
  std::vector<double> gz, gw;
  gll_nodes(N, gz, gw);
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  // quadratic Lagrange weights through -1, 0, 1 at the GLL nodes
  std::vector<std::array<double, 3>> L(Nq);
  for (int i = 0; i < Nq; ++i) {
    const double t = gz[i];
    L[i] = {0.5 * t * (t - 1.0), (1.0 - t) * (1.0 + t), 0.5 * t * (t + 1.0)};
  }
  const double zg[3] = {-1.0, 0.0, 1.0};
  auto hat = [&](int end, int i) {   // linear blend of end 0 / 2 at point i
    return end == 0 ? 0.5 * (1.0 - zg[i]) : 0.5 * (1.0 + zg[i]);
  };
  for (int e = 0; e < E; ++e) {
    if (curved[e].empty()) continue;
    for (int d = 0; d < 3; ++d) {
      const double *vc = (d == 0 ? vx : d == 1 ? vy : vz) + (int64_t)e * 8;
      double q[27];
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j)
          for (int i = 0; i < 3; ++i) {
            double s = 0.0;
            for (int c = 0; c < 8; ++c)
              s += hat(c & 1 ? 2 : 0, i) * hat(c & 2 ? 2 : 0, j) * hat(c & 4 ? 2 : 0, k) * vc[c];
            q[i + 3 * j + 9 * k] = s;
          }
      for (int r : curved[e]) q[kEdgeMid[side[r] - 1]] = params[(int64_t)r * 5 + d];
      // vertex interpolant v, then v + the 12 edge corrections
      double v[27] = {0}, g[27];
      for (int kk = 0; kk < 3; kk += 2)
        for (int jj = 0; jj < 3; jj += 2)
          for (int ii = 0; ii < 3; ii += 2)
            for (int n = 0; n < 27; ++n)
              v[n] += hat(ii, n % 3) * hat(jj, (n / 3) % 3) * hat(kk, n / 9) * q[ii + 3 * jj + 9 * kk];
      for (int n = 0; n < 27; ++n) g[n] = 0.0;
      for (int n = 0; n < 27; ++n) {
        const int i = n % 3, j = (n / 3) % 3, k = n / 9;
        for (int kk = 0; kk < 3; kk += 2)
          for (int jj = 0; jj < 3; jj += 2) {   // x-edges
            const int m = i + 3 * jj + 9 * kk;
            g[n] += hat(jj, j) * hat(kk, k) * (q[m] - v[m]);
          }
        for (int kk = 0; kk < 3; kk += 2)
          for (int ii = 0; ii < 3; ii += 2) {   // y-edges
            const int m = ii + 3 * j + 9 * kk;
            g[n] += hat(ii, i) * hat(kk, k) * (q[m] - v[m]);
          }
        for (int jj = 0; jj < 3; jj += 2)
          for (int ii = 0; ii < 3; ii += 2) {   // z-edges
            const int m = ii + 3 * jj + 9 * k;
            g[n] += hat(ii, i) * hat(jj, j) * (q[m] - v[m]);
          }
      }
      for (int n = 0; n < 27; ++n) g[n] += v[n];
      double *out = (d == 0 ? x : d == 1 ? y : z) + (int64_t)e * Np;
      for (int k = 0; k < Nq; ++k)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i) {
            double s = 0.0;
            for (int c = 0; c < 3; ++c)
              for (int b2 = 0; b2 < 3; ++b2)
                for (int a = 0; a < 3; ++a) s += L[i][a] * L[j][b2] * L[k][c] * g[a + 3 * b2 + 9 * c];
            out[i + Nq * (j + Nq * k)] = s;
          }
    }
  }
}

}  // namespace nrs
