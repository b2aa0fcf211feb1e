// re2.hpp -- .re2 mesh files and the numbering of unstructured hexahedral meshes.
#pragma once
#include <vector>

#include "common.hpp"

namespace nrs {

struct Re2Header {
  int nelgt = 0, nelgv = 0;      // elements (all, fluid)
  int ncurve = 0, nbc = 0;       // curved-side records, boundary records of the first field
  bool swap = false;             // file written with the other byte order
};

Re2Header re2_header(const char *path);
// vx, vy, vz [nelgt][8] corners in lexicographic order (i + 2 j + 4 k); group
// [nelgt] (may be null); cbc [nelgt][6][3] boundary codes of the first field in
// the solver's face order (0 t-, 1 s-, 2 r+, 3 s+, 4 r-, 5 t+; blanks inside)
void re2_read(const char *path, double *vx, double *vy, double *vz, int *group, char *cbc,
              int *peer = nullptr);
// 1-based ids of the corners [E][8]: corners closer than tol share an id
// (tol <= 0: 1e-3 of the shortest element edge)
std::vector<long long> vertex_ids(int E, const double *vx, const double *vy, const double *vz, double tol);
// GLL node coordinates (trilinear map of the corners) and global ids at order N;
// peer [E*6*2] (or null): periodic face pairs whose corners are identified
void mesh_from_vertices(int N, int E, const double *vx, const double *vy, const double *vz, double tol,
                        double *x, double *y, double *z, long long *gids, const int *peer = nullptr);

// Curved-side records in file order: elem 0-based, side 1..12 (Nek's edge or
// face number), params [ncurve][5], type one character per record
void re2_curves(const char *path, int *elem, int *side, double *params, char *type);
// Overwrites x, y, z [E][(N+1)^3] of the elements with midside-node ('m')
// curved edges by their triquadratic Gordon-Hall map (Nek5000's xyzquad);
// other curve types throw NRS_ERR_UNSUPPORTED
void midside_map(int N, int E, const double *vx, const double *vy, const double *vz, int ncurve,
                 const int *elem, const int *side, const double *params, const char *type, double *x,
                 double *y, double *z);

}  // namespace nrs
