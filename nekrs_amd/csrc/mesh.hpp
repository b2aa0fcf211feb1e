// mesh.hpp -- host-side basis and box-mesh helpers.
#pragma once
#include <vector>

namespace nrs {
void gll_nodes(int N, std::vector<double> &z, std::vector<double> &w);
std::vector<double> dmatrix(const std::vector<double> &z);
std::vector<double> interp_matrix(const std::vector<double> &zin, const std::vector<double> &zout);
void box_mesh(int nx, int ny, int nz, int N, int map, double eps, double *x, double *y, double *z,
              long long *gid, int *EToB);
// True when every element's GLL nodes are the trilinear map of its eight corner
// nodes (within 1e-12 of the element's size); corners then holds [E][3][8]
// (x, y, z of corner a + 2b + 4c at node (aN, bN, cN)).
bool trilinear_corners(int N, int E, const double *x, const double *y, const double *z,
                       std::vector<double> &corners);
}  // namespace nrs
