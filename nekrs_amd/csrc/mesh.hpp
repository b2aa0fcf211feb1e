// mesh.hpp -- host-side basis and box-mesh helpers.
#pragma once
#include <vector>

namespace nrs {
void gll_nodes(int N, std::vector<double> &z, std::vector<double> &w);
std::vector<double> dmatrix(const std::vector<double> &z);
std::vector<double> interp_matrix(const std::vector<double> &zin, const std::vector<double> &zout);
void box_mesh(int nx, int ny, int nz, int N, int map, double eps, double *x, double *y, double *z,
              long long *gid, int *EToB);
}  // namespace nrs
