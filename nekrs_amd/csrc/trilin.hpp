// trilin.hpp -- geometric factors of a trilinear hexahedron formed at a node
// from the element's eight corners, for kernels that would otherwise stream
// them from memory.
//
// Corners [3][8] (x, y, z of corner a + 2b + 4c at (r, s, t) = (+-1, +-1, +-1)).
// A lane that owns the node column (r_i, s_j) keeps the corner differences
// the map's derivatives take there (init, once per element); at a plane t_k
// the derivatives, the Jacobian, the inverse by cofactors and JW = J w_i w_j
// w_k follow geometricFactorsHex3D (kernels/mesh/geometricFactorsHex3D.okl:
// 27-143), whose spectral derivatives D x of a trilinear map equal these up
// to rounding.
#pragma once
#include <hip/hip_runtime.h>

namespace nrs {

template <typename T> struct TriGeo {
  T dRB[3], dRT[3], dSB[3], dST[3], dT[3], wij;

  __device__ __forceinline__ void init(const T *__restrict__ P0, T r, T s, T wi, T wj)
  {
    const T hr0 = T(0.5) * (T(1) - r), hr1 = T(0.5) * (T(1) + r);
    const T hs0 = T(0.5) * (T(1) - s), hs1 = T(0.5) * (T(1) + s);
    wij = wi * wj;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T *P = P0 + 8 * c;
      const T p0 = P[0], p1 = P[1], p2 = P[2], p3 = P[3], p4 = P[4], p5 = P[5], p6 = P[6], p7 = P[7];
      dRB[c] = T(0.5) * (hs0 * (p1 - p0) + hs1 * (p3 - p2));
      dRT[c] = T(0.5) * (hs0 * (p5 - p4) + hs1 * (p7 - p6));
      dSB[c] = T(0.5) * (hr0 * (p2 - p0) + hr1 * (p3 - p1));
      dST[c] = T(0.5) * (hr0 * (p6 - p4) + hr1 * (p7 - p5));
      dT[c] = T(0.5) * (hs0 * (hr0 * (p4 - p0) + hr1 * (p5 - p1)) + hs1 * (hr0 * (p6 - p2) + hr1 * (p7 - p3)));
    }
  }

  // g = rx ry rz sx sy sz tx ty tz JW at the plane t (GLL weight wk)
  __device__ __forceinline__ void at(T t, T wk, T (&g)[10]) const
  {
    const T ht0 = T(0.5) * (T(1) - t), ht1 = T(0.5) * (T(1) + t);
    const T xr = ht0 * dRB[0] + ht1 * dRT[0], yr = ht0 * dRB[1] + ht1 * dRT[1], zr = ht0 * dRB[2] + ht1 * dRT[2];
    const T xs = ht0 * dSB[0] + ht1 * dST[0], ys = ht0 * dSB[1] + ht1 * dST[1], zs = ht0 * dSB[2] + ht1 * dST[2];
    const T xt = dT[0], yt = dT[1], zt = dT[2];
    const T J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
    const T Jinv = T(1) / J;
    g[0] = (ys * zt - zs * yt) * Jinv;
    g[1] = -(xs * zt - zs * xt) * Jinv;
    g[2] = (xs * yt - ys * xt) * Jinv;
    g[3] = -(yr * zt - zr * yt) * Jinv;
    g[4] = (xr * zt - zr * xt) * Jinv;
    g[5] = -(xr * yt - yr * xt) * Jinv;
    g[6] = (yr * zs - zr * ys) * Jinv;
    g[7] = -(xr * zs - zr * xs) * Jinv;
    g[8] = (xr * ys - yr * xs) * Jinv;
    g[9] = J * wij * wk;
  }
};

}  // namespace nrs
