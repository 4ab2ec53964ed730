// Isoparametric geometry of a P2 triangle at one Gauss point (device and
// CPU-emulator code).
#pragma once
#include "sfx_portable.h"

namespace sfx {

// Inverse transpose of the Jacobian of the map from the reference triangle
// at one Gauss point, from the six node coordinates rn, zn and the reference
// derivatives d = (dx, dy) pairs of the six shape functions at that point:
// dr(k,k1) = SUM(rr(k,jj(:,m))*dd(k1,:,l)), dw = J^{-T} dd
// (SOURCES/GAUSS_POINTS/gauss_points_2d.F90:108-121).  Exact for curved
// elements, whose midpoint nodes sit on a curved boundary
// (SOURCES/MESH/refine_mesh.F90:342-353); equal to the vertex formula for
// straight-sided ones.  d/dr = a11 d/dx + a12 d/dy, d/dz = a21 d/dx + a22 d/dy.
SFX_DEV void p2_inverse_jacobian(const double (&rn)[6], const double (&zn)[6], const double* d, double& a11,
                                 double& a12, double& a21, double& a22) {
  double j11 = 0, j12 = 0, j21 = 0, j22 = 0;   // dr/dx, dr/dy, dz/dx, dz/dy
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    j11 = fma(rn[n], d[2 * n], j11);
    j12 = fma(rn[n], d[2 * n + 1], j12);
    j21 = fma(zn[n], d[2 * n], j21);
    j22 = fma(zn[n], d[2 * n + 1], j22);
  }
  const double idet = 1.0 / (j11 * j22 - j12 * j21);
  a11 = j22 * idet;  a12 = -j21 * idet;
  a21 = -j12 * idet; a22 = j11 * idet;
}

}  // namespace sfx
