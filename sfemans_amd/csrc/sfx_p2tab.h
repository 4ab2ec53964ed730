// Reference P2 triangle tables (host side).
#pragma once
#include <cmath>

namespace sfx {

// Shape functions and derivatives of the P2 triangle at Stroud's 7 points
// (SOURCES/GAUSS_POINTS/GP_2D_p2.F90:9-88): w(6,7) then d(2,6,7), column-major.
inline void p2_reference_tables(double* tab) {
  const double half = 0.5, one = 1, two = 2, three = 3, four = 4;
  const double sq = std::sqrt(15.0);
  const double r = one / three;
  const double s1 = (6 - sq) / 21, t1 = (9 + 2 * sq) / 21;
  const double s2 = (6 + sq) / 21, t2 = (9 - 2 * sq) / 21;
  const double xx[7] = {r, s1, s1, t1, s2, s2, t2};
  const double yy[7] = {r, s1, t1, s1, s2, t2, s2};
  for (int l = 0; l < 7; ++l) {
    const double x = xx[l], y = yy[l];
    double* w = tab + 6 * l;
    double* d = tab + 42 + 12 * l;
    w[0] = (half - x - y) * (one - x - y) * two;
    w[1] = x * (x - half) * two;
    w[2] = y * (y - half) * two;
    w[3] = x * y * four;
    w[4] = y * (one - x - y) * four;
    w[5] = x * (one - x - y) * four;
    d[0] = -three + four * (x + y);     d[1] = -three + four * (x + y);
    d[2] = (two * x - half) * two;      d[3] = 0;
    d[4] = 0;                           d[5] = (two * y - half) * two;
    d[6] = y * four;                    d[7] = x * four;
    d[8] = -y * four;                   d[9] = (one - x - two * y) * four;
    d[10] = (one - two * x - y) * four; d[11] = -x * four;
  }
}

// Stroud's 7-point weights on the reference triangle (GP_2D_p2.F90:53-67).
inline void p2_stroud_weights(double* p) {
  const double area = 0.5, sq = std::sqrt(15.0);
  const double a = area * 9.0 / 40, b1 = area * (155 - sq) / 1200, b2 = area * (155 + sq) / 1200;
  p[0] = a; p[1] = p[2] = p[3] = b1; p[4] = p[5] = p[6] = b2;
}

// Determinant of the isoparametric Jacobian of element J (1-based node ids)
// at Gauss point l: dr(k,k1) = SUM(rr(k,jj(:,m))*dd(k1,:,l)) (gauss_points_2d.F90:108-114).
inline double p2_jacobian_det(const int* J, const double* rr, const double* tab, int l) {
  const double* d = tab + 42 + 12 * l;
  double j11 = 0, j12 = 0, j21 = 0, j22 = 0;
  for (int n = 0; n < 6; ++n) {
    const double r = rr[2 * (J[n] - 1)], z = rr[2 * (J[n] - 1) + 1];
    j11 += r * d[2 * n];
    j12 += r * d[2 * n + 1];
    j21 += z * d[2 * n];
    j22 += z * d[2 * n + 1];
  }
  return j11 * j22 - j12 * j21;
}

// Per element and Gauss point: rj(l,m) * ray(l,m) = |det J(l)| p(l) ray(l)
// (gauss_points_2d.F90:122; ray = SUM(rr(1,jj(:,m))*ww(:,l))), with the
// Jacobian of the isoparametric map at the point (curved elements included).
inline void p2_assembly_weights(const int* jj, const double* rr, long long me, const double* tab, double* wgt) {
  double p[7];
  p2_stroud_weights(p);
  for (long long e = 0; e < me; ++e) {
    const int* J = jj + 6 * e;
    for (int l = 0; l < 7; ++l) {
      double ray = 0;
      for (int n = 0; n < 6; ++n) ray += rr[2 * (J[n] - 1)] * tab[n + 6 * l];
      wgt[7 * e + l] = std::fabs(p2_jacobian_det(J, rr, tab, l)) * p[l] * ray;
    }
  }
}

}  // namespace sfx
