/* CPU restatement of the nonlinear term at the Gauss points, in C with
 * OpenMP over element blocks.
 *
 * TEST INFRASTRUCTURE ONLY: a second, compiled restatement of the reference
 * algorithm, used by tests/ as a cross-check of the numpy oracle and by the
 * cpu_baseline leg of bench.py as the CPU timing.  The product path never
 * loads it.
 *
 * sfx_cpu_smb_cross_prod_gauss restates smb_cross_prod_gauss_sft_par without
 * precession (SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:514-630):
 *   - P2 tables and Stroud rule (SOURCES/GAUSS_POINTS/GP_2D_p2.F90:9-88), the
 *     Jacobian and dw = J^{-T} dd from all six nodes
 *     (SOURCES/GAUSS_POINTS/gauss_points_2d.F90:102-121), the radius
 *     SUM(rr(1,jj(:,m))*ww(:,l)) (subroutine_ns_with_velocity.F90:574);
 *   - W and RotV in cos/sin form (:577-608);
 *   - FFT_PAR_CROSS_PROD_DCL(RotV, W) with m_max_pad = 3M/2
 *     (SOURCES/POST_PROCESSING/fft_parallel.F90:147-328, nb_procs = 1):
 *     synthesis on the reference grid N = 2*m_max_pad - 1 of
 *     x(n) = a_0 + sum_{m>=1} a_m cos(2 pi m n/N) + b_m sin(2 pi m n/N)
 *     (the packing c_m = (a_m - i b_m)/2 followed by the unnormalised c2r
 *     transform, :227-267), the cross product on the grid (:270-274), and
 *     the analysis scaled 1/N, kept as mode 0 and 2*conj for m >= 1
 *     (:277-324): a_m = (2/N) sum_n x(n) cos, b_m = (2/N) sum_n x(n) sin,
 *     a_0 = (1/N) sum_n x(n), b_0 = 0.
 * The transforms are direct sums over a block of points (the inner loop runs
 * over points and vectorises), folded over the pairs n, N - n of the odd
 * grid; they are sums of the same terms an FFT would form, in another order.
 *
 * sfx_cpu_smb_cross_prod_gauss_fft is the same computation with the
 * azimuthal transforms done by a mixed-radix (4, 2, 3) Stockham FFT on the
 * smooth grid N = 3*2^k >= 3M - 2 (the reference's FFTW plans have the same
 * role on N = 2*m_max_pad - 1; the product is bilinear, so every grid with
 * N >= 3M - 2 gives the same coefficients up to rounding).  Two real fields
 * share one complex transform (z = x1 + i x2), as FFTW's batched c2r does
 * per point: three inverse transforms for the six fields, two forward ones
 * for the three components of the product.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BE 8            /* elements per block */
#define BP (7 * BE)     /* Gauss points per block */

static void p2_tables(double w[7][6], double d[7][6][2], double p[7]) {
  const double sq = sqrt(15.0), r = 1.0 / 3.0, area = 0.5;
  const double s1 = (6 - sq) / 21, t1 = (9 + 2 * sq) / 21, b1 = area * (155 - sq) / 1200;
  const double s2 = (6 + sq) / 21, t2 = (9 - 2 * sq) / 21, b2 = area * (155 + sq) / 1200;
  const double xx[7] = {r, s1, s1, t1, s2, s2, t2};
  const double yy[7] = {r, s1, t1, s1, s2, t2, s2};
  const double pp[7] = {area * 9.0 / 40, b1, b1, b1, b2, b2, b2};
  for (int l = 0; l < 7; ++l) {
    const double x = xx[l], y = yy[l];
    p[l] = pp[l];
    w[l][0] = (0.5 - x - y) * (1 - x - y) * 2;
    w[l][1] = x * (x - 0.5) * 2;
    w[l][2] = y * (y - 0.5) * 2;
    w[l][3] = x * y * 4;
    w[l][4] = y * (1 - x - y) * 4;
    w[l][5] = x * (1 - x - y) * 4;
    d[l][0][0] = -3 + 4 * (x + y);   d[l][0][1] = -3 + 4 * (x + y);
    d[l][1][0] = (2 * x - 0.5) * 2;  d[l][1][1] = 0;
    d[l][2][0] = 0;                  d[l][2][1] = (2 * y - 0.5) * 2;
    d[l][3][0] = y * 4;              d[l][3][1] = x * 4;
    d[l][4][0] = -y * 4;             d[l][4][1] = (1 - x - 2 * y) * 4;
    d[l][5][0] = (1 - 2 * x - y) * 4; d[l][5][1] = -x * 4;
  }
}

/* RotV (columns 0..5) and W (columns 6..11) at the 7*ne Gauss points of the
 * elements e0 .. e0+ne-1, A[(column * M + mode) * BP + point]
 * (subroutine_ns_with_velocity.F90:559-608). */
static void gauss_block(const int32_t* jj, const double* rr, int64_t np, const int32_t* list_mode, int32_t M,
                        const double* V_in, const double w[7][6], const double dd[7][6][2], int64_t e0, int ne,
                        double* A) {
  memset(A, 0, sizeof(double) * 12 * (size_t)M * BP);
  for (int el = 0; el < ne; ++el) {
    const int64_t e = e0 + el;
    int64_t j[6];
    for (int n = 0; n < 6; ++n) j[n] = (int64_t)jj[6 * e + n] - 1;
    for (int l = 0; l < 7; ++l) {
      double dr[2][2] = {{0, 0}, {0, 0}};
      for (int n = 0; n < 6; ++n)
        for (int k = 0; k < 2; ++k) {
          dr[k][0] += rr[2 * j[n] + k] * dd[l][n][0];
          dr[k][1] += rr[2 * j[n] + k] * dd[l][n][1];
        }
      const double rjac = dr[0][0] * dr[1][1] - dr[0][1] * dr[1][0];
      double dwr[6], dwz[6], ray = 0;
      for (int n = 0; n < 6; ++n) {
        dwr[n] = (dd[l][n][0] * dr[1][1] - dd[l][n][1] * dr[1][0]) / rjac;
        dwz[n] = (-dd[l][n][0] * dr[0][1] + dd[l][n][1] * dr[0][0]) / rjac;
        ray += rr[2 * j[n]] * w[l][n];
      }
      const int p = 7 * el + l;
      for (int i = 0; i < M; ++i) {
        double W[6] = {0}, Dr[6] = {0}, Dz[6] = {0};
        for (int n = 0; n < 6; ++n)
          for (int k = 0; k < 6; ++k) {
            const double v = V_in[j[n] + np * (k + 6 * (int64_t)i)];
            W[k] += v * w[l][n];
            Dr[k] += v * dwr[n];
            Dz[k] += v * dwz[n];
          }
        const double mr = (double)list_mode[i] / ray, ir = 1.0 / ray;
        double R[6];
        R[0] = mr * W[5] - Dz[2];
        R[3] = Dz[1] - Dr[5];
        R[4] = ir * W[2] + Dr[2] - mr * W[1];
        R[1] = -mr * W[4] - Dz[3];
        R[2] = Dz[0] - Dr[4];
        R[5] = ir * W[3] + Dr[3] + mr * W[0];
        for (int k = 0; k < 6; ++k) {
          A[((size_t)k * M + i) * BP + p] = R[k];
          A[((size_t)(6 + k) * M + i) * BP + p] = W[k];
        }
      }
    }
  }
}

/* V_in (np, 6, M) and V_out (7 me, 6, M), column-major (point fastest);
 * jj (6, me) 1-based, rr (2, np).  Returns 0, or 1 on bad arguments. */
int sfx_cpu_smb_cross_prod_gauss(const int32_t* jj, const double* rr, int64_t me, int64_t np,
                                 const int32_t* list_mode, int32_t M, const double* V_in,
                                 double* V_out, int32_t nthreads) {
  if (M <= 0 || me < 0 || np <= 0) return 1;
  const int mpad = 3 * M / 2, N = 2 * mpad - 1;
  const int64_t G = 7 * me;
  double w[7][6], dd[7][6][2], pp[7];
  p2_tables(w, dd, pp);
  /* cos / sin tables, (m n mod N) for accuracy */
  double* C = (double*)malloc(sizeof(double) * (size_t)M * N);
  double* S = (double*)malloc(sizeof(double) * (size_t)M * N);
  if (!C || !S) { free(C); free(S); return 1; }
  const double two_pi = 6.283185307179586476925286766559;
  for (int m = 0; m < M; ++m)
    for (int n = 0; n < N; ++n) {
      const double t = two_pi * (double)(((int64_t)m * n) % N) / N;
      C[(size_t)m * N + n] = cos(t);
      S[(size_t)m * N + n] = sin(t);
    }
  const int64_t nblk = (me + BE - 1) / BE;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    /* coefficients [field 0..11][mode][point]: RotV (6 columns) then W */
    double* A = (double*)malloc(sizeof(double) * 12 * (size_t)M * BP);
    /* grid values [field 0..5][n][point] and the product [3][n][point] */
    double* X = (double*)malloc(sizeof(double) * 6 * (size_t)N * BP);
    double* P = (double*)malloc(sizeof(double) * 3 * (size_t)N * BP);
#pragma omp for schedule(dynamic, 4)
    for (int64_t b = 0; b < nblk; ++b) {
      const int64_t e0 = b * BE;
      const int ne = (int)((me - e0) < BE ? (me - e0) : BE);
      gauss_block(jj, rr, np, list_mode, M, V_in, (const double(*)[6])w, (const double(*)[6][2])dd, e0, ne, A);
      const int npt = 7 * ne;
      /* synthesis of the six real fields (RotV r, theta, z; W r, theta, z).
       * N is odd: cos is even and sin odd under n -> N - n, so the cosine
       * part E and the sine part O give x(n) = E + O and x(N - n) = E - O. */
      for (int f = 0; f < 6; ++f) {
        const double* a = A + (size_t)(2 * f) * M * BP;       /* cos column */
        const double* bb = A + (size_t)(2 * f + 1) * M * BP;  /* sin column */
        double* x0 = X + ((size_t)f * N) * BP;
        for (int p = 0; p < npt; ++p) x0[p] = a[p];            /* mode 0: a_0 (b_0 dropped) */
        for (int m = 1; m < M; ++m)
          for (int p = 0; p < npt; ++p) x0[p] += a[(size_t)m * BP + p];
        for (int n = 1; n <= (N - 1) / 2; ++n) {
          double E[BP], O[BP];
          for (int p = 0; p < npt; ++p) { E[p] = a[p]; O[p] = 0.0; }
          for (int m = 1; m < M; ++m) {
            const double c = C[(size_t)m * N + n], s = S[(size_t)m * N + n];
            const double* am = a + (size_t)m * BP;
            const double* bm = bb + (size_t)m * BP;
            for (int p = 0; p < npt; ++p) { E[p] += am[p] * c; O[p] += bm[p] * s; }
          }
          double* xp = X + ((size_t)f * N + n) * BP;
          double* xm = X + ((size_t)f * N + (N - n)) * BP;
          for (int p = 0; p < npt; ++p) { xp[p] = E[p] + O[p]; xm[p] = E[p] - O[p]; }
        }
      }
      /* cross product RotV x W on the grid */
      for (int n = 0; n < N; ++n) {
        const double* r0 = X + ((size_t)0 * N + n) * BP;
        const double* r1 = X + ((size_t)1 * N + n) * BP;
        const double* r2 = X + ((size_t)2 * N + n) * BP;
        const double* w0 = X + ((size_t)3 * N + n) * BP;
        const double* w1 = X + ((size_t)4 * N + n) * BP;
        const double* w2 = X + ((size_t)5 * N + n) * BP;
        double* c0 = P + ((size_t)0 * N + n) * BP;
        double* c1 = P + ((size_t)1 * N + n) * BP;
        double* c2 = P + ((size_t)2 * N + n) * BP;
        for (int p = 0; p < npt; ++p) {
          c0[p] = r1[p] * w2[p] - r2[p] * w1[p];
          c1[p] = r2[p] * w0[p] - r0[p] * w2[p];
          c2[p] = r0[p] * w1[p] - r1[p] * w0[p];
        }
      }
      /* analysis, truncated to modes 0..M-1, stored as (cos, sin) columns;
       * pairs n, N - n share cos (sum) and sin (difference) */
      for (int c = 0; c < 3; ++c) {
        double* xs = X;   /* reuse: sums and differences of the product */
        const double* x = P + ((size_t)c * N) * BP;
        for (int n = 1; n <= (N - 1) / 2; ++n) {
          const double* xp = x + (size_t)n * BP;
          const double* xm = x + (size_t)(N - n) * BP;
          double* su = xs + (size_t)(2 * n) * BP;
          double* di = xs + (size_t)(2 * n + 1) * BP;
          for (int p = 0; p < npt; ++p) { su[p] = xp[p] + xm[p]; di[p] = xp[p] - xm[p]; }
        }
        for (int m = 0; m < M; ++m) {
          double sa[BP], sb[BP];
          for (int p = 0; p < npt; ++p) { sa[p] = x[p]; sb[p] = 0; }
          for (int n = 1; n <= (N - 1) / 2; ++n) {
            const double cc = C[(size_t)m * N + n], ss = S[(size_t)m * N + n];
            const double* su = xs + (size_t)(2 * n) * BP;
            const double* di = xs + (size_t)(2 * n + 1) * BP;
            for (int p = 0; p < npt; ++p) { sa[p] += su[p] * cc; sb[p] += di[p] * ss; }
          }
          const double sc = (m == 0 ? 1.0 : 2.0) / N;
          for (int p = 0; p < npt; ++p) {
            const int64_t g = 7 * e0 + p;
            V_out[g + G * (2 * c + 6 * (int64_t)m)] = sc * sa[p];
            V_out[g + G * (2 * c + 1 + 6 * (int64_t)m)] = (m == 0) ? 0.0 : sc * sb[p];
          }
        }
      }
    }
    free(A); free(X); free(P);
  }
  free(C); free(S);
  return 0;
}

/* ---------------------------------------------------------------------------
 * FFT variant.  Complex vectors of BP points: re[k * BP + p], im[k * BP + p].
 * ------------------------------------------------------------------------- */

/* Radices of N = 3 * 2^k: fours, then a two if k is odd, then the three. */
static int fft_radices(int N, int* rad) {
  int n = N, k = 0;
  if (n % 3) return -1;
  n /= 3;
  while (n % 4 == 0) { rad[k++] = 4; n /= 4; }
  if (n == 2) { rad[k++] = 2; n = 1; }
  if (n != 1) return -1;
  rad[k++] = 3;
  return k;
}

/* Stockham decimation-in-frequency pass of radix r over a length-n
 * sub-transform at stride s (n = r * m): inputs x[q + s (p + t m)],
 * outputs y[q + s (r p + u)] = (sum_t x_t w_r^{t u}) w_n^{p u}, with
 * w_n = exp(sgn 2 pi i / n).  Self-sorting: after all passes the result is
 * in natural order.  tw holds cos/sin(2 pi j / N) for j < N. */
static void fft_pass(int r, int n, int s, int N, int sgn, const double* tc, const double* ts,
                     const double* xr, const double* xi, double* yr, double* yi) {
  const int m = n / r;
  const double h = 0.86602540378443864676;   /* sqrt(3)/2 */
  for (int p = 0; p < m; ++p) {
    double wr[4], wi[4];
    for (int u = 0; u < r; ++u) {
      const int j = (int)(((int64_t)p * u * (N / n)) % N);
      wr[u] = tc[j];
      wi[u] = sgn * ts[j];
    }
    for (int q = 0; q < s; ++q) {
      const double* ar[4] = {xr, xr, xr, xr};
      const double* ai[4] = {xi, xi, xi, xi};
      double* br[4] = {yr, yr, yr, yr};
      double* bi[4] = {yi, yi, yi, yi};
      for (int t = 0; t < r; ++t) {
        ar[t] = xr + (size_t)(q + s * (p + t * m)) * BP;
        ai[t] = xi + (size_t)(q + s * (p + t * m)) * BP;
        br[t] = yr + (size_t)(q + s * (r * p + t)) * BP;
        bi[t] = yi + (size_t)(q + s * (r * p + t)) * BP;
      }
      if (r == 2) {
        for (int k = 0; k < BP; ++k) {
          const double dr = ar[0][k] - ar[1][k], di = ai[0][k] - ai[1][k];
          br[0][k] = ar[0][k] + ar[1][k];
          bi[0][k] = ai[0][k] + ai[1][k];
          br[1][k] = dr * wr[1] - di * wi[1];
          bi[1][k] = dr * wi[1] + di * wr[1];
        }
      } else if (r == 4) {
        /* w_4 = sgn * i: y1 = a0 + w a1 - a2 - w a3, y3 = a0 - w a1 - a2 + w a3 */
        for (int k = 0; k < BP; ++k) {
          const double s0r = ar[0][k] + ar[2][k], s0i = ai[0][k] + ai[2][k];
          const double d0r = ar[0][k] - ar[2][k], d0i = ai[0][k] - ai[2][k];
          const double s1r = ar[1][k] + ar[3][k], s1i = ai[1][k] + ai[3][k];
          const double d1r = ar[1][k] - ar[3][k], d1i = ai[1][k] - ai[3][k];
          /* sgn i (d1) */
          const double jr = -sgn * d1i, ji = sgn * d1r;
          const double y1r = d0r + jr, y1i = d0i + ji;
          const double y2r = s0r - s1r, y2i = s0i - s1i;
          const double y3r = d0r - jr, y3i = d0i - ji;
          br[0][k] = s0r + s1r;
          bi[0][k] = s0i + s1i;
          br[1][k] = y1r * wr[1] - y1i * wi[1];
          bi[1][k] = y1r * wi[1] + y1i * wr[1];
          br[2][k] = y2r * wr[2] - y2i * wi[2];
          bi[2][k] = y2r * wi[2] + y2i * wr[2];
          br[3][k] = y3r * wr[3] - y3i * wi[3];
          bi[3][k] = y3r * wi[3] + y3i * wr[3];
        }
      } else {
        /* w_3 = -1/2 + sgn i sqrt(3)/2 */
        for (int k = 0; k < BP; ++k) {
          const double tr = ar[1][k] + ar[2][k], ti = ai[1][k] + ai[2][k];
          const double dr = ar[1][k] - ar[2][k], di = ai[1][k] - ai[2][k];
          const double cr = ar[0][k] - 0.5 * tr, ci = ai[0][k] - 0.5 * ti;
          /* sgn i h (a1 - a2) */
          const double er = -sgn * h * di, ei = sgn * h * dr;
          const double y1r = cr + er, y1i = ci + ei;
          const double y2r = cr - er, y2i = ci - ei;
          br[0][k] = ar[0][k] + tr;
          bi[0][k] = ai[0][k] + ti;
          br[1][k] = y1r * wr[1] - y1i * wi[1];
          bi[1][k] = y1r * wi[1] + y1i * wr[1];
          br[2][k] = y2r * wr[2] - y2i * wi[2];
          bi[2][k] = y2r * wi[2] + y2i * wr[2];
        }
      }
    }
  }
}

/* Unnormalised DFT of length N (sgn = -1 forward, +1 inverse) of the BP-point
 * vector signal (re, im); the result is left in (re, im); (wr, wi) scratch. */
static void fft_run(int N, int nrad, const int* rad, int sgn, const double* tc, const double* ts, double* re,
                    double* im, double* wre, double* wim) {
  double *xr = re, *xi = im, *yr = wre, *yi = wim;
  int n = N, s = 1;
  for (int k = 0; k < nrad; ++k) {
    fft_pass(rad[k], n, s, N, sgn, tc, ts, xr, xi, yr, yi);
    n /= rad[k];
    s *= rad[k];
    double* t = xr; xr = yr; yr = t;
    t = xi; xi = yi; yi = t;
  }
  if (xr != re) {
    memcpy(re, xr, sizeof(double) * (size_t)N * BP);
    memcpy(im, xi, sizeof(double) * (size_t)N * BP);
  }
}

/* As sfx_cpu_smb_cross_prod_gauss, with the transforms done by the FFT on
 * N = 3 * 2^k, the smallest such N >= 3M - 2 (N >= 3). */
int sfx_cpu_smb_cross_prod_gauss_fft(const int32_t* jj, const double* rr, int64_t me, int64_t np,
                                     const int32_t* list_mode, int32_t M, const double* V_in,
                                     double* V_out, int32_t nthreads) {
  if (M <= 0 || me < 0 || np <= 0) return 1;
  int N = 3;
  while (N < 3 * M - 2) N *= 2;
  int rad[32];
  const int nrad = fft_radices(N, rad);
  if (nrad < 0) return 1;
  const int64_t G = 7 * me;
  double w[7][6], dd[7][6][2], pp[7];
  p2_tables(w, dd, pp);
  double* tc = (double*)malloc(sizeof(double) * N);
  double* ts = (double*)malloc(sizeof(double) * N);
  if (!tc || !ts) { free(tc); free(ts); return 1; }
  const double two_pi = 6.283185307179586476925286766559;
  for (int j = 0; j < N; ++j) {
    tc[j] = cos(two_pi * j / N);
    ts[j] = sin(two_pi * j / N);
  }
  const int64_t nblk = (me + BE - 1) / BE;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    double* A = (double*)malloc(sizeof(double) * 12 * (size_t)M * BP);
    /* three complex signals (RotV r + i RotV t, RotV z + i W r, W t + i W z) */
    double* Zr = (double*)malloc(sizeof(double) * 3 * (size_t)N * BP);
    double* Zi = (double*)malloc(sizeof(double) * 3 * (size_t)N * BP);
    double* Sr = (double*)malloc(sizeof(double) * (size_t)N * BP);
    double* Si = (double*)malloc(sizeof(double) * (size_t)N * BP);
#pragma omp for schedule(dynamic, 4)
    for (int64_t b = 0; b < nblk; ++b) {
      const int64_t e0 = b * BE;
      const int ne = (int)((me - e0) < BE ? (me - e0) : BE);
      const int npt = 7 * ne;
      gauss_block(jj, rr, np, list_mode, M, V_in, (const double(*)[6])w, (const double(*)[6][2])dd, e0, ne, A);
      /* spectra: field f has X_0 = a_0, X_m = (a_m - i b_m)/2, X_{N-m} = conj;
       * Z = X_f1 + i X_f2 (fftpack-style packing of two real signals) */
      memset(Zr, 0, sizeof(double) * 3 * (size_t)N * BP);
      memset(Zi, 0, sizeof(double) * 3 * (size_t)N * BP);
      for (int c = 0; c < 3; ++c) {
        const int f1 = 2 * c, f2 = 2 * c + 1;   /* fields 0..5: RotV r,t,z then W r,t,z */
        const double* a1 = A + (size_t)(2 * f1) * M * BP;
        const double* b1 = A + (size_t)(2 * f1 + 1) * M * BP;
        const double* a2 = A + (size_t)(2 * f2) * M * BP;
        const double* b2 = A + (size_t)(2 * f2 + 1) * M * BP;
        double* zr = Zr + (size_t)c * N * BP;
        double* zi = Zi + (size_t)c * N * BP;
        for (int k = 0; k < BP; ++k) { zr[k] = a1[k]; zi[k] = a2[k]; }
        for (int m = 1; m < M; ++m) {
          double* pr = zr + (size_t)m * BP;
          double* pi = zi + (size_t)m * BP;
          double* qr = zr + (size_t)(N - m) * BP;
          double* qi = zi + (size_t)(N - m) * BP;
          const double* A1 = a1 + (size_t)m * BP;
          const double* B1 = b1 + (size_t)m * BP;
          const double* A2 = a2 + (size_t)m * BP;
          const double* B2 = b2 + (size_t)m * BP;
          for (int k = 0; k < BP; ++k) {
            /* X1 = (A1 - i B1)/2, X2 = (A2 - i B2)/2; Z_m = X1 + i X2, Z_{N-m} = conj X1 + i conj X2 */
            pr[k] = 0.5 * (A1[k] + B2[k]);
            pi[k] = 0.5 * (A2[k] - B1[k]);
            qr[k] = 0.5 * (A1[k] - B2[k]);
            qi[k] = 0.5 * (A2[k] + B1[k]);
          }
        }
        fft_run(N, nrad, rad, +1, tc, ts, zr, zi, Sr, Si);
      }
      /* the product on the grid: z0 = (Rr, Rt), z1 = (Rz, Wr), z2 = (Wt, Wz);
       * P = RotV x W goes to Sr + i Si = Pr + i Pt, and Zr[0] = Pz (imag 0) */
      {
        const double *R0 = Zr, *R1 = Zi, *R2 = Zr + (size_t)N * BP, *W0 = Zi + (size_t)N * BP;
        const double *W1 = Zr + (size_t)2 * N * BP, *W2 = Zi + (size_t)2 * N * BP;
        double* P2 = Zr + (size_t)2 * N * BP;   /* z2 is free once W t, W z are read */
        double* P2i = Zi + (size_t)2 * N * BP;
        for (size_t k = 0; k < (size_t)N * BP; ++k) {
          const double r0 = R0[k], r1 = R1[k], r2 = R2[k], w0 = W0[k], w1 = W1[k], w2 = W2[k];
          Sr[k] = r1 * w2 - r2 * w1;
          Si[k] = r2 * w0 - r0 * w2;
          P2[k] = r0 * w1 - r1 * w0;
          P2i[k] = 0.0;
        }
      }
      memcpy(Zr, Sr, sizeof(double) * (size_t)N * BP);
      memcpy(Zi, Si, sizeof(double) * (size_t)N * BP);
      fft_run(N, nrad, rad, -1, tc, ts, Zr, Zi, Sr, Si);
      fft_run(N, nrad, rad, -1, tc, ts, Zr + (size_t)2 * N * BP, Zi + (size_t)2 * N * BP, Sr, Si);
      /* unpack Y1 = (Z_m + conj Z_{N-m})/2, Y2 = (Z_m - conj Z_{N-m})/(2i);
       * a_m = 2 Re Y_m / N, b_m = -2 Im Y_m / N, a_0 = Y_0 / N, b_0 = 0 */
      for (int m = 0; m < M; ++m) {
        const double sc = (m == 0 ? 1.0 : 2.0) / N;
        const int mm = (m == 0) ? 0 : N - m;
        for (int c = 0; c < 3; ++c) {
          const double* zr = (c < 2) ? Zr : Zr + (size_t)2 * N * BP;
          const double* zi = (c < 2) ? Zi : Zi + (size_t)2 * N * BP;
          const double* pr = zr + (size_t)m * BP;
          const double* pi = zi + (size_t)m * BP;
          const double* qr = zr + (size_t)mm * BP;
          const double* qi = zi + (size_t)mm * BP;
          double* oa = V_out + 7 * e0 + G * (2 * c + 6 * (int64_t)m);
          double* ob = V_out + 7 * e0 + G * (2 * c + 1 + 6 * (int64_t)m);
          for (int p = 0; p < npt; ++p) {
            double yr, yi;
            if (c == 0) {            /* real part of the packed pair */
              yr = 0.5 * (pr[p] + qr[p]);
              yi = 0.5 * (pi[p] - qi[p]);
            } else if (c == 1) {     /* imaginary part of the packed pair */
              yr = 0.5 * (pi[p] + qi[p]);
              yi = -0.5 * (pr[p] - qr[p]);
            } else {
              yr = pr[p];
              yi = pi[p];
            }
            oa[p] = sc * yr;
            ob[p] = (m == 0) ? 0.0 : -sc * yi;
          }
        }
      }
    }
    free(A); free(Zr); free(Zi); free(Sr); free(Si);
  }
  free(tc); free(ts);
  return 0;
}
