// nnls.cpp -- non-negative least squares for the cubature fit.
//
// The reference fits cubature weights with the Lawson-Hanson active-set
// algorithm (NNLS_SOLVER, reference src/linearalgebra/NNLS.h:156-208, called
// from CUBATURE_GENERATOR_EIGEN::generateImportanceSampledCubature,
// src/cubature/CUBATURE_GENERATOR_EIGEN.cpp:654-663), capped at
// maxIter * n inner iterations (maxIter = 5 by default).  This is the same
// published algorithm (Lawson & Hanson, "Solving Least Squares Problems",
// ch. 23): the columns entering the passive set P are triangularised by
// Householder reflections applied to the whole matrix and right-hand side
// (so the dual vector of the remaining columns is read off the untouched
// rows), and a column leaving P is removed with Givens rotations.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "../../include/cfs_plume.h"

namespace {

// Reflection H = I - u u^T / (sigma * u0) that maps x[p..m) of a column to
// (-s sigma, 0, ...); returns u0 (u[p] after the update) or 0 when x[p..m) is 0.
double house_make(double* col, int p, int m, double* sig) {
  double scale = 0.0;
  for (int i = p; i < m; i++) scale = std::max(scale, std::fabs(col[i]));
  if (scale == 0.0) {
    *sig = 0.0;
    return 0.0;
  }
  double s = 0.0;
  for (int i = p; i < m; i++) {
    const double t = col[i] / scale;
    s += t * t;
  }
  double sigma = scale * std::sqrt(s);
  if (col[p] > 0.0) sigma = -sigma;
  const double u0 = col[p] - sigma;  // u = x - sigma e_p
  col[p] = sigma;                    // the transformed pivot
  *sig = sigma;
  return u0;
}

// Applies H (u0 at p, u[p+1..m) = v[p+1..m)) to y[p..m).
void house_apply(const double* v, double u0, double sigma, int p, int m, double* y) {
  if (u0 == 0.0 || sigma == 0.0) return;
  const double b = u0 * sigma;  // = -|u|^2 / 2 (negative)
  double d = u0 * y[p];
  for (int i = p + 1; i < m; i++) d += v[i] * y[i];
  const double f = d / b;
  y[p] += f * u0;
  for (int i = p + 1; i < m; i++) y[i] += f * v[i];
}

// Rotation (c, s) with c a + s b = sqrt(a^2 + b^2), -s a + c b = 0.
void givens(double a, double b, double* c, double* s, double* r) {
  if (std::fabs(a) > std::fabs(b)) {
    const double t = b / a, u = std::copysign(std::sqrt(1.0 + t * t), a);
    *c = 1.0 / u;
    *s = *c * t;
    *r = a * u;
  } else if (b != 0.0) {
    const double t = a / b, u = std::copysign(std::sqrt(1.0 + t * t), b);
    *s = 1.0 / u;
    *c = *s * t;
    *r = b * u;
  } else {
    *c = 1.0;
    *s = 0.0;
    *r = 0.0;
  }
}

void parallel_cols(int n, const std::function<void(int, int)>& fn) {
  const int nt = n >= 512 ? (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1;
  if (nt == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) {
    const int a = (int)((long)n * t / nt), e = (int)((long)n * (t + 1) / nt);
    th.emplace_back([&fn, a, e] { fn(a, e); });
  }
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

int cfs_nnls(const double* A_in, int m, int n, const double* b_in, int max_iter_factor, double* x, double* rnorm,
             int* iterations) {
  if (!A_in || !b_in || !x || m < 1 || n < 1 || max_iter_factor < 1) return CFS_ERR_ARG;
  std::vector<double> A;
  try {
    A.assign(A_in, A_in + (size_t)m * n);
  } catch (...) {
    return CFS_ERR_NOMEM;
  }
  std::vector<double> b(b_in, b_in + m), zz(m), w(n, 0.0);
  std::vector<int> index(n);
  for (int j = 0; j < n; j++) {
    index[j] = j;
    x[j] = 0.0;
  }
  auto col = [&](int j) { return A.data() + (size_t)j * m; };
  int nsetp = 0;  // index[0..nsetp) is P, index[nsetp..n) is Z
  const long itmax = (long)max_iter_factor * n;
  long iter = 0;
  int status = CFS_OK;

  // z[0..nsetp) = R^-1 b[0..nsetp) with R the leading columns of P
  auto solve_triangular = [&]() {
    for (int i = 0; i < nsetp; i++) zz[i] = b[i];
    for (int i = nsetp - 1; i >= 0; i--) {
      const double* c = col(index[i]);
      zz[i] /= c[i];
      for (int k = 0; k < i; k++) zz[k] -= c[k] * zz[i];
    }
  };

  for (;;) {
    if (nsetp >= n || nsetp >= m) break;
    // dual vector of the columns in Z: w_j = a_j[nsetp..m) . b[nsetp..m)
    parallel_cols(n - nsetp, [&](int a0, int a1) {
      for (int z = nsetp + a0; z < nsetp + a1; z++) {
        const double* c = col(index[z]);
        double s = 0.0;
        for (int i = nsetp; i < m; i++) s += c[i] * b[i];
        w[index[z]] = s;
      }
    });
    // pick the most positive dual entry whose column can enter P
    int chosen = -1;
    double u0 = 0.0, sigma = 0.0;
    for (;;) {
      int iz = -1;
      double wmax = 0.0;
      for (int z = nsetp; z < n; z++)
        if (w[index[z]] > wmax) {
          wmax = w[index[z]];
          iz = z;
        }
      if (iz < 0) break;  // no positive entry: the Kuhn-Tucker conditions hold
      const int j = index[iz];
      double* c = col(j);
      const double saved = c[nsetp];
      double unorm = 0.0;
      for (int i = 0; i < nsetp; i++) unorm += c[i] * c[i];
      unorm = std::sqrt(unorm);
      double sg = 0.0;
      const double uu = house_make(c, nsetp, m, &sg);
      // the new pivot must be significant next to the column's part in P's
      // rows, and the tentative new coefficient must be positive
      if (uu != 0.0 && unorm + std::fabs(c[nsetp]) * 0.01 > unorm) {
        for (int i = 0; i < m; i++) zz[i] = b[i];
        house_apply(c, uu, sg, nsetp, m, zz.data());
        if (zz[nsetp] / c[nsetp] > 0.0) {
          chosen = iz;
          u0 = uu;
          sigma = sg;
          break;
        }
      }
      c[nsetp] = saved;  // reject: restore and try the next candidate
      w[j] = 0.0;
    }
    if (chosen < 0) break;
    // move the column into P and apply its reflection to b and the columns of Z
    const int jt = index[chosen];
    std::swap(index[chosen], index[nsetp]);
    for (int i = 0; i < m; i++) b[i] = zz[i];
    {
      const double* v = col(jt);
      const int p = nsetp;
      parallel_cols(n - nsetp - 1, [&](int a0, int a1) {
        for (int z = p + 1 + a0; z < p + 1 + a1; z++) house_apply(v, u0, sigma, p, m, col(index[z]));
      });
      double* c = col(jt);
      for (int i = p + 1; i < m; i++) c[i] = 0.0;
    }
    nsetp++;
    w[jt] = 0.0;
    solve_triangular();

    // inner loop: step toward z until every coefficient in P is positive
    for (;;) {
      if (++iter > itmax) {
        status = 1;  // iteration cap reached (the reference's mode 3)
        break;
      }
      double alpha = 2.0;
      int jj = -1;
      for (int i = 0; i < nsetp; i++) {
        if (zz[i] <= 0.0) {
          const int j = index[i];
          const double t = -x[j] / (zz[i] - x[j]);
          if (alpha > t) {
            alpha = t;
            jj = i;
          }
        }
      }
      if (jj < 0) break;  // z is feasible
      for (int i = 0; i < nsetp; i++) {
        const int j = index[i];
        x[j] += alpha * (zz[i] - x[j]);
      }
      // move to Z every coefficient of P that reached zero (at least the one at jj)
      int i = jj;
      for (;;) {
        const int j = index[i];
        x[j] = 0.0;
        // delete position i of P: shift later columns down, restoring the
        // triangle with a rotation of rows (k, k+1) for each
        for (int k = i; k < nsetp - 1; k++) {
          double* ck = col(index[k + 1]);
          double c, s, r;
          givens(ck[k], ck[k + 1], &c, &s, &r);
          ck[k] = r;
          ck[k + 1] = 0.0;
          for (int z = 0; z < n; z++) {
            if (z == k + 1) continue;
            double* cz = col(index[z]);
            const double a0 = cz[k], a1 = cz[k + 1];
            cz[k] = c * a0 + s * a1;
            cz[k + 1] = -s * a0 + c * a1;
          }
          const double b0 = b[k], b1 = b[k + 1];
          b[k] = c * b0 + s * b1;
          b[k + 1] = -s * b0 + c * b1;
          std::swap(index[k], index[k + 1]);
        }
        nsetp--;
        // any other coefficient of P that is not positive leaves too
        i = -1;
        for (int q = 0; q < nsetp; q++)
          if (x[index[q]] <= 0.0) {
            i = q;
            break;
          }
        if (i < 0) break;
      }
      solve_triangular();
    }
    if (status) break;
    for (int i = 0; i < nsetp; i++) x[index[i]] = zz[i];
  }
  if (rnorm) {
    double s = 0.0;
    for (int i = nsetp; i < m; i++) s += b[i] * b[i];
    *rnorm = std::sqrt(s);
  }
  // a capped run (status 1) still leaves a non-negative x; the caller sees
  // the cap as *iterations > max_iter_factor * n
  if (iterations) *iterations = (int)std::min<long>(iter, 0x7fffffff);
  (void)status;
  return CFS_OK;
}

}  // extern "C"
