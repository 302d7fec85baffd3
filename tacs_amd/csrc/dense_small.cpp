// Small dense host-side linear algebra for the deflated-restart GMRES: the
// eigenvalues of a real n x n matrix (Householder reduction to Hessenberg
// form, then implicit double-shift Francis QR steps -- Golub & Van Loan,
// Matrix Computations, Algorithms 7.4.2 and 7.5.1/7.5.2), eigenvectors of
// selected eigenvalues by inverse iteration, and helpers for least squares
// with Givens rotations.  Sizes are at most a few hundred, so plain loops.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <complex>
#include <vector>

#include "../../include/tacs_hip.h"

namespace {

typedef std::complex<double> cplx;

// Householder vector of x[0..len): (I - beta v v^T) x = alpha e_0, v[0] free
inline void house(int len, const double* x, double* v, double& beta) {
  double s = 0.0;
  for (int i = 1; i < len; i++) s += x[i] * x[i];
  for (int i = 0; i < len; i++) v[i] = x[i];
  if (s == 0.0) {
    beta = 0.0;
    return;
  }
  double nrm = sqrt(x[0] * x[0] + s);
  v[0] = (x[0] <= 0.0) ? x[0] - nrm : -s / (x[0] + nrm);
  beta = 2.0 * v[0] * v[0] / (s + v[0] * v[0]);
  double v0 = v[0];
  for (int i = 0; i < len; i++) v[i] /= v0;
}

// Reduce A (row-major n x n) to upper Hessenberg form in place by similarity;
// when Q is not NULL it receives the orthogonal Q (row-major) with
// A_in = Q A_out Q^T.
void hessenberg(int n, double* A, double* Q = NULL) {
  std::vector<double> x(n), v(n);
  if (Q)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) Q[i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int k = 0; k + 2 < n; k++) {
    int len = n - k - 1;
    for (int i = 0; i < len; i++) x[i] = A[(k + 1 + i) * n + k];
    double beta;
    house(len, x.data(), v.data(), beta);
    if (beta == 0.0) continue;
    // left: rows k+1..n-1, all columns k..n-1
    for (int j = k; j < n; j++) {
      double s = 0.0;
      for (int i = 0; i < len; i++) s += v[i] * A[(k + 1 + i) * n + j];
      s *= beta;
      for (int i = 0; i < len; i++) A[(k + 1 + i) * n + j] -= s * v[i];
    }
    // right: all rows, columns k+1..n-1
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int j = 0; j < len; j++) s += A[i * n + k + 1 + j] * v[j];
      s *= beta;
      for (int j = 0; j < len; j++) A[i * n + k + 1 + j] -= s * v[j];
    }
    for (int i = k + 2; i < n; i++) A[i * n + k] = 0.0;
    if (Q) {
      // Q <- Q (I - beta v v^T) on columns k+1..n-1
      for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < len; j++) s += Q[i * n + k + 1 + j] * v[j];
        s *= beta;
        for (int j = 0; j < len; j++) Q[i * n + k + 1 + j] -= s * v[j];
      }
    }
  }
}

// Eigenvalues of the 2 x 2 block [a b; c d]
inline void eig2(double a, double b, double c, double d, double& r1, double& i1, double& r2,
                 double& i2) {
  double p = 0.5 * (a - d), q = p * p + b * c, mu = 0.5 * (a + d);
  if (q >= 0.0) {
    double z = sqrt(q);
    double l1 = mu + (p >= 0.0 ? z : -z);
    double det = a * d - b * c;
    double l2 = (l1 != 0.0) ? det / l1 : mu - (p >= 0.0 ? z : -z);
    r1 = l1;
    r2 = l2;
    i1 = i2 = 0.0;
  } else {
    r1 = r2 = mu;
    i1 = sqrt(-q);
    i2 = -i1;
  }
}

// Eigenvalues of an upper Hessenberg matrix (destroyed) by implicit
// double-shift QR on the active window [lo, hi].  Returns 0 on success.
int hqr_eigvals(int n, double* H, double* wr, double* wi) {
  auto h = [&](int i, int j) -> double& { return H[i * n + j]; };
  double anorm = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = std::max(0, i - 1); j < n; j++) anorm += fabs(h(i, j));
  if (anorm == 0.0) anorm = 1.0;
  const double eps = 2.220446049250313e-16;
  int hi = n - 1, iter = 0, total = 0;
  while (hi >= 0) {
    int lo = hi;
    while (lo > 0) {
      double s = fabs(h(lo - 1, lo - 1)) + fabs(h(lo, lo));
      if (s == 0.0) s = anorm;
      if (fabs(h(lo, lo - 1)) <= eps * s) {
        h(lo, lo - 1) = 0.0;
        break;
      }
      lo--;
    }
    if (lo == hi) {
      wr[hi] = h(hi, hi);
      wi[hi] = 0.0;
      hi--;
      iter = 0;
      continue;
    }
    if (lo == hi - 1) {
      eig2(h(hi - 1, hi - 1), h(hi - 1, hi), h(hi, hi - 1), h(hi, hi), wr[hi - 1], wi[hi - 1],
           wr[hi], wi[hi]);
      hi -= 2;
      iter = 0;
      continue;
    }
    if (++total > 60 * n) return 1;
    iter++;
    // shifts from the trailing 2 x 2 block (exceptional shift every 10 steps)
    double s = h(hi - 1, hi - 1) + h(hi, hi);
    double t = h(hi - 1, hi - 1) * h(hi, hi) - h(hi - 1, hi) * h(hi, hi - 1);
    if (iter % 10 == 0) {
      double w = fabs(h(hi, hi - 1)) + fabs(h(hi - 1, hi - 2));
      s = 1.5 * w;
      t = w * w;
    }
    double x = h(lo, lo) * h(lo, lo) + h(lo, lo + 1) * h(lo + 1, lo) - s * h(lo, lo) + t;
    double y = h(lo + 1, lo) * (h(lo, lo) + h(lo + 1, lo + 1) - s);
    double z = h(lo + 1, lo) * h(lo + 2, lo + 1);
    for (int k = lo; k <= hi - 2; k++) {
      double xv[3] = {x, y, z}, v[3], beta;
      house(3, xv, v, beta);
      if (beta != 0.0) {
        int q = std::max(lo, k - 1);
        for (int j = q; j <= hi; j++) {
          double sm = beta * (v[0] * h(k, j) + v[1] * h(k + 1, j) + v[2] * h(k + 2, j));
          h(k, j) -= sm * v[0];
          h(k + 1, j) -= sm * v[1];
          h(k + 2, j) -= sm * v[2];
        }
        int r = std::min(k + 3, hi);
        for (int i = lo; i <= r; i++) {
          double sm = beta * (h(i, k) * v[0] + h(i, k + 1) * v[1] + h(i, k + 2) * v[2]);
          h(i, k) -= sm * v[0];
          h(i, k + 1) -= sm * v[1];
          h(i, k + 2) -= sm * v[2];
        }
      }
      x = h(k + 1, k);
      y = h(k + 2, k);
      if (k < hi - 2) z = h(k + 3, k);
    }
    double xv[2] = {x, y}, v[2], beta;
    house(2, xv, v, beta);
    if (beta != 0.0) {
      for (int j = hi - 2; j <= hi; j++) {
        double sm = beta * (v[0] * h(hi - 1, j) + v[1] * h(hi, j));
        h(hi - 1, j) -= sm * v[0];
        h(hi, j) -= sm * v[1];
      }
      for (int i = lo; i <= hi; i++) {
        double sm = beta * (h(i, hi - 1) * v[0] + h(i, hi) * v[1]);
        h(i, hi - 1) -= sm * v[0];
        h(i, hi) -= sm * v[1];
      }
    }
  }
  return 0;
}

// Inverse iteration on an upper Hessenberg H (row-major n x n) at lambda
// (complex): LU of H - lambda' I with partial pivoting between neighbouring
// rows (the only rows with entries in a pivot column), O(n^2) per vector
// instead of the O(n^3) of a dense factorisation; lambda' = lambda slightly
// perturbed, four solves from a fixed start vector s of the original basis
// (Q^T s here, so the iterates are those of the dense iteration in that basis).
void inverse_iteration_hess(int n, const double* Hh, const double* Q, cplx lambda, double gnorm,
                            std::vector<cplx>& v) {
  std::vector<cplx> M((size_t)n * n);
  cplx lam = lambda + cplx(1e-10 * gnorm, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      M[(size_t)i * n + j] = (j >= i - 1) ? Hh[(size_t)i * n + j] - (i == j ? lam : 0.0) : 0.0;
  std::vector<cplx> fac(n, cplx(0.0, 0.0));
  std::vector<char> swp(n, 0);
  for (int k = 0; k + 1 < n; k++) {
    cplx* rk = &M[(size_t)k * n];
    cplx* r1 = &M[(size_t)(k + 1) * n];
    if (std::abs(r1[k]) > std::abs(rk[k])) {
      for (int j = k; j < n; j++) std::swap(rk[j], r1[j]);
      swp[k] = 1;
    }
    if (std::abs(rk[k]) == 0.0) rk[k] = cplx(1e-14 * gnorm, 0.0);
    cplx f = r1[k] / rk[k];
    fac[k] = f;
    r1[k] = 0.0;
    for (int j = k + 1; j < n; j++) r1[j] -= f * rk[j];
  }
  if (std::abs(M[(size_t)(n - 1) * n + n - 1]) == 0.0)
    M[(size_t)(n - 1) * n + n - 1] = cplx(1e-14 * gnorm, 0.0);
  // the fixed start vector s of the original basis, as Q^T s
  std::vector<cplx> s0(n);
  for (int i = 0; i < n; i++) s0[i] = cplx(1.0 + 0.1 * ((i * 7) % 11), 0.05 * (i % 3));
  v.assign(n, cplx(0.0, 0.0));
  for (int c = 0; c < n; c++) {
    cplx acc(0.0, 0.0);
    for (int r = 0; r < n; r++) acc += Q[(size_t)r * n + c] * s0[r];
    v[c] = acc;
  }
  for (int it = 0; it < 4; it++) {
    for (int k = 0; k + 1 < n; k++) {
      if (swp[k]) std::swap(v[k], v[k + 1]);
      v[k + 1] -= fac[k] * v[k];
    }
    for (int i = n - 1; i >= 0; i--) {
      cplx s = v[i];
      for (int j = i + 1; j < n; j++) s -= M[(size_t)i * n + j] * v[j];
      v[i] = s / M[(size_t)i * n + i];
    }
    double nrm = 0.0;
    for (int i = 0; i < n; i++) nrm += std::norm(v[i]);
    nrm = sqrt(nrm);
    for (int i = 0; i < n; i++) v[i] /= nrm;
  }
}

}  // namespace

extern "C" {

// Eigenvalues of G (row-major n x n), sorted by ascending modulus, and a real
// basis V (column-major n x kout) of the eigenvectors of the k eigenvalues of
// smallest modulus (a complex pair contributes its real and imaginary parts;
// a pair split by the cut is kept whole, so kout <= k + 1).  Returns kout, or
// -1 when the QR iteration fails.
int tacs_dense_smallest_eigvecs(int n, const double* G, int k, double* V, double* wr,
                                double* wi) {
  if (n <= 0) return 0;
  std::vector<double> H(G, G + (size_t)n * n), Q((size_t)n * n);
  hessenberg(n, H.data(), Q.data());
  const std::vector<double> Hh(H);  // the QR iteration overwrites H
  std::vector<double> er(n), ei(n);
  if (hqr_eigvals(n, H.data(), er.data(), ei.data())) return -1;
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    double ma = std::hypot(er[a], ei[a]), mb = std::hypot(er[b], ei[b]);
    if (ma != mb) return ma < mb;
    return ei[a] > ei[b];  // conjugate pairs: positive imaginary part first
  });
  if (wr && wi)
    for (int i = 0; i < n; i++) {
      wr[i] = er[idx[i]];
      wi[i] = ei[idx[i]];
    }
  double gnorm = 0.0;
  for (size_t i = 0; i < (size_t)n * n; i++) gnorm = std::max(gnorm, fabs(G[i]));
  if (gnorm == 0.0) gnorm = 1.0;
  int kout = 0;
  std::vector<cplx> v, vh;
  for (int s = 0; s < n && kout < k; s++) {
    int i = idx[s];
    if (ei[i] < 0.0) continue;  // the conjugate of a pair already taken
    // eigenvector of the Hessenberg form, then G's: v = Q v_h
    inverse_iteration_hess(n, Hh.data(), Q.data(), cplx(er[i], ei[i]), gnorm, vh);
    v.assign(n, cplx(0.0, 0.0));
    for (int r = 0; r < n; r++) {
      cplx acc(0.0, 0.0);
      for (int c = 0; c < n; c++) acc += Q[(size_t)r * n + c] * vh[c];
      v[r] = acc;
    }
    if (ei[i] == 0.0) {
      // real eigenvector: rotate the phase so the largest entry is real
      int p = 0;
      for (int j = 1; j < n; j++)
        if (std::abs(v[j]) > std::abs(v[p])) p = j;
      cplx ph = std::conj(v[p]) / std::abs(v[p]);
      for (int j = 0; j < n; j++) V[(size_t)kout * n + j] = (v[j] * ph).real();
      kout++;
    } else {
      for (int j = 0; j < n; j++) V[(size_t)kout * n + j] = v[j].real();
      kout++;
      for (int j = 0; j < n; j++) V[(size_t)kout * n + j] = v[j].imag();
      kout++;
    }
  }
  return kout;
}

}  // extern "C"
