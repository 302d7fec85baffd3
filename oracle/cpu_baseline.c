/*
  Multi-core CPU baseline of the path: the oracle's restatement of the
  reference algorithms (tacs_oracle.c) run over all host cores with OpenMP,
  the way the reference's threaded path spreads its work
  (TACSAssembler_thread.cpp:149: worker threads take elements from a queue;
  its scatter is serialised by a mutex, here by an element colouring that
  lets every element of a colour scatter at once; BCSRMatMult3.cpp:28 and
  TACSBVec's BLAS-1 over row / entry ranges).

  TEST INFRASTRUCTURE ONLY, like the rest of this library: bench.py's
  cpu_baseline leg times it; the product path never calls it.  The serial
  functions of tacs_oracle.c stay the parity checker; these give the same
  results up to summation order (tests/test_oracle.py compares them).
*/
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void orc_solid_element(int order, const double *X, const double *u, const double C21[21],
                       double t, double rho, double alpha, double gamma, double *res,
                       double *mat);

int orc_par_threads(void) { return omp_get_max_threads(); }

/* Greedy element colouring: no two elements of one colour share a node, so
   a colour's elements scatter into the matrix and residual without races.
   Element e gets the lowest colour not yet taken by an element touching one
   of its nodes.  color_ptr[ncolors + 1] / color_elems[nelem] list the
   elements of each colour in ascending order.  Returns ncolors (at most 64),
   or -1 when 64 colours do not suffice. */
int orc_color_elements(int nelem, int npe, int nnodes, const int *conn, int *color_ptr,
                       int *color_elems) {
  uint64_t *used = (uint64_t *)calloc(nnodes > 0 ? nnodes : 1, sizeof(uint64_t));
  int *color = (int *)malloc((nelem > 0 ? nelem : 1) * sizeof(int));
  int ncolors = 0;
  for (int e = 0; e < nelem; e++) {
    uint64_t taken = 0;
    for (int a = 0; a < npe; a++) taken |= used[conn[(size_t)npe * e + a]];
    if (taken == ~(uint64_t)0) {
      free(used);
      free(color);
      return -1;
    }
    int c = 0;
    while (taken & ((uint64_t)1 << c)) c++;
    color[e] = c;
    if (c + 1 > ncolors) ncolors = c + 1;
    for (int a = 0; a < npe; a++) used[conn[(size_t)npe * e + a]] |= (uint64_t)1 << c;
  }
  for (int c = 0; c <= ncolors; c++) color_ptr[c] = 0;
  for (int e = 0; e < nelem; e++) color_ptr[color[e] + 1]++;
  for (int c = 0; c < ncolors; c++) color_ptr[c + 1] += color_ptr[c];
  int *pos = (int *)malloc((ncolors + 1) * sizeof(int));
  memcpy(pos, color_ptr, (ncolors + 1) * sizeof(int));
  for (int e = 0; e < nelem; e++) color_elems[pos[color[e]]++] = e;
  free(pos);
  free(used);
  free(color);
  return ncolors;
}

double orc_hex_quadrature(int order, int n, double pt[3]);
void orc_hex_basis(int order, const double pt[3], double *N, double *Nxi);

/* Element Jacobian of the isotropic solid as the baseline computes it: per
   quadrature point the physical basis gradients g_a, then every 3x3 block
   K_ab += w det(J) t [lam g_a g_b^T + mu g_b g_a^T + mu (g_a . g_b) I] (the
   nonzero coefficient pairs of TACSLinearElasticity3D::evalWeakMatrix,
   TACSLinearElasticity.cpp:1158, for an isotropic tangent).  Same matrix as
   orc_solid_element (tests/test_oracle.py), written for speed so the
   baseline is not slower than the reference's element loop.  mat (nv x nv,
   node-major dofs) is overwritten. */
void orc_iso_solid_jacobian(int order, const double *X, double lam, double mu, double t,
                            double *mat) {
  const int npe = order * order * order, nv = 3 * npe;
  memset(mat, 0, (size_t)nv * nv * sizeof(double));
  for (int q = 0; q < npe; q++) {
    double pt[3], N[27], Nxi[81], g[81], Xd[9] = {0}, J[9];
    const double w = orc_hex_quadrature(order, q, pt);
    orc_hex_basis(order, pt, N, Nxi);
    for (int a = 0; a < npe; a++)
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) Xd[3 * i + k] += X[3 * a + i] * Nxi[3 * a + k];
    const double det = Xd[8] * (Xd[0] * Xd[4] - Xd[3] * Xd[1]) -
                       Xd[7] * (Xd[0] * Xd[5] - Xd[3] * Xd[2]) +
                       Xd[6] * (Xd[1] * Xd[5] - Xd[2] * Xd[4]);
    const double r = 1.0 / det;
    J[0] = (Xd[4] * Xd[8] - Xd[5] * Xd[7]) * r;
    J[1] = -(Xd[1] * Xd[8] - Xd[2] * Xd[7]) * r;
    J[2] = (Xd[1] * Xd[5] - Xd[2] * Xd[4]) * r;
    J[3] = -(Xd[3] * Xd[8] - Xd[5] * Xd[6]) * r;
    J[4] = (Xd[0] * Xd[8] - Xd[2] * Xd[6]) * r;
    J[5] = -(Xd[0] * Xd[5] - Xd[2] * Xd[3]) * r;
    J[6] = (Xd[3] * Xd[7] - Xd[4] * Xd[6]) * r;
    J[7] = -(Xd[0] * Xd[7] - Xd[1] * Xd[6]) * r;
    J[8] = (Xd[0] * Xd[4] - Xd[1] * Xd[3]) * r;
    const double f = w * det * t;
    for (int a = 0; a < npe; a++)
      for (int d = 0; d < 3; d++)
        g[3 * a + d] = Nxi[3 * a] * J[d] + Nxi[3 * a + 1] * J[3 + d] + Nxi[3 * a + 2] * J[6 + d];
    for (int a = 0; a < npe; a++) {
      const double ga0 = f * g[3 * a], ga1 = f * g[3 * a + 1], ga2 = f * g[3 * a + 2];
      const double la[3] = {lam * ga0, lam * ga1, lam * ga2};
      double *row0 = &mat[(size_t)(3 * a) * nv];
      for (int b = 0; b < npe; b++) {
        const double gb0 = g[3 * b], gb1 = g[3 * b + 1], gb2 = g[3 * b + 2];
        const double mb[3] = {mu * gb0, mu * gb1, mu * gb2};
        const double dd = mu * (ga0 * gb0 + ga1 * gb1 + ga2 * gb2);
        const double gb[3] = {gb0, gb1, gb2}, ga[3] = {ga0, ga1, ga2};
        for (int j = 0; j < 3; j++) {
          double *rj = row0 + (size_t)j * nv + 3 * b;
          for (int k = 0; k < 3; k++) rj[k] += la[j] * gb[k] + mb[j] * ga[k];
          rj[j] += dd;
        }
      }
    }
  }
}

static long find_slot_par(const int *rowp, const int *cols, int row, int col) {
  int lo = rowp[row], hi = rowp[row + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (cols[mid] == col) return mid;
    if (cols[mid] < col) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

/* orc_assemble_solid (the element loop of TACSAssembler::assembleJacobian,
   src/TACSAssembler.cpp:4352-4383, with BCSRMat::addRowValues' add rule)
   over the colours: within a colour the elements are spread over the
   threads.  A and res are accumulated.  Returns missing pattern entries. */
long orc_par_assemble_solid(int order, int ncolors, const int *color_ptr, const int *color_elems,
                            const int *conn, const double *X, const double *u,
                            const double *t_elem, const double C21[21], double rho,
                            double alpha, double gamma, const int *rowp, const int *cols,
                            double *A, double *res) {
  const int npe = order * order * order;
  const int nv = 3 * npe;
  long missing = 0;
  // isotropic C[21]: the matrix-only assembly takes the (lambda, mu) kernel
  const double lam = C21[1], mu = C21[15];
  int iso = C21[0] == C21[6] && C21[0] == C21[11] && C21[1] == C21[2] && C21[1] == C21[7] &&
            C21[15] == C21[18] && C21[15] == C21[20] && fabs(C21[0] - lam - 2.0 * mu) <= 1e-12 * C21[0];
  for (int k = 0; k < 21; k++)
    if (k != 0 && k != 1 && k != 2 && k != 6 && k != 7 && k != 11 && k != 15 && k != 18 &&
        k != 20 && C21[k] != 0.0)
      iso = 0;
  for (int c = 0; c < ncolors; c++) {
#pragma omp parallel reduction(+ : missing)
    {
      double Xe[81], ue[81], re[81];
      double *ke = (double *)malloc((size_t)nv * nv * sizeof(double));
#pragma omp for schedule(dynamic, 64)
      for (int k = color_ptr[c]; k < color_ptr[c + 1]; k++) {
        const int e = color_elems[k];
        const int *cn = &conn[(size_t)npe * e];
        for (int a = 0; a < npe; a++)
          for (int d = 0; d < 3; d++) {
            Xe[3 * a + d] = X[3 * (size_t)cn[a] + d];
            ue[3 * a + d] = u ? u[3 * (size_t)cn[a] + d] : 0.0;
          }
        memset(re, 0, nv * sizeof(double));
        if (A && iso && !res && gamma == 0.0 && alpha == 1.0) {
          orc_iso_solid_jacobian(order, Xe, lam, mu, t_elem[e], ke);
        } else {
          if (A) memset(ke, 0, (size_t)nv * nv * sizeof(double));
          orc_solid_element(order, Xe, ue, C21, t_elem[e], rho, alpha, gamma, res ? re : NULL,
                            A ? ke : NULL);
        }
        if (res)
          for (int a = 0; a < npe; a++)
            for (int d = 0; d < 3; d++) res[3 * (size_t)cn[a] + d] += re[3 * a + d];
        if (A)
          for (int a = 0; a < npe; a++)
            for (int b = 0; b < npe; b++) {
              long s = find_slot_par(rowp, cols, cn[a], cn[b]);
              if (s < 0) {
                missing++;
                continue;
              }
              double *blk = &A[9 * s];
              for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++) blk[3 * r + q] += ke[(3 * a + r) * nv + 3 * b + q];
            }
      }
      free(ke);
    }
  }
  return missing;
}

/* y = A x (BCSRMatVecMult3/6 src/bpmat/BCSRMatMult3.cpp:28), rows over the
   threads, each row summed in storage order as the serial oracle does. */
void orc_par_bsr_mult(int b, int nrows, const int *rowp, const int *cols, const double *A,
                      const double *x, double *y) {
  if (b == 3) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nrows; i++) {
      double y0 = 0.0, y1 = 0.0, y2 = 0.0;
      for (int j = rowp[i]; j < rowp[i + 1]; j++) {
        const double *a = &A[9 * (size_t)j];
        const double *xx = &x[3 * (size_t)cols[j]];
        y0 += (a[0] * xx[0] + a[1] * xx[1]) + a[2] * xx[2];
        y1 += (a[3] * xx[0] + a[4] * xx[1]) + a[5] * xx[2];
        y2 += (a[6] * xx[0] + a[7] * xx[1]) + a[8] * xx[2];
      }
      y[3 * (size_t)i] = y0;
      y[3 * (size_t)i + 1] = y1;
      y[3 * (size_t)i + 2] = y2;
    }
    return;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nrows; i++) {
    double acc[8] = {0};
    for (int j = rowp[i]; j < rowp[i + 1]; j++) {
      const double *a = &A[(size_t)b * b * j];
      const double *xx = &x[(size_t)b * cols[j]];
      for (int r = 0; r < b; r++) {
        double v = 0.0;
        for (int c = 0; c < b; c++) v += a[b * r + c] * xx[c];
        acc[r] += v;
      }
    }
    for (int r = 0; r < b; r++) y[(size_t)b * i + r] = acc[r];
  }
}

static double pdot(size_t n, const double *x, const double *y) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (size_t i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}
static void paxpy(size_t n, double a, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) y[i] += a * x[i];
}
static void pscale(size_t n, double a, double *x) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) x[i] *= a;
}
static void ppc(int b, int nrows, const double *Dinv, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nrows; i++) {
    const double *d = &Dinv[(size_t)b * b * i];
    for (int r = 0; r < b; r++) {
      double v = 0.0;
      for (int c = 0; c < b; c++) v += d[b * r + c] * x[(size_t)b * i + c];
      y[(size_t)b * i + r] = v;
    }
  }
}

/* orc_gmres_capped (GMRES::solve src/bpmat/KSM.cpp:785-958 with
   ModifiedGramSchmidt :526 and block-Jacobi) with its vector operations and
   products over the threads: the same recurrence, step for step. */
int orc_par_gmres(int b, int nrows, const int *rowp, const int *cols, const double *A,
                  const double *Dinv, const double *rhs, double *x, int m, int nrestart,
                  double rtol, double atol, int zero_guess, int max_its, int *iters,
                  double *resnorm_out) {
  size_t n = (size_t)b * nrows;
  double **W = (double **)malloc((m + 1) * sizeof(double *));
  for (int i = 0; i <= m; i++) W[i] = (double *)malloc(n * sizeof(double));
  double *work = (double *)malloc(n * sizeof(double));
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *res = (double *)calloc(m + 1, sizeof(double));
  double *Qc = (double *)calloc(m, sizeof(double));
  double *Qs = (double *)calloc(m, sizeof(double));
  double rhs_norm = 0.0, resNorm = 0.0;
  int solve_flag = 0, total = 0;
  for (int count = 0; count < nrestart + 1; count++) {
    if (zero_guess && count == 0) {
      memset(x, 0, n * sizeof(double));
      memcpy(W[0], rhs, n * sizeof(double));
      res[0] = sqrt(pdot(n, W[0], W[0]));
      pscale(n, 1.0 / res[0], W[0]);
    } else {
      orc_par_bsr_mult(b, nrows, rowp, cols, A, x, W[0]);
      paxpy(n, -1.0, rhs, W[0]);
      res[0] = sqrt(pdot(n, W[0], W[0]));
      pscale(n, -1.0 / res[0], W[0]);
    }
    if (count == 0) {
      rhs_norm = res[0];
      resNorm = rhs_norm;
    }
    int niters = 0;
    if (res[0] < atol) {
      solve_flag = 1;
      break;
    }
    for (int i = 0; i < m; i++) {
      double *h = &H[(size_t)(m + 1) * i];
      ppc(b, nrows, Dinv, W[i], work);
      orc_par_bsr_mult(b, nrows, rowp, cols, A, work, W[i + 1]);
      for (int j = 0; j <= i; j++) {
        h[j] = pdot(n, W[j], W[i + 1]);
        paxpy(n, -h[j], W[j], W[i + 1]);
      }
      h[i + 1] = sqrt(pdot(n, W[i + 1], W[i + 1]));
      pscale(n, 1.0 / h[i + 1], W[i + 1]);
      for (int k = 0; k < i; k++) {
        double h1 = h[k], h2 = h[k + 1];
        h[k] = h1 * Qc[k] + h2 * Qs[k];
        h[k + 1] = -h1 * Qs[k] + h2 * Qc[k];
      }
      double h1 = h[i], h2 = h[i + 1];
      double sq = sqrt(h1 * h1 + h2 * h2);
      Qc[i] = h1 / sq;
      Qs[i] = h2 / sq;
      h[i] = h1 * Qc[i] + h2 * Qs[i];
      h[i + 1] = -h1 * Qs[i] + h2 * Qc[i];
      h1 = res[i];
      res[i] = h1 * Qc[i];
      res[i + 1] = -h1 * Qs[i];
      niters++;
      resNorm = fabs(res[i + 1]);
      if (resNorm < atol || resNorm < rtol * rhs_norm) {
        solve_flag = 1;
        break;
      }
      if (max_its > 0 && total + niters >= max_its) break;
    }
    total += niters;
    for (int i = niters - 1; i >= 0; i--) {
      for (int j = i + 1; j < niters; j++) res[i] -= H[(size_t)(m + 1) * j + i] * res[j];
      res[i] /= H[(size_t)(m + 1) * i + i];
    }
    memset(work, 0, n * sizeof(double));
    for (int i = 0; i < niters; i++) paxpy(n, res[i], W[i], work);
    ppc(b, nrows, Dinv, work, W[0]);
    paxpy(n, 1.0, W[0], x);
    if (solve_flag || (max_its > 0 && total >= max_its)) break;
  }
  if (iters) *iters = total;
  if (resnorm_out) *resnorm_out = resNorm;
  for (int i = 0; i <= m; i++) free(W[i]);
  free(W);
  free(work);
  free(H);
  free(res);
  free(Qc);
  free(Qs);
  return solve_flag;
}

/* The oracle's element Jacobian over nelem elements (3 npe coordinates
   each) in one loop, for the speed ratio against the reference's element
   loop on the same inputs; mat receives the last element's matrix. */
void orc_solid_jacobian_loop(int order, int nelem, const double *Xe, const double C21[21],
                             double t, double *res, double *mat) {
  const int npe = order * order * order, nv = 3 * npe;
  double u[81] = {0};
  for (int e = 0; e < nelem; e++) {
    memset(res, 0, nv * sizeof(double));
    memset(mat, 0, (size_t)nv * nv * sizeof(double));
    orc_solid_element(order, &Xe[(size_t)3 * npe * e], u, C21, t, 0.0, 1.0, 0.0, res, mat);
  }
}

void orc_quad4_element(const double X[12], const double v[24], double E, double nu, double t,
                       double alpha, double *res, double *mat);

/* The oracle's Quad4 shell element Jacobian over nelem elements (12
   coordinates each), for the speed ratio against the reference's loop. */
void orc_quad4_jacobian_loop(int nelem, const double *X, double E, double nu, double t,
                             double *res, double *mat) {
  double v[24] = {0};
  for (int e = 0; e < nelem; e++) {
    memset(res, 0, 24 * sizeof(double));
    memset(mat, 0, 576 * sizeof(double));
    orc_quad4_element(&X[(size_t)12 * e], v, E, nu, t, 1.0, res, mat);
  }
}

/* orc_assemble_shell (TACSAssembler::assembleJacobian's element loop over
   TACSQuad4Shell, 6x6 blocks) over the colours, elements of a colour spread
   over the threads; A and res accumulated. */
long orc_par_assemble_shell(int ncolors, const int *color_ptr, const int *color_elems,
                            const int *conn, const double *X, const double *u,
                            const double *t_elem, double E, double nu, const int *rowp,
                            const int *cols, double *A, double *res) {
  long missing = 0;
  for (int c = 0; c < ncolors; c++) {
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : missing)
    for (int k = color_ptr[c]; k < color_ptr[c + 1]; k++) {
      const int e = color_elems[k];
      const int *cn = &conn[4 * (size_t)e];
      double Xe[12], ue[24], re[24] = {0}, ke[576];
      for (int a = 0; a < 4; a++) {
        for (int d = 0; d < 3; d++) Xe[3 * a + d] = X[3 * (size_t)cn[a] + d];
        for (int d = 0; d < 6; d++) ue[6 * a + d] = u ? u[6 * (size_t)cn[a] + d] : 0.0;
      }
      if (A) memset(ke, 0, sizeof(ke));
      orc_quad4_element(Xe, ue, E, nu, t_elem[e], 1.0, res ? re : NULL, A ? ke : NULL);
      if (res)
        for (int a = 0; a < 4; a++)
          for (int d = 0; d < 6; d++) res[6 * (size_t)cn[a] + d] += re[6 * a + d];
      if (A)
        for (int a = 0; a < 4; a++)
          for (int b = 0; b < 4; b++) {
            long s = find_slot_par(rowp, cols, cn[a], cn[b]);
            if (s < 0) {
              missing++;
              continue;
            }
            double *blk = &A[36 * s];
            for (int r = 0; r < 6; r++)
              for (int q = 0; q < 6; q++) blk[6 * r + q] += ke[(6 * a + r) * 24 + 6 * b + q];
          }
    }
  }
  return missing;
}
