/*
  CPU restatement of TACS's solid-element assembly / block-CSR / GMRES /
  adjoint path, in plain C and FP64.

  TEST INFRASTRUCTURE ONLY.  This library is the parity oracle for the HIP
  implementation in tacs_amd/: it may be loaded by tests/, by
  __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and by
  nothing else.  The product path never calls it.

  Every routine cites the reference routine it restates (paths relative to
  the TACS source tree).  Pins (each checked in tests/test_oracle_*.py):
    - orc_hex_quadrature / orc_hex_basis: closed forms (partition of unity,
      exact integration of polynomials, literal Gauss constants
      src/elements/basis/TACSGaussQuadrature.h:26,29).
    - orc_iso_C21: closed form of isotropic Hooke's law; Voigt round trip.
    - orc_solid_geo_element / orc_assemble_geo: the reference's own
      TACSElement3D::getMatType geometric stiffness (oracle/_ref); symmetry;
      the closed form G_ab = (grad N_a . sigma . grad N_b) I for a uniform
      stress state.
    - orc_solid_matvec_element / orc_matfree_product: the reference's own
      TACSElement3D matrix-free product (oracle/_ref) and K x from the
      assembled element matrices of orc_solid_element (mass and stiffness).
    - orc_solid_element: symmetric Ke, rigid-body null space, patch test,
      residual = Ke*u, finite differences of the residual, golden compliance
      of tests/integration_tests/test_elast_linhexa_element_3d.py:11-27.
    - orc_node_csr / orc_assemble_solid / BC routines: brute-force dense
      assembly on tiny meshes, golden compliance and KS above.
    - orc_bsr_mult: dense matvec.
    - orc_gmres: dense LU solve (numpy) on small systems.
    - functions and sensitivities: golden compliance / KS values, central
      finite differences, adjoint identity dC/dt_e = -u_e^T K0_e u_e.
*/
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Gauss points: the reference uses these literals, not 1/sqrt(3), sqrt(0.6)
   (src/elements/basis/TACSGaussQuadrature.h:26-30). */
static const double gpts2[2] = {-0.577350269189626, 0.577350269189626};
static const double gwts2[2] = {1.0, 1.0};
static const double gpts3[3] = {-0.774596669241483, 0.0, 0.774596669241483};
static const double gwts3[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};

/* 1D Lagrange shape functions on the Gauss-Lobatto knots {-1,1} (order 2)
   and {-1,0,1} (order 3): src/elements/basis/TACSHexaBasis.cpp:139-200,
   :392-460 */
static void lagr1d(int order, double x, double *n, double *d) {
  if (order == 2) {
    n[0] = 0.5 * (1.0 - x);
    n[1] = 0.5 * (1.0 + x);
    d[0] = -0.5;
    d[1] = 0.5;
  } else {
    n[0] = -0.5 * x * (1.0 - x);
    n[1] = (1.0 - x) * (1.0 + x);
    n[2] = 0.5 * (1.0 + x) * x;
    d[0] = -0.5 + x;
    d[1] = -2.0 * x;
    d[2] = 0.5 + x;
  }
}

/* Quadrature point n of the order^3 tensor Gauss rule, x-fastest
   (TACSHexaBasis.cpp:96-110 linear, :337-351 quadratic). */
double orc_hex_quadrature(int order, int n, double pt[3]) {
  const double *p = (order == 2 ? gpts2 : gpts3);
  const double *w = (order == 2 ? gwts2 : gwts3);
  int i = n % order, j = (n / order) % order, k = n / (order * order);
  pt[0] = p[i];
  pt[1] = p[j];
  pt[2] = p[k];
  return w[i] * w[j] * w[k];
}

/* Face quadrature point: face/2 selects the fixed parametric direction,
   face%2 its sign; TACSHexaBasis.cpp:116-135 and tangents :21-70. */
double orc_hex_face_quadrature(int order, int face, int n, double pt[3],
                               double t[6]) {
  const double *p = (order == 2 ? gpts2 : gpts3);
  const double *w = (order == 2 ? gwts2 : gwts3);
  int a = n % order, b = n / order;
  double fixed = -1.0 + 2.0 * (face % 2);
  if (face / 2 == 0) {
    pt[0] = fixed; pt[1] = p[a]; pt[2] = p[b];
  } else if (face / 2 == 1) {
    pt[0] = p[a]; pt[1] = fixed; pt[2] = p[b];
  } else {
    pt[0] = p[a]; pt[1] = p[b]; pt[2] = fixed;
  }
  /* The two tangents are ordered so that t1 x t2 is the outward normal. */
  static const double tang[6][6] = {
      {0, 0, 1, 0, 1, 0}, {0, 1, 0, 0, 0, 1}, {1, 0, 0, 0, 0, 1},
      {0, 0, 1, 1, 0, 0}, {0, 1, 0, 1, 0, 0}, {1, 0, 0, 0, 1, 0}};
  memcpy(t, tang[face], 6 * sizeof(double));
  return w[a] * w[b];
}

/* Tensor-product basis values N[npe] and parametric gradients Nxi[3*npe]
   (node a = i + order*j + order^2*k). */
void orc_hex_basis(int order, const double pt[3], double *N, double *Nxi) {
  double n1[3], d1[3], n2[3], d2[3], n3[3], d3[3];
  lagr1d(order, pt[0], n1, d1);
  lagr1d(order, pt[1], n2, d2);
  lagr1d(order, pt[2], n3, d3);
  int a = 0;
  for (int k = 0; k < order; k++) {
    for (int j = 0; j < order; j++) {
      for (int i = 0; i < order; i++, a++) {
        N[a] = n1[i] * n2[j] * n3[k];
        Nxi[3 * a] = d1[i] * n2[j] * n3[k];
        Nxi[3 * a + 1] = n1[i] * d2[j] * n3[k];
        Nxi[3 * a + 2] = n1[i] * n2[j] * d3[k];
      }
    }
  }
}

/* inv3x3 / det3x3: src/elements/TACSElementAlgebra.h:1811,1980 */
static double det3(const double A[9]) {
  return A[8] * (A[0] * A[4] - A[3] * A[1]) -
         A[7] * (A[0] * A[5] - A[3] * A[2]) +
         A[6] * (A[1] * A[5] - A[2] * A[4]);
}
static double inv3(const double A[9], double B[9]) {
  double det = det3(A);
  double r = 1.0 / det;
  B[0] = (A[4] * A[8] - A[5] * A[7]) * r;
  B[1] = -(A[1] * A[8] - A[2] * A[7]) * r;
  B[2] = (A[1] * A[5] - A[2] * A[4]) * r;
  B[3] = -(A[3] * A[8] - A[5] * A[6]) * r;
  B[4] = (A[0] * A[8] - A[2] * A[6]) * r;
  B[5] = -(A[0] * A[5] - A[2] * A[3]) * r;
  B[6] = (A[3] * A[7] - A[4] * A[6]) * r;
  B[7] = -(A[0] * A[7] - A[1] * A[6]) * r;
  B[8] = (A[0] * A[4] - A[1] * A[3]) * r;
  return det;
}

/* Isotropic packed tangent stiffness C[21]:
   src/constitutive/TACSMaterialProperties.cpp:270-292 (G = E/(2(1+nu)),
   :31). Packed upper triangle by rows. */
void orc_iso_C21(double E, double nu, double C[21]) {
  double D = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double G = 0.5 * E / (1.0 + nu);
  memset(C, 0, 21 * sizeof(double));
  C[0] = C[6] = C[11] = (1.0 - nu) * D;
  C[1] = C[2] = C[7] = nu * D;
  C[15] = C[18] = C[20] = G;
}

/* Unpack C[21] to a dense symmetric 6x6. */
static void unpack21(const double C[21], double D[36]) {
  int idx = 0;
  for (int i = 0; i < 6; i++) {
    for (int j = i; j < 6; j++, idx++) {
      D[6 * i + j] = C[idx];
      D[6 * j + i] = C[idx];
    }
  }
}

/* Voigt index of the symmetric tensor pair (i,j), order xx,yy,zz,yz,xz,xy
   (src/elements/TACSLinearElasticity.cpp:896-902). */
static int voigt(int i, int j) {
  if (i == j) return i;
  int s = i + j;
  return (s == 3 ? 3 : (s == 2 ? 4 : 5));
}

/* Strain (engineering shears) from the displacement gradient
   Ux[3*j + d] = du_j/dx_d (TACSLinearElasticity.cpp:896-902). */
static void strain(const double Ux[9], double e[6]) {
  e[0] = Ux[0];
  e[1] = Ux[4];
  e[2] = Ux[8];
  e[3] = Ux[5] + Ux[7];
  e[4] = Ux[2] + Ux[6];
  e[5] = Ux[1] + Ux[3];
}

static void matvec6(const double D[36], const double e[6], double s[6]) {
  for (int i = 0; i < 6; i++) {
    double v = 0.0;
    for (int j = 0; j < 6; j++) v += D[6 * i + j] * e[j];
    s[i] = v;
  }
}

/* Per-quadrature-point geometry and field gradients, following
   TACSElementBasis::getFieldGradient (basis/TACSElementBasis.cpp:266-320):
   Xd = dX/dxi, J = Xd^{-1}, Nx[3a+d] = sum_k Nxi[3a+k] J[3k+d]. Returns
   detXd. */
static double qp_geometry(int npe, const double *Nxi, const double *X,
                          double *Nx, double J[9]) {
  double Xd[9] = {0};
  for (int a = 0; a < npe; a++) {
    for (int i = 0; i < 3; i++) {
      for (int k = 0; k < 3; k++) Xd[3 * i + k] += X[3 * a + i] * Nxi[3 * a + k];
    }
  }
  double det = inv3(Xd, J);
  for (int a = 0; a < npe; a++) {
    for (int d = 0; d < 3; d++) {
      Nx[3 * a + d] = Nxi[3 * a] * J[d] + Nxi[3 * a + 1] * J[3 + d] +
                      Nxi[3 * a + 2] * J[6 + d];
    }
  }
  return det;
}

static void grad_field(int npe, const double *Nx, const double *u, double Ux[9]) {
  memset(Ux, 0, 9 * sizeof(double));
  for (int a = 0; a < npe; a++) {
    for (int j = 0; j < 3; j++) {
      for (int d = 0; d < 3; d++) Ux[3 * j + d] += u[3 * a + j] * Nx[3 * a + d];
    }
  }
}

#define MAXN 27

/*
  Element residual and Jacobian of TACSElement3D + TACSLinearElasticity3D
  (linear strain) + TACSSolidConstitutive:
    residual  src/elements/TACSElement3D.cpp:139-172,
              TACSLinearElasticity.cpp:875-951 (evalWeakIntegrand),
              basis/TACSElementBasis.cpp:801-838 (addWeakResidual)
    Jacobian  TACSElement3D.cpp:176-224, TACSLinearElasticity.cpp:1158-
              (evalWeakMatrix: coefficient of du_j/dx_d x du_k/dx_e is
              t*C[voigt(j,d)][voigt(k,e)]; mass pairs rho*t),
              basis/TACSElementBasis.cpp:840-930 (scale by w*detJ*alpha
              for stiffness, w*detJ*gamma for acceleration pairs).
  mat is nv x nv row-major with node-major dofs (row = 3a + j).
  Static analysis has zero velocities/accelerations, so the inertial
  residual term vanishes; gamma adds the consistent mass to the matrix.
*/
void orc_solid_element(int order, const double *X, const double *u,
                       const double C21[21], double t, double rho,
                       double alpha, double gamma, double *res,
                       double *mat) {
  const int npe = order * order * order;
  const int nv = 3 * npe;
  double D[36];
  unpack21(C21, D);
  for (int n = 0; n < npe; n++) {
    double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9], Ux[9];
    double w = orc_hex_quadrature(order, n, pt);
    orc_hex_basis(order, pt, N, Nxi);
    double det = qp_geometry(npe, Nxi, X, Nx, J);
    double wd = w * det;
    if (res) {
      double e[6], s[6];
      grad_field(npe, Nx, u, Ux);
      strain(Ux, e);
      matvec6(D, e, s);
      for (int i = 0; i < 6; i++) s[i] *= t;
      /* DUx[3j+d] = sigma_{jd} */
      for (int a = 0; a < npe; a++) {
        for (int j = 0; j < 3; j++) {
          double r = 0.0;
          for (int d = 0; d < 3; d++) r += s[voigt(j, d)] * Nx[3 * a + d];
          res[3 * a + j] += wd * r;
        }
      }
    }
    if (mat) {
      for (int a = 0; a < npe; a++) {
        for (int b = 0; b < npe; b++) {
          for (int j = 0; j < 3; j++) {
            for (int k = 0; k < 3; k++) {
              double v = 0.0;
              for (int d = 0; d < 3; d++) {
                for (int e = 0; e < 3; e++) {
                  v += D[6 * voigt(j, d) + voigt(k, e)] * Nx[3 * a + d] *
                       Nx[3 * b + e];
                }
              }
              double m = alpha * wd * t * v;
              if (j == k) m += gamma * wd * rho * t * N[a] * N[b];
              mat[(3 * a + j) * nv + 3 * b + k] += m;
            }
          }
        }
      }
    }
  }
}

/*
  Geometric stiffness of one element at the state u, restating
  TACSElement3D::getMatType (src/elements/TACSElement3D.cpp:316-360) for
  TACS_GEOMETRIC_STIFFNESS_MATRIX: TACSLinearElasticity3D::evalWeakMatrix
  (TACSLinearElasticity.cpp:1704-) gives the coefficient sigma_de = (t C e)_de
  of du_j/dx_d x du_j/dx_e (same component j only), scaled by w detJ.
  mat (nv x nv, node-major dofs) is added to.
*/
void orc_solid_geo_element(int order, const double *X, const double *u,
                           const double C21[21], double t, double *mat) {
  const int npe = order * order * order;
  const int nv = 3 * npe;
  double D[36];
  unpack21(C21, D);
  for (int n = 0; n < npe; n++) {
    double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9], Ux[9], e[6], sg[6];
    double w = orc_hex_quadrature(order, n, pt);
    orc_hex_basis(order, pt, N, Nxi);
    double det = qp_geometry(npe, Nxi, X, Nx, J);
    grad_field(npe, Nx, u, Ux);
    strain(Ux, e);
    matvec6(D, e, sg);
    for (int i = 0; i < 6; i++) sg[i] *= t;
    for (int a = 0; a < npe; a++) {
      for (int b = 0; b < npe; b++) {
        double v = 0.0;
        for (int d = 0; d < 3; d++) {
          for (int f = 0; f < 3; f++) v += Nx[3 * a + d] * sg[voigt(d, f)] * Nx[3 * b + f];
        }
        v *= w * det;
        for (int j = 0; j < 3; j++) mat[(3 * a + j) * nv + 3 * b + j] += v;
      }
    }
  }
}

static long find_slot(const int *rowp, const int *cols, int row, int col);

/* A += gc * (geometric stiffness at u) over the mesh, element by element
   (the element loop of TACSAssembler::assembleMatType src/TACSAssembler.cpp:
   4469-4494 for this type). */
long orc_assemble_geo(int order, int nelem, const int *conn, const double *X,
                      const double *u, const double *t_elem, const double C21[21],
                      double gc, const int *rowp, const int *cols, double *A) {
  const int npe = order * order * order;
  const int nv = 3 * npe;
  double Xe[3 * MAXN], ue[3 * MAXN];
  double *ke = (double *)malloc((size_t)nv * nv * sizeof(double));
  long missing = 0;
  for (int e = 0; e < nelem; e++) {
    const int *c = &conn[npe * e];
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) {
        Xe[3 * a + k] = X[3 * c[a] + k];
        ue[3 * a + k] = u[3 * c[a] + k];
      }
    }
    memset(ke, 0, (size_t)nv * nv * sizeof(double));
    orc_solid_geo_element(order, Xe, ue, C21, t_elem[e], ke);
    for (int a = 0; a < npe; a++) {
      for (int b = 0; b < npe; b++) {
        long s = find_slot(rowp, cols, c[a], c[b]);
        if (s < 0) {
          missing++;
          continue;
        }
        for (int r = 0; r < 3; r++) {
          for (int q = 0; q < 3; q++) A[9 * s + 3 * r + q] += gc * ke[(3 * a + r) * nv + 3 * b + q];
        }
      }
    }
  }
  free(ke);
  return missing;
}

/*
  Matrix-free element product y += K_e x_e with K_e = kc * stiffness +
  mc * consistent mass, restating TACSElement3D::getMatVecProductData
  (src/elements/TACSElement3D.cpp:379-425: per quadrature point J = Xd^-1 and
  the weak-form Jacobian entries of TACSLinearElasticity3D::evalWeakMatrix
  scaled by TACSElementBasis::scaleWeakMatrix, basis/TACSElementBasis.cpp:840)
  and TACSElementBasis::addMatVecProduct (basis/TACSElementBasis.cpp:940-
  1050): interpolate U and dU/dxi of x, map the gradient with J^T, apply the
  entries (stress t C e, mass rho t U), map DUx back with J and add the
  transposed interpolation.
*/
void orc_solid_matvec_element(int order, const double *X, const double *x,
                              const double C21[21], double t, double rho,
                              double kc, double mc, double *y) {
  const int npe = order * order * order;
  double D[36];
  unpack21(C21, D);
  for (int n = 0; n < npe; n++) {
    double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9];
    double w = orc_hex_quadrature(order, n, pt);
    orc_hex_basis(order, pt, N, Nxi);
    double det = qp_geometry(npe, Nxi, X, Nx, J);
    double U[3] = {0, 0, 0}, Ud[9] = {0}, Ux[9];
    for (int a = 0; a < npe; a++) {
      for (int c = 0; c < 3; c++) {
        U[c] += N[a] * x[3 * a + c];
        for (int k = 0; k < 3; k++) Ud[3 * c + k] += x[3 * a + c] * Nxi[3 * a + k];
      }
    }
    /* mat3x3MultTrans(J, Ud): Ux[c][d] = sum_k J[3k + d] Ud[c][k] */
    for (int c = 0; c < 3; c++) {
      for (int d = 0; d < 3; d++) {
        Ux[3 * c + d] = J[d] * Ud[3 * c] + J[3 + d] * Ud[3 * c + 1] + J[6 + d] * Ud[3 * c + 2];
      }
    }
    double e[6], sg[6];
    strain(Ux, e);
    matvec6(D, e, sg);
    const double wd = w * det;
    double DUx[9], DU[3], ud[9];
    for (int c = 0; c < 3; c++) {
      DU[c] = mc * wd * rho * t * U[c];
      for (int d = 0; d < 3; d++) DUx[3 * c + d] = kc * wd * t * sg[voigt(c, d)];
    }
    /* mat3x3Mult(J, DUx): ud[c][k] = sum_d J[3k + d] DUx[c][d] */
    for (int c = 0; c < 3; c++) {
      for (int k = 0; k < 3; k++) {
        ud[3 * c + k] = J[3 * k] * DUx[3 * c] + J[3 * k + 1] * DUx[3 * c + 1] +
                        J[3 * k + 2] * DUx[3 * c + 2];
      }
    }
    for (int a = 0; a < npe; a++) {
      for (int c = 0; c < 3; c++) {
        double v = N[a] * DU[c];
        for (int k = 0; k < 3; k++) v += Nxi[3 * a + k] * ud[3 * c + k];
        y[3 * a + c] += v;
      }
    }
  }
}

/* Mesh-level matrix-free product y += sum_e P_e^T K_e P_e x, element by
   element in ascending order (TACSAssembler::addMatrixFreeVecProduct
   src/TACSAssembler.cpp:5645-5717, before its applyBCs). */
void orc_matfree_product(int order, int nelem, const int *conn, const double *X,
                         const double *t_elem, const double C21[21], double rho,
                         double kc, double mc, const double *x, double *y) {
  const int npe = order * order * order;
  double Xe[3 * MAXN], xe[3 * MAXN], ye[3 * MAXN];
  for (int e = 0; e < nelem; e++) {
    const int *c = &conn[npe * e];
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) {
        Xe[3 * a + k] = X[3 * c[a] + k];
        xe[3 * a + k] = x[3 * c[a] + k];
        ye[3 * a + k] = 0.0;
      }
    }
    orc_solid_matvec_element(order, Xe, xe, C21, t_elem[e], rho, kc, mc, ye);
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) y[3 * c[a] + k] += ye[3 * a + k];
    }
  }
}

/*
  Traction on one element face, TACSTraction3D::addResidual
  (src/elements/TACSTraction3D.cpp:87-140) with a constant traction given
  per coordinate direction (tractionCoordinateComponent = 1): res -= int_face
  N * trac dA, with dA from the face normal (basis/TACSElementBasis.cpp:37-64).
*/
void orc_traction_face(int order, int face, const double *X,
                       const double trac[3], double *res) {
  const int npe = order * order * order;
  const int nfq = order * order;
  for (int n = 0; n < nfq; n++) {
    double pt[3], tg[6], N[MAXN], Nxi[3 * MAXN], Xd[9] = {0};
    double w = orc_hex_face_quadrature(order, face, n, pt, tg);
    orc_hex_basis(order, pt, N, Nxi);
    for (int a = 0; a < npe; a++) {
      for (int i = 0; i < 3; i++) {
        for (int k = 0; k < 3; k++) Xd[3 * i + k] += X[3 * a + i] * Nxi[3 * a + k];
      }
    }
    double t1[3], t2[3], nrm[3];
    for (int i = 0; i < 3; i++) {
      t1[i] = Xd[3 * i] * tg[0] + Xd[3 * i + 1] * tg[1] + Xd[3 * i + 2] * tg[2];
      t2[i] = Xd[3 * i] * tg[3] + Xd[3 * i + 1] * tg[4] + Xd[3 * i + 2] * tg[5];
    }
    nrm[0] = t1[1] * t2[2] - t1[2] * t2[1];
    nrm[1] = t1[2] * t2[0] - t1[0] * t2[2];
    nrm[2] = t1[0] * t2[1] - t1[1] * t2[0];
    double A = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    double wa = w * A;
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) res[3 * a + k] -= wa * N[a] * trac[k];
    }
  }
}

/*
  TACSPressure3D::addResidual (src/elements/TACSPressure3D.cpp:60): at each
  face quadrature point the unit normal n = (t1 x t2)/|t1 x t2| with the face
  tangents of TACSHexaBasis (getFaceNormal basis/TACSElementBasis.cpp:37),
  DUt = -p n, res += w |t1 x t2| N DUt.
*/
void orc_pressure_face(int order, int face, const double *X, double p, double *res) {
  const int npe = order * order * order;
  const int nfq = order * order;
  for (int n = 0; n < nfq; n++) {
    double pt[3], tg[6], N[MAXN], Nxi[3 * MAXN], Xd[9] = {0};
    double w = orc_hex_face_quadrature(order, face, n, pt, tg);
    orc_hex_basis(order, pt, N, Nxi);
    for (int a = 0; a < npe; a++) {
      for (int i = 0; i < 3; i++) {
        for (int k = 0; k < 3; k++) Xd[3 * i + k] += X[3 * a + i] * Nxi[3 * a + k];
      }
    }
    double t1[3], t2[3], nrm[3];
    for (int i = 0; i < 3; i++) {
      t1[i] = Xd[3 * i] * tg[0] + Xd[3 * i + 1] * tg[1] + Xd[3 * i + 2] * tg[2];
      t2[i] = Xd[3 * i] * tg[3] + Xd[3 * i + 1] * tg[4] + Xd[3 * i + 2] * tg[5];
    }
    nrm[0] = t1[1] * t2[2] - t1[2] * t2[1];
    nrm[1] = t1[2] * t2[0] - t1[0] * t2[2];
    nrm[2] = t1[0] * t2[1] - t1[1] * t2[0];
    double A = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    double DUt[3];
    for (int k = 0; k < 3; k++) DUt[k] = -p * nrm[k] / A;
    double wa = w * A;
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) res[3 * a + k] += wa * N[a] * DUt[k];
    }
  }
}

/* Pressure aux elements (TACSPressure3D), summed into the node residual. */
void orc_add_pressures(int order, int naux, const int *aux_elem, const int *aux_face,
                       const double *aux_p, const int *conn, const double *X, double *res) {
  const int npe = order * order * order;
  double Xe[3 * MAXN], re[3 * MAXN];
  for (int q = 0; q < naux; q++) {
    const int *c = &conn[npe * aux_elem[q]];
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) Xe[3 * a + k] = X[3 * c[a] + k];
    }
    memset(re, 0, sizeof(re));
    orc_pressure_face(order, aux_face[q], Xe, aux_p[q], re);
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) res[3 * c[a] + k] += re[3 * a + k];
    }
  }
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/*
  Node-to-node CSR: row i = sorted, unique union of the nodes of every
  element containing i, diagonal kept (TACSAssembler::
  computeLocalNodeToNodeCSR src/TACSAssembler.cpp:1899, via
  TacsSortAndUniquifyCSR src/TacsUtilities.h:102).  Returns nnz, or -1 if
  cap is too small.  With cols == NULL only counts.
*/
long orc_node_csr(int nnodes, int nelem, int npe, const int *conn, int *rowp,
                  int *cols, long cap) {
  int *cnt = (int *)calloc(nnodes + 1, sizeof(int));
  for (int e = 0; e < nelem; e++) {
    for (int a = 0; a < npe; a++) cnt[conn[npe * e + a] + 1]++;
  }
  for (int i = 0; i < nnodes; i++) cnt[i + 1] += cnt[i];
  int *n2e = (int *)malloc((size_t)cnt[nnodes] * sizeof(int));
  int *pos = (int *)malloc(nnodes * sizeof(int));
  memcpy(pos, cnt, nnodes * sizeof(int));
  for (int e = 0; e < nelem; e++) {
    for (int a = 0; a < npe; a++) n2e[pos[conn[npe * e + a]]++] = e;
  }
  int *tmp = (int *)malloc((size_t)64 * npe * sizeof(int) + 64);
  long nnz = 0;
  rowp[0] = 0;
  for (int i = 0; i < nnodes; i++) {
    int m = 0;
    int ne = cnt[i + 1] - cnt[i];
    if ((size_t)ne * npe > (size_t)64 * npe) {
      free(tmp);
      tmp = (int *)malloc((size_t)ne * npe * sizeof(int));
    }
    for (int p = cnt[i]; p < cnt[i + 1]; p++) {
      int e = n2e[p];
      for (int a = 0; a < npe; a++) tmp[m++] = conn[npe * e + a];
    }
    qsort(tmp, m, sizeof(int), cmp_int);
    int u = 0;
    for (int q = 0; q < m; q++) {
      if (q == 0 || tmp[q] != tmp[q - 1]) tmp[u++] = tmp[q];
    }
    if (cols) {
      if (nnz + u > cap) {
        nnz = -1;
        break;
      }
      memcpy(&cols[nnz], tmp, u * sizeof(int));
    }
    nnz += u;
    rowp[i + 1] = (int)nnz;
  }
  free(tmp);
  free(pos);
  free(n2e);
  free(cnt);
  return nnz;
}

/* Column search within a sorted row (TacsSearchArray src/TacsUtilities.h:50). */
static long find_slot(const int *rowp, const int *cols, int row, int col) {
  int lo = rowp[row], hi = rowp[row + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (cols[mid] == col) return mid;
    if (cols[mid] < col) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

/*
  Whole-mesh assembly, the element loop of TACSAssembler::assembleJacobian
  (src/TACSAssembler.cpp:4352-4383) / assembleRes (:4190-4227) for one
  solid element type, scattering into b=3 BCSR with the add rule of
  BCSRMat::addRowValues (src/bpmat/BCSRMat.cpp:1800-1840): block (r,c) of
  slot(row=conn[a], col=conn[b]) += mat[(3a+r)*nv + 3b + c].  A and res are
  accumulated (callers zero them, as the reference does at :4300).
  Returns the number of missing pattern entries (0 on success).
*/
long orc_assemble_solid(int order, int nelem, const int *conn, const double *X,
                        const double *u, const double *t_elem,
                        const double C21[21], double rho, double alpha,
                        double gamma, const int *rowp, const int *cols,
                        double *A, double *res) {
  const int npe = order * order * order;
  const int nv = 3 * npe;
  double *Xe = (double *)malloc(nv * sizeof(double));
  double *ue = (double *)malloc(nv * sizeof(double));
  double *re = (double *)malloc(nv * sizeof(double));
  double *ke = (double *)malloc((size_t)nv * nv * sizeof(double));
  long missing = 0;
  for (int e = 0; e < nelem; e++) {
    const int *c = &conn[npe * e];
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) {
        Xe[3 * a + k] = X[3 * c[a] + k];
        ue[3 * a + k] = u ? u[3 * c[a] + k] : 0.0;
      }
    }
    memset(re, 0, nv * sizeof(double));
    if (A) memset(ke, 0, (size_t)nv * nv * sizeof(double));
    orc_solid_element(order, Xe, ue, C21, t_elem[e], rho, alpha, gamma,
                      res ? re : NULL, A ? ke : NULL);
    if (res) {
      for (int a = 0; a < npe; a++) {
        for (int k = 0; k < 3; k++) res[3 * c[a] + k] += re[3 * a + k];
      }
    }
    if (A) {
      for (int a = 0; a < npe; a++) {
        for (int b = 0; b < npe; b++) {
          long s = find_slot(rowp, cols, c[a], c[b]);
          if (s < 0) {
            missing++;
            continue;
          }
          double *blk = &A[9 * s];
          for (int r = 0; r < 3; r++) {
            for (int q = 0; q < 3; q++) blk[3 * r + q] += ke[(3 * a + r) * nv + 3 * b + q];
          }
        }
      }
    }
  }
  free(Xe);
  free(ue);
  free(re);
  free(ke);
  return missing;
}

/* Traction aux elements: res[conn] += traction residual of (elem, face). */
void orc_add_tractions(int order, int naux, const int *aux_elem,
                       const int *aux_face, const double *aux_trac,
                       const int *conn, const double *X, double *res) {
  const int npe = order * order * order;
  double Xe[3 * MAXN], re[3 * MAXN];
  for (int q = 0; q < naux; q++) {
    const int *c = &conn[npe * aux_elem[q]];
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) Xe[3 * a + k] = X[3 * c[a] + k];
    }
    memset(re, 0, sizeof(re));
    orc_traction_face(order, aux_face[q], Xe, &aux_trac[3 * q], re);
    for (int a = 0; a < npe; a++) {
      for (int k = 0; k < 3; k++) res[3 * c[a] + k] += re[3 * a + k];
    }
  }
}

/* BCSRMat::zeroRow with ident=1 (src/bpmat/BCSRMat.cpp:2027-2053), applied
   per BC node as TACSParallelMat::applyBCs does (TACSParallelMat.cpp:343).
   Columns are not touched. */
void orc_apply_bc_mat(int b, int nbc, const int *nodes, const int *vars,
                      const int *rowp, const int *cols, double *A) {
  for (int q = 0; q < nbc; q++) {
    int row = nodes[q];
    for (int j = rowp[row]; j < rowp[row + 1]; j++) {
      double *a = &A[(size_t)b * b * j];
      for (int ii = 0; ii < b; ii++) {
        if (vars[q] & (1 << ii)) {
          for (int jj = 0; jj < b; jj++) a[b * ii + jj] = 0.0;
        }
      }
      if (cols[j] == row) {
        for (int ii = 0; ii < b; ii++) {
          if (vars[q] & (1 << ii)) a[(b + 1) * ii] = 1.0;
        }
      }
    }
  }
}

/* TACSBVec::applyBCs (src/bpmat/TACSBVec.cpp:546-598): constrained entries
   become u - lambda*value when u is given, else 0. */
void orc_apply_bc_vec(int b, int nbc, const int *nodes, const int *vars,
                      const double *vals, const double *u, double lambda,
                      double *x) {
  for (int q = 0; q < nbc; q++) {
    for (int k = 0; k < b; k++) {
      if (vars[q] & (1 << k)) {
        size_t i = (size_t)b * nodes[q] + k;
        x[i] = u ? u[i] - lambda * (vals ? vals[b * q + k] : 0.0) : 0.0;
      }
    }
  }
}

/* y = A x for b x b BCSR, row-major blocks; per block row the sum runs over
   the row's blocks in storage order (BCSRMatVecMult3 src/bpmat/
   BCSRMatMult3.cpp:28, BCSRMatVecMult6 BCSRMatMult6.cpp:82). */
void orc_bsr_mult(int b, int nrows, const int *rowp, const int *cols,
                  const double *A, const double *x, double *y) {
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

/* Dense in-place inverse of a b x b block by Gauss-Jordan with partial
   pivoting; returns 0 on success. */
static int block_inverse(int b, const double *a, double *inv) {
  double m[64], r[64];
  memcpy(m, a, b * b * sizeof(double));
  for (int i = 0; i < b * b; i++) r[i] = 0.0;
  for (int i = 0; i < b; i++) r[(b + 1) * i] = 1.0;
  for (int c = 0; c < b; c++) {
    int p = c;
    for (int i = c + 1; i < b; i++) {
      if (fabs(m[b * i + c]) > fabs(m[b * p + c])) p = i;
    }
    if (m[b * p + c] == 0.0) return 1;
    if (p != c) {
      for (int k = 0; k < b; k++) {
        double t = m[b * c + k]; m[b * c + k] = m[b * p + k]; m[b * p + k] = t;
        t = r[b * c + k]; r[b * c + k] = r[b * p + k]; r[b * p + k] = t;
      }
    }
    double d = 1.0 / m[b * c + c];
    for (int k = 0; k < b; k++) {
      m[b * c + k] *= d;
      r[b * c + k] *= d;
    }
    for (int i = 0; i < b; i++) {
      if (i != c) {
        double f = m[b * i + c];
        for (int k = 0; k < b; k++) {
          m[b * i + k] -= f * m[b * c + k];
          r[b * i + k] -= f * r[b * c + k];
        }
      }
    }
  }
  memcpy(inv, r, b * b * sizeof(double));
  return 0;
}

/* Inverse diagonal blocks: the block-Jacobi preconditioner built from the
   same quantity as BCSRMat::factorDiag (src/bpmat/BCSRMat.cpp:1527).
   Returns the number of singular blocks. */
int orc_bsr_diag_inv(int b, int nrows, const int *rowp, const int *cols,
                     const double *A, double *Dinv) {
  int bad = 0;
  for (int i = 0; i < nrows; i++) {
    long s = find_slot(rowp, cols, i, i);
    if (s < 0 || block_inverse(b, &A[(size_t)b * b * s], &Dinv[(size_t)b * b * i])) bad++;
  }
  return bad;
}

static void apply_pc(int b, int nrows, const double *Dinv, const double *x,
                     double *y) {
  if (!Dinv) {
    memcpy(y, x, (size_t)b * nrows * sizeof(double));
    return;
  }
  for (int i = 0; i < nrows; i++) {
    const double *d = &Dinv[(size_t)b * b * i];
    for (int r = 0; r < b; r++) {
      double v = 0.0;
      for (int c = 0; c < b; c++) v += d[b * r + c] * x[(size_t)b * i + c];
      y[(size_t)b * i + r] = v;
    }
  }
}

int orc_gmres_capped(int b, int nrows, const int *rowp, const int *cols,
                     const double *A, const double *Dinv, const double *rhs,
                     double *x, int m, int nrestart, double rtol, double atol,
                     int zero_guess, int max_its, int *iters,
                     double *resnorm_out);

static double vdot(size_t n, const double *x, const double *y) {
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}
static void vaxpy(size_t n, double a, const double *x, double *y) {
  for (size_t i = 0; i < n; i++) y[i] += a * x[i];
}
static void vscale(size_t n, double a, double *x) {
  for (size_t i = 0; i < n; i++) x[i] *= a;
}

/*
  Restarted right-preconditioned GMRES with modified Gram-Schmidt,
  restating GMRES::solve (src/bpmat/KSM.cpp:785-958) with
  ModifiedGramSchmidt (:526): Givens rotations on the column-packed
  Hessenberg, stop when |res| < atol or < rtol*|r0| where r0 is the very
  first residual, solution update x += M^{-1} (W y).  zero_guess as in the
  reference.  Returns solve_flag (1 converged); *iters = total Arnoldi steps.
*/
int orc_gmres(int b, int nrows, const int *rowp, const int *cols,
              const double *A, const double *Dinv, const double *rhs,
              double *x, int m, int nrestart, double rtol, double atol,
              int zero_guess, int *iters, double *resnorm_out) {
  return orc_gmres_capped(b, nrows, rowp, cols, A, Dinv, rhs, x, m, nrestart,
                          rtol, atol, zero_guess, 0, iters, resnorm_out);
}

/* Same, stopping after max_its Arnoldi steps when max_its > 0 (used only to
   time a bounded sample of iterations). */
int orc_gmres_capped(int b, int nrows, const int *rowp, const int *cols,
                     const double *A, const double *Dinv, const double *rhs,
                     double *x, int m, int nrestart, double rtol, double atol,
                     int zero_guess, int max_its, int *iters,
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
      res[0] = sqrt(vdot(n, W[0], W[0]));
      vscale(n, 1.0 / res[0], W[0]);
    } else {
      orc_bsr_mult(b, nrows, rowp, cols, A, x, W[0]);
      vaxpy(n, -1.0, rhs, W[0]);
      res[0] = sqrt(vdot(n, W[0], W[0]));
      vscale(n, -1.0 / res[0], W[0]);
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
      double *h = &H[(size_t)(m + 1) * i]; /* column i, length i+2 */
      apply_pc(b, nrows, Dinv, W[i], work);
      orc_bsr_mult(b, nrows, rowp, cols, A, work, W[i + 1]);
      for (int j = 0; j <= i; j++) {
        h[j] = vdot(n, W[j], W[i + 1]);
        vaxpy(n, -h[j], W[j], W[i + 1]);
      }
      h[i + 1] = sqrt(vdot(n, W[i + 1], W[i + 1]));
      vscale(n, 1.0 / h[i + 1], W[i + 1]);
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
    for (int i = 0; i < niters; i++) vaxpy(n, res[i], W[i], work);
    apply_pc(b, nrows, Dinv, work, W[0]);
    vaxpy(n, 1.0, W[0], x);
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

/*
  Function integrands and their derivatives for the solid element.
  mode 0: compliance = sum w*detXd*(e.s), s = t*C*e  (TACSCompliance::
          elementWiseEval src/functions/TACSCompliance.cpp:69-93 with
          TACS_STRAIN_ENERGY_DENSITY, TACSLinearElasticity.cpp:1980-2008;
          no factor 1/2).
  Failure index fail = vonMises(C*e)/ys (TACSSolidConstitutive::evalFailure
  src/constitutive/TACSSolidConstitutive.cpp:274, TACSMaterialProperties::
  vonMisesFailure3D :439); t does not enter.
*/
static void elem_gather(int npe, const int *c, const double *X, const double *u,
                        double *Xe, double *ue) {
  for (int a = 0; a < npe; a++) {
    for (int k = 0; k < 3; k++) {
      Xe[3 * a + k] = X[3 * c[a] + k];
      if (ue) ue[3 * a + k] = u[3 * c[a] + k];
    }
  }
}

static double von_mises(const double s[6]) {
  return sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[0] - s[2]) * (s[0] - s[2]) +
                     (s[1] - s[2]) * (s[1] - s[2]) +
                     6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
}

double orc_compliance(int order, int nelem, const int *conn, const double *X,
                      const double *u, const double *t_elem,
                      const double C21[21]) {
  const int npe = order * order * order;
  double D[36], Xe[3 * MAXN], ue[3 * MAXN], total = 0.0;
  unpack21(C21, D);
  for (int el = 0; el < nelem; el++) {
    elem_gather(npe, &conn[npe * el], X, u, Xe, ue);
    for (int n = 0; n < npe; n++) {
      double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9], Ux[9], e[6], s[6];
      double w = orc_hex_quadrature(order, n, pt);
      orc_hex_basis(order, pt, N, Nxi);
      double det = qp_geometry(npe, Nxi, Xe, Nx, J);
      grad_field(npe, Nx, ue, Ux);
      strain(Ux, e);
      matvec6(D, e, s);
      double U0 = 0.0;
      for (int i = 0; i < 6; i++) U0 += e[i] * t_elem[el] * s[i];
      total += w * det * U0;
    }
  }
  return total;
}

/* Derivative of the compliance w.r.t. u (TACSCompliance::getElementSVSens
   :95-125 -> TACSLinearElasticity3D::evalPointQuantitySens :2123 gives
   dfdUx = 2*dfdq*sigma; addFieldGradientSVSens basis/TACSElementBasis.cpp:
   407).  Adds into dfdu (no BCs, like TACSAssembler::addSVSens :4941). */
void orc_compliance_svsens(int order, int nelem, const int *conn,
                           const double *X, const double *u,
                           const double *t_elem, const double C21[21],
                           double alpha, double *dfdu) {
  const int npe = order * order * order;
  double D[36], Xe[3 * MAXN], ue[3 * MAXN];
  unpack21(C21, D);
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[npe * el];
    elem_gather(npe, c, X, u, Xe, ue);
    for (int n = 0; n < npe; n++) {
      double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9], Ux[9], e[6], s[6];
      double w = orc_hex_quadrature(order, n, pt);
      orc_hex_basis(order, pt, N, Nxi);
      double det = qp_geometry(npe, Nxi, Xe, Nx, J);
      grad_field(npe, Nx, ue, Ux);
      strain(Ux, e);
      matvec6(D, e, s);
      double f = 2.0 * w * det * alpha * t_elem[el];
      for (int a = 0; a < npe; a++) {
        for (int j = 0; j < 3; j++) {
          double r = 0.0;
          for (int d = 0; d < 3; d++) r += s[voigt(j, d)] * Nx[3 * a + d];
          dfdu[3 * c[a] + j] += f * r;
        }
      }
    }
  }
}

/* Per-element derivative of the compliance w.r.t. the element thickness
   DV: scale * sum w*detXd*(e . C e)   (TACSCompliance::addElementDVSens
   :162 -> TACSLinearElasticity3D::addPointQuantityDVSens :2049 ->
   TACSSolidConstitutive::addStressDVSens :181). */
void orc_compliance_dvsens(int order, int nelem, const int *conn,
                           const double *X, const double *u,
                           const double C21[21], double scale,
                           double *dfdt_elem) {
  const int npe = order * order * order;
  double D[36], Xe[3 * MAXN], ue[3 * MAXN];
  unpack21(C21, D);
  for (int el = 0; el < nelem; el++) {
    elem_gather(npe, &conn[npe * el], X, u, Xe, ue);
    double acc = 0.0;
    for (int n = 0; n < npe; n++) {
      double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9], Ux[9], e[6], s[6];
      double w = orc_hex_quadrature(order, n, pt);
      orc_hex_basis(order, pt, N, Nxi);
      double det = qp_geometry(npe, Nxi, Xe, Nx, J);
      grad_field(npe, Nx, ue, Ux);
      strain(Ux, e);
      matvec6(D, e, s);
      double v = 0.0;
      for (int i = 0; i < 6; i++) v += s[i] * e[i];
      acc += scale * w * det * v;
    }
    dfdt_elem[el] += acc;
  }
}

/* Per-element adjoint-residual product psi^T dR/dt (TACSElement3D::
   addAdjResProduct src/elements/TACSElement3D.cpp:227-257 ->
   TACSLinearElasticity3D::addWeakAdjProduct TACSLinearElasticity.cpp:953 ->
   addStressDVSens(scale*w*detXd, e(u), e(psi))).  Static: no inertia. */
void orc_adj_res_product(int order, int nelem, const int *conn,
                         const double *X, const double *u, const double *psi,
                         const double C21[21], double scale,
                         double *dfdt_elem) {
  const int npe = order * order * order;
  double D[36], Xe[3 * MAXN], ue[3 * MAXN], pe[3 * MAXN];
  unpack21(C21, D);
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[npe * el];
    elem_gather(npe, c, X, u, Xe, ue);
    elem_gather(npe, c, X, psi, Xe, pe);
    double acc = 0.0;
    for (int n = 0; n < npe; n++) {
      double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9], Ux[9], Px[9];
      double e[6], phi[6], s[6];
      double w = orc_hex_quadrature(order, n, pt);
      orc_hex_basis(order, pt, N, Nxi);
      double det = qp_geometry(npe, Nxi, Xe, Nx, J);
      grad_field(npe, Nx, ue, Ux);
      grad_field(npe, Nx, pe, Px);
      strain(Ux, e);
      strain(Px, phi);
      matvec6(D, e, s);
      double v = 0.0;
      for (int i = 0; i < 6; i++) v += s[i] * phi[i];
      acc += scale * det * w * v;
    }
    dfdt_elem[el] += acc;
  }
}

/*
  KS failure, two-stage KS_CONTINUOUS (TACSKSFailure.cpp:74-78 value,
  :125-189 elementWiseEval): maxFail = max over qp of sf*fail;
  ksSum = sum w*detXd*exp(ksWeight*(sf*fail - maxFail));
  KS = maxFail + log(ksSum/alpha)/ksWeight.  out[0]=KS, [1]=maxFail,
  [2]=ksSum.
*/
void orc_ks_failure(int order, int nelem, const int *conn, const double *X,
                    const double *u, const double C21[21], double ys,
                    double ksWeight, double alpha, double safety,
                    double out[3]) {
  const int npe = order * order * order;
  double D[36], Xe[3 * MAXN], ue[3 * MAXN];
  unpack21(C21, D);
  double maxf = -1e20, sum = 0.0;
  for (int pass = 0; pass < 2; pass++) {
    for (int el = 0; el < nelem; el++) {
      elem_gather(npe, &conn[npe * el], X, u, Xe, ue);
      for (int n = 0; n < npe; n++) {
        double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9], Ux[9], e[6], s[6];
        double w = orc_hex_quadrature(order, n, pt);
        orc_hex_basis(order, pt, N, Nxi);
        double det = qp_geometry(npe, Nxi, Xe, Nx, J);
        grad_field(npe, Nx, ue, Ux);
        strain(Ux, e);
        matvec6(D, e, s);
        double f = safety * von_mises(s) / ys;
        if (pass == 0) {
          if (f > maxf) maxf = f;
        } else {
          sum += w * det * exp(ksWeight * (f - maxf));
        }
      }
    }
  }
  out[0] = maxf + log(sum / alpha) / ksWeight;
  out[1] = maxf;
  out[2] = sum;
}

/* d KS / du (TACSKSFailure::getElementSVSens :221-283 ->
   TACSLinearElasticity3D::evalPointQuantitySens :2123 (failure branch) ->
   TACSSolidConstitutive::evalFailureStrainSens :301 ->
   vonMisesFailure3DStressSens TACSMaterialProperties.cpp:449). */
void orc_ks_svsens(int order, int nelem, const int *conn, const double *X,
                   const double *u, const double C21[21], double ys,
                   double ksWeight, double safety, double maxFail,
                   double ksSum, double alpha, double *dfdu) {
  const int npe = order * order * order;
  double D[36], Xe[3 * MAXN], ue[3 * MAXN];
  unpack21(C21, D);
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[npe * el];
    elem_gather(npe, c, X, u, Xe, ue);
    for (int n = 0; n < npe; n++) {
      double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9], Ux[9], e[6], s[6];
      double w = orc_hex_quadrature(order, n, pt);
      orc_hex_basis(order, pt, N, Nxi);
      double det = qp_geometry(npe, Nxi, Xe, Nx, J);
      grad_field(npe, Nx, ue, Ux);
      strain(Ux, e);
      matvec6(D, e, s);
      double vm = von_mises(s);
      double f = safety * vm / ys;
      double dfdq = exp(ksWeight * (f - maxFail)) / ksSum * w * det * safety;
      double ds[6] = {0}, de[6];
      if (vm != 0.0) {
        double fact = 0.5 / (ys * vm);
        ds[0] = fact * (2.0 * s[0] - s[1] - s[2]);
        ds[1] = fact * (2.0 * s[1] - s[0] - s[2]);
        ds[2] = fact * (2.0 * s[2] - s[0] - s[1]);
        ds[3] = 6.0 * fact * s[3];
        ds[4] = 6.0 * fact * s[4];
        ds[5] = 6.0 * fact * s[5];
      }
      matvec6(D, ds, de);
      for (int a = 0; a < npe; a++) {
        for (int j = 0; j < 3; j++) {
          double r = 0.0;
          for (int d = 0; d < 3; d++) r += de[voigt(j, d)] * Nx[3 * a + d];
          dfdu[3 * c[a] + j] += alpha * dfdq * r;
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* Node-coordinate sensitivities                                       */

/* 3x3 stress tensor of Voigt s (xx, yy, zz, yz, xz, xy) */
static void sig_tensor(const double s[6], double S[9]) {
  S[0] = s[0]; S[4] = s[1]; S[8] = s[2];
  S[5] = S[7] = s[3];
  S[2] = S[6] = s[4];
  S[1] = S[3] = s[5];
}

/* dfdX_a += f * (W I - sum_p Ux_p^T P_p) grad_x N_a for one quadrature point
   (the derivative of w detXd W(Ux) with Ux = Ud Xd^{-1} w.r.t. the nodes) */
static void add_xpt_term(int npe, const double *Nx, double f, double W,
                         int np, const double *Ux[], const double *P[],
                         const int *c, double *dfdX) {
  double M[9];
  for (int i = 0; i < 3; i++) {
    for (int l = 0; l < 3; l++) {
      double v = (i == l) ? W : 0.0;
      for (int p = 0; p < np; p++) {
        for (int k = 0; k < 3; k++) v -= Ux[p][3 * k + i] * P[p][3 * k + l];
      }
      M[3 * i + l] = v;
    }
  }
  for (int a = 0; a < npe; a++) {
    for (int i = 0; i < 3; i++) {
      double v = 0.0;
      for (int l = 0; l < 3; l++) v += M[3 * i + l] * Nx[3 * a + l];
      dfdX[3 * c[a] + i] += f * v;
    }
  }
}

/*
  Partial derivative of the compliance w.r.t. the node coordinates at fixed
  u: TACSAssembler::addXptSens (src/TACSAssembler.cpp:4858) ->
  TACSCompliance::getElementXptSens (src/functions/TACSCompliance.cpp:129:
  dfdq = w detXd, dfddetXd = w U0) -> TACSElement3D::addPointQuantityXptSens
  (src/elements/TACSElement3D.cpp:618).  Per quadrature point the derivative
  of w detXd W, W = e.(tCe), is w detXd (W I - Ux^T 2 sigma) grad_x N_a.
*/
void orc_compliance_xptsens(int order, int nelem, const int *conn,
                            const double *X, const double *u,
                            const double *t_elem, const double C21[21],
                            double scale, double *dfdX) {
  const int npe = order * order * order;
  double D[36], Xe[3 * MAXN], ue[3 * MAXN];
  unpack21(C21, D);
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[npe * el];
    elem_gather(npe, c, X, u, Xe, ue);
    for (int n = 0; n < npe; n++) {
      double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9], Ux[9], e[6], s[6], S[9];
      double w = orc_hex_quadrature(order, n, pt);
      orc_hex_basis(order, pt, N, Nxi);
      double det = qp_geometry(npe, Nxi, Xe, Nx, J);
      grad_field(npe, Nx, ue, Ux);
      strain(Ux, e);
      matvec6(D, e, s);
      for (int i = 0; i < 6; i++) s[i] *= t_elem[el];
      double W = 0.0;
      for (int i = 0; i < 6; i++) W += e[i] * s[i];
      sig_tensor(s, S);
      for (int i = 0; i < 9; i++) S[i] *= 2.0;
      const double *Ups[1] = {Ux};
      const double *Ps[1] = {S};
      add_xpt_term(npe, Nx, scale * w * det, W, 1, Ups, Ps, c, dfdX);
    }
  }
}

/*
  scale * psi^T dR/dX of the element residual: TACSAssembler::
  addAdjointResXptSensProducts (src/TACSAssembler.cpp:5136) ->
  TACSElement3D::addAdjResXptProduct (src/elements/TACSElement3D.cpp:260).
  psi^T R_e = sum_q w detXd e(psi).(t C e(u)); its node derivative per
  quadrature point is w detXd (g I - Ux_u^T sigma_psi - Ux_psi^T sigma_u)
  grad_x N_a with g = e(psi).sigma_u.
*/
void orc_adj_res_xpt_product(int order, int nelem, const int *conn,
                             const double *X, const double *u,
                             const double *psi, const double *t_elem,
                             const double C21[21], double scale, double *dfdX) {
  const int npe = order * order * order;
  double D[36], Xe[3 * MAXN], ue[3 * MAXN], pe[3 * MAXN];
  unpack21(C21, D);
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[npe * el];
    elem_gather(npe, c, X, u, Xe, ue);
    elem_gather(npe, c, X, psi, Xe, pe);
    for (int n = 0; n < npe; n++) {
      double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9];
      double Ux[9], Px[9], eu[6], ep[6], su[6], sp[6], Su[9], Sp[9];
      double w = orc_hex_quadrature(order, n, pt);
      orc_hex_basis(order, pt, N, Nxi);
      double det = qp_geometry(npe, Nxi, Xe, Nx, J);
      grad_field(npe, Nx, ue, Ux);
      grad_field(npe, Nx, pe, Px);
      strain(Ux, eu);
      strain(Px, ep);
      matvec6(D, eu, su);
      matvec6(D, ep, sp);
      for (int i = 0; i < 6; i++) {
        su[i] *= t_elem[el];
        sp[i] *= t_elem[el];
      }
      double g = 0.0;
      for (int i = 0; i < 6; i++) g += ep[i] * su[i];
      sig_tensor(su, Su);
      sig_tensor(sp, Sp);
      const double *Ups[2] = {Ux, Px};
      const double *Ps[2] = {Sp, Su};
      add_xpt_term(npe, Nx, scale * w * det, g, 2, Ups, Ps, c, dfdX);
    }
  }
}

/*
  scale * psi^T dR/dX of coordinate-component traction aux elements.  The
  residual is -sum_q w A N_a t with A = |t1 x t2| (TACSTraction3D::addResidual
  src/elements/TACSTraction3D.cpp:87, TACSElementBasis::getFaceNormal
  basis/TACSElementBasis.cpp:37).  The reference has no analytic version (it
  differences the residual forward with step 1e-7, TACSElement::
  addAdjResXptProduct src/elements/TACSElement.cpp:270); this is the exact
  derivative: dA/dX_a = (t2 x n)(grad N_a . tg1) + (n x t1)(grad N_a . tg2).
*/
void orc_traction_xpt_product(int order, int naux, const int *aux_elem,
                              const int *aux_face, const double *aux_trac,
                              const int *conn, const double *X,
                              const double *psi, double scale, double *dfdX) {
  const int npe = order * order * order;
  const int nfq = order * order;
  for (int q = 0; q < naux; q++) {
    const int *c = &conn[npe * aux_elem[q]];
    const double *tr = &aux_trac[3 * q];
    double Xe[3 * MAXN], pe[3 * MAXN];
    elem_gather(npe, c, X, psi, Xe, pe);
    for (int n = 0; n < nfq; n++) {
      double pt[3], tg[6], N[MAXN], Nxi[3 * MAXN], Xd[9] = {0};
      double w = orc_hex_face_quadrature(order, aux_face[q], n, pt, tg);
      orc_hex_basis(order, pt, N, Nxi);
      for (int a = 0; a < npe; a++) {
        for (int i = 0; i < 3; i++) {
          for (int k = 0; k < 3; k++) Xd[3 * i + k] += Xe[3 * a + i] * Nxi[3 * a + k];
        }
      }
      double t1[3], t2[3], nv[3];
      for (int i = 0; i < 3; i++) {
        t1[i] = Xd[3 * i] * tg[0] + Xd[3 * i + 1] * tg[1] + Xd[3 * i + 2] * tg[2];
        t2[i] = Xd[3 * i] * tg[3] + Xd[3 * i + 1] * tg[4] + Xd[3 * i + 2] * tg[5];
      }
      nv[0] = t1[1] * t2[2] - t1[2] * t2[1];
      nv[1] = t1[2] * t2[0] - t1[0] * t2[2];
      nv[2] = t1[0] * t2[1] - t1[1] * t2[0];
      double A = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
      double nh[3] = {nv[0] / A, nv[1] / A, nv[2] / A};
      double a1[3] = {t2[1] * nh[2] - t2[2] * nh[1], t2[2] * nh[0] - t2[0] * nh[2],
                      t2[0] * nh[1] - t2[1] * nh[0]};
      double a2[3] = {nh[1] * t1[2] - nh[2] * t1[1], nh[2] * t1[0] - nh[0] * t1[2],
                      nh[0] * t1[1] - nh[1] * t1[0]};
      /* psi^T R = -w A P, P = sum_a N_a psi_a . t */
      double P = 0.0;
      for (int a = 0; a < npe; a++) {
        P += N[a] * (pe[3 * a] * tr[0] + pe[3 * a + 1] * tr[1] + pe[3 * a + 2] * tr[2]);
      }
      for (int a = 0; a < npe; a++) {
        double g1 = Nxi[3 * a] * tg[0] + Nxi[3 * a + 1] * tg[1] + Nxi[3 * a + 2] * tg[2];
        double g2 = Nxi[3 * a] * tg[3] + Nxi[3 * a + 1] * tg[4] + Nxi[3 * a + 2] * tg[5];
        for (int i = 0; i < 3; i++) {
          dfdX[3 * c[a] + i] += scale * (-w * P) * (a1[i] * g1 + a2[i] * g2);
        }
      }
    }
  }
}

/*
  Partial derivative of the continuous KS failure w.r.t. the node coordinates
  at fixed u (TACSKSFailure::getElementXptSens src/functions/TACSKSFailure.cpp
  -> TACSElement3D::addPointQuantityXptSens :618): with
  Phi_q = exp(rho (f_q - max)) / (rho S), KS = max + log(sum_q w detXd
  exp(rho (f_q - max))) / rho has node derivative sum_q w detXd (Phi I -
  Ux^T rho Phi df/dUx) grad_x N_a; f = safety * von Mises / ys is independent
  of the thickness for solids.
*/
void orc_ks_xptsens(int order, int nelem, const int *conn, const double *X,
                    const double *u, const double C21[21], double ys,
                    double ksWeight, double safety, double maxFail,
                    double ksSum, double scale, double *dfdX) {
  const int npe = order * order * order;
  double D[36], Xe[3 * MAXN], ue[3 * MAXN];
  unpack21(C21, D);
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[npe * el];
    elem_gather(npe, c, X, u, Xe, ue);
    for (int n = 0; n < npe; n++) {
      double pt[3], N[MAXN], Nxi[3 * MAXN], Nx[3 * MAXN], J[9], Ux[9], e[6], s[6];
      double w = orc_hex_quadrature(order, n, pt);
      orc_hex_basis(order, pt, N, Nxi);
      double det = qp_geometry(npe, Nxi, Xe, Nx, J);
      grad_field(npe, Nx, ue, Ux);
      strain(Ux, e);
      matvec6(D, e, s);
      double vm = von_mises(s);
      double f = safety * vm / ys;
      double Phi = exp(ksWeight * (f - maxFail)) / (ksWeight * ksSum);
      double ds[6] = {0}, de[6], P[9];
      if (vm != 0.0) {
        double fact = 0.5 / (ys * vm);
        ds[0] = fact * (2.0 * s[0] - s[1] - s[2]);
        ds[1] = fact * (2.0 * s[1] - s[0] - s[2]);
        ds[2] = fact * (2.0 * s[2] - s[0] - s[1]);
        ds[3] = 6.0 * fact * s[3];
        ds[4] = 6.0 * fact * s[4];
        ds[5] = 6.0 * fact * s[5];
      }
      matvec6(D, ds, de);
      sig_tensor(de, P);
      for (int i = 0; i < 9; i++) P[i] *= ksWeight * Phi * safety;
      const double *Ups[1] = {Ux};
      const double *Ps[1] = {P};
      add_xpt_term(npe, Nx, scale * w * det, Phi, 1, Ups, Ps, c, dfdX);
    }
  }
}
