/*
  CPU restatement of TACS's linear MITC4 shell (TACSQuad4Shell =
  TACSShellElement<TACSQuadLinearQuadrature, TACSShellQuadBasis<2>,
  TACSLinearizedRotation, TACSShellLinearModel>, src/elements/shell/
  TACSShellElementDefs.h:18-20) with TACSShellNaturalTransform and
  TACSIsoShellConstitutive, in plain C and FP64.

  TEST INFRASTRUCTURE ONLY (part of the oracle library, see tacs_oracle.c):
  loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
  never by the product path.

  Form of the restatement.  With the linear strain model and linearized
  rotations the strain e[9] at a quadrature point is linear in the element
  variables: e = B_q v.  orc_quad4_strains() evaluates the reference's
  forward strain map (node normals, drill strain at the nodes, tying strains,
  displacement gradient; TACSShellElement::addResidual,
  TACSShellElement.h:295-420) for one variable vector; B_q is obtained from
  it column by column, and the element stiffness is
      K = sum_q w_q detXd_q B_q^T Cs B_q,       R = K v,
  which is the reference's element Jacobian / residual for a static state
  (the inertial terms multiply the zero accelerations).

  Pins (tests/test_oracle_shell.py): residual, Jacobian, stiffness and mass
  matrices, thickness adjoint product, failure index, detXd and pressure
  load against values written by
  the reference's own element code (tests/golden/elements.npz); mesh-level
  compliance and KS failure of the 10x10 clamped plate against the golden
  values of tests/integration_tests/test_shell_plate_quad.py:32-46; rigid
  body null space and symmetry of K; central differences for the
  sensitivities.
*/
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SH_NN 4
#define SH_NV 24

/* TacsGaussQuadPts2 literal (src/elements/basis/TACSGaussQuadrature.h:26) */
static const double sh_gp[2] = {-0.577350269189626, 0.577350269189626};

/* Lobatto shape functions of order 2 on {-1, 1}; also the Lagrange functions
   on the tying knots TacsShellLinearTyingPoints = {-1, 1}
   (TACSShellElementQuadBasis.h:116). */
static void sh_lagr(double x, double n[2], double d[2]) {
  n[0] = 0.5 * (1.0 - x);
  n[1] = 0.5 * (1.0 + x);
  d[0] = -0.5;
  d[1] = 0.5;
}

/* interpFields / interpFieldsGrad of TACSShellQuadBasis<2>
   (TACSShellElementQuadBasis.h:170-232): node (i, j) = i + 2 j, field k of
   node a at values[stride*a + k]; grad[2k] = d/dxi, grad[2k+1] = d/deta. */
static void sh_interp(int stride, int m, const double pt[2], const double *v,
                      double *f) {
  double na[2], nb[2], da[2], db[2];
  sh_lagr(pt[0], na, da);
  sh_lagr(pt[1], nb, db);
  for (int k = 0; k < m; k++) f[k] = 0.0;
  for (int j = 0; j < 2; j++) {
    for (int i = 0; i < 2; i++) {
      const double *p = &v[stride * (i + 2 * j)];
      for (int k = 0; k < m; k++) f[k] += na[i] * nb[j] * p[k];
    }
  }
}

static void sh_grad(int stride, int m, const double pt[2], const double *v,
                    double *g) {
  double na[2], nb[2], da[2], db[2];
  sh_lagr(pt[0], na, da);
  sh_lagr(pt[1], nb, db);
  for (int k = 0; k < 2 * m; k++) g[k] = 0.0;
  for (int j = 0; j < 2; j++) {
    for (int i = 0; i < 2; i++) {
      const double *p = &v[stride * (i + 2 * j)];
      for (int k = 0; k < m; k++) {
        g[2 * k] += da[i] * nb[j] * p[k];
        g[2 * k + 1] += na[i] * db[j] * p[k];
      }
    }
  }
}

static void sh_cross(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

static double sh_det3(const double A[9]) {
  return A[8] * (A[0] * A[4] - A[3] * A[1]) -
         A[7] * (A[0] * A[5] - A[3] * A[2]) +
         A[6] * (A[1] * A[5] - A[2] * A[4]);
}

static double sh_inv3(const double A[9], double B[9]) {
  double det = sh_det3(A);
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

/* C = A B and C = A^T B for row-major 3x3 */
static void sh_mm(const double A[9], const double B[9], double C[9]) {
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) {
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
    }
  }
}
static void sh_mtm(const double A[9], const double B[9], double C[9]) {
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) {
      C[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
    }
  }
}

/* Frame [a | b | c] with columns a = (Xxi[0], Xxi[2], Xxi[4]),
   b = (Xxi[1], Xxi[3], Xxi[5]), c = n (zero column when n is NULL):
   TacsShellAssembleFrame (TACSShellUtilities.h:8-35). */
static void sh_frame(const double Xxi[6], const double *n, double Xd[9]) {
  for (int k = 0; k < 3; k++) {
    Xd[3 * k] = Xxi[2 * k];
    Xd[3 * k + 1] = Xxi[2 * k + 1];
    Xd[3 * k + 2] = n ? n[k] : 0.0;
  }
}

/* The element frame in use: the natural transform, or
   TACSShellRefAxisTransform (TACSShellElementTransform.h:95-160) after
   orc_shell_set_ref_axis(axis) (NULL returns to the natural transform).  The
   setting is global to the oracle's shell functions. */
static int sh_ref_axis = 0;
static double sh_axis[3] = {0.0, 0.0, 0.0};

void orc_shell_set_ref_axis(const double *axis) {
  sh_ref_axis = axis != NULL;
  if (axis) {
    double nrm = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    double inv = nrm != 0.0 ? 1.0 / nrm : 0.0;
    for (int k = 0; k < 3; k++) sh_axis[k] = inv * axis[k];
  }
}

/* TACSShellNaturalTransform::computeTransform
   (TACSShellElementTransform.h:25-92).  The reference removes the normal
   component from the first entry of t1 only, three times over; that is what
   it computes, so it is kept. Columns of T: t1, t2 = n x t1, n. */
static void sh_transform(const double Xxi[6], const double n0[3], double T[9]) {
  double n[3] = {n0[0], n0[1], n0[2]};
  double inv = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] *= inv;
  n[1] *= inv;
  n[2] *= inv;
  double t1[3] = {Xxi[0], Xxi[2], Xxi[4]};
  if (sh_ref_axis) {
    /* TACSShellRefAxisTransform::computeTransform (:116-160) */
    double an = sh_axis[0] * n[0] + sh_axis[1] * n[1] + sh_axis[2] * n[2];
    for (int k = 0; k < 3; k++) t1[k] = sh_axis[k] - an * n[k];
  } else {
    double d = n[0] * t1[0] + n[1] * t1[1] + n[2] * t1[2];
    t1[0] = t1[0] - d * n[0];
    t1[0] = t1[0] - d * n[0];
    t1[0] = t1[0] - d * n[0];
  }
  inv = 1.0 / sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
  t1[0] *= inv;
  t1[1] *= inv;
  t1[2] *= inv;
  double t2[3];
  sh_cross(n, t1, t2);
  for (int k = 0; k < 3; k++) {
    T[3 * k] = t1[k];
    T[3 * k + 1] = t2[k];
    T[3 * k + 2] = n[k];
  }
}

/* Parametric location of node i: TACSShellQuadBasis::getNodePoint (:147) */
static void sh_node_pt(int i, double pt[2]) {
  pt[0] = -1.0 + 2.0 * (i % 2);
  pt[1] = -1.0 + 2.0 * (i / 2);
}

/* Tying points in storage order G11(2), G22(2), G12(1), G23(2), G13(2):
   getTyingField / getTyingPoint (TACSShellElementQuadBasis.h:487-560) with
   order knots {-1, 1} and reduced knot {0}. */
static const double sh_ty_pt[9][2] = {{0.0, -1.0}, {0.0, 1.0},  {-1.0, 0.0},
                                      {1.0, 0.0},  {0.0, 0.0},  {-1.0, 0.0},
                                      {1.0, 0.0},  {0.0, -1.0}, {0.0, 1.0}};

/* Node normals fn (unit Xxi_1 x Xxi_2 at the node) and node frames Xdn:
   TacsShellComputeNodeNormals (TACSShellUtilities.h:301-341). */
static void sh_node_normals(const double X[12], double fn[12], double Xdn[36]) {
  for (int i = 0; i < SH_NN; i++) {
    double pt[2], Xxi[6];
    sh_node_pt(i, pt);
    sh_grad(3, 3, pt, X, Xxi);
    double a[3] = {Xxi[0], Xxi[2], Xxi[4]}, b[3] = {Xxi[1], Xxi[3], Xxi[5]};
    double *f = &fn[3 * i];
    sh_cross(a, b, f);
    double nrm = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (nrm != 0.0) {
      f[0] *= 1.0 / nrm;
      f[1] *= 1.0 / nrm;
      f[2] *= 1.0 / nrm;
    }
    if (Xdn) sh_frame(Xxi, f, &Xdn[9 * i]);
  }
}

/*
  Strains e[q][9] at the four quadrature points for element variables v[24]
  (6 per node: u, v, w, q1, q2, q3) and w_q*detXd_q.  drill = 0 sets e[8] = 0
  (the failure evaluation, TACSShellElement::evalPointQuantity).
    - drill strain at the nodes: TacsShellComputeDrillStrain
      (TACSShellUtilities.h:649-700), C = I - q^x
      (TACSLinearizedRotation::computeRotationMat, TACSDirector.h:25-34),
      et = 0.5 (Ct[3] + u0x[3] - Ct[1] - u0x[1]) (TACSDirector.h:562-566);
    - director d = q x fn (computeDirectorRates, TACSDirector.h:210-227);
    - tying strains: TACSShellLinearModel::computeTyingStrain
      (TACSShellElementModel.h:29-80), interpolated with
      interpTyingStrain (TACSShellElementQuadBasis.h:651-675);
    - displacement gradient: TacsShellComputeDispGrad
      (TACSShellUtilities.h:361-430);
    - e = [e0ty0, e0ty3, 2 e0ty1, u1x0, u1x4, u1x1 + u1x3, 2 e0ty4,
      2 e0ty2, et] (TACSShellLinearModel::evalStrain :717-731).
*/
static void sh_strains(const double X[12], const double v[24], int drill, int nonlinear,
                       double e[4][9], double detw[4]);

void orc_quad4_strains(const double X[12], const double v[24], int drill,
                       double e[4][9], double detw[4]) {
  sh_strains(X, v, drill, 0, e, detw);
}

/* Strains of TACSShellNonlinearModel (TACSShellElementModel.h:1067): the
   tying strains with their quadratic terms (computeTyingStrain :1088-1140:
   g11 += Uxi_1.Uxi_1 / 2, g22 += Uxi_2.Uxi_2 / 2, g12 += Uxi_1.Uxi_2 / 2,
   g13 += d0.Uxi_1 / 2, g23 += d0.Uxi_2 / 2) and the bending strains with
   u0x^T u1x (evalStrain :1996-2012); the drill strain as in the linear
   model. */
void orc_quad4_strains_nl(const double X[12], const double v[24], int drill,
                          double e[4][9], double detw[4]) {
  sh_strains(X, v, drill, 1, e, detw);
}

static void sh_strains(const double X[12], const double v[24], int drill, int nonlinear,
                       double e[4][9], double detw[4]) {
  double fn[12], Xdn[36];
  sh_node_normals(X, fn, Xdn);

  double etn[SH_NN];
  for (int i = 0; i < SH_NN; i++) {
    double pt[2], Xxi[6], Tn[9], u0xi[6], Xdinv[9], XdinvT[9], u0d[9];
    sh_node_pt(i, pt);
    const double *Xd = &Xdn[9 * i];
    Xxi[0] = Xd[0];
    Xxi[1] = Xd[1];
    Xxi[2] = Xd[3];
    Xxi[3] = Xd[4];
    Xxi[4] = Xd[6];
    Xxi[5] = Xd[7];
    sh_transform(Xxi, &fn[3 * i], Tn);
    sh_grad(6, 3, pt, v, u0xi);
    sh_inv3(Xd, Xdinv);
    sh_mm(Xdinv, Tn, XdinvT);
    sh_frame(u0xi, NULL, u0d);
    const double *q = &v[6 * i + 3];
    double C[9] = {1.0, q[2], -q[1], -q[2], 1.0, q[0], q[1], -q[0], 1.0};
    double tmp[9], Ct[9], u0x[9];
    sh_mtm(Tn, C, tmp);
    sh_mm(tmp, Tn, Ct);
    sh_mm(u0d, XdinvT, tmp);
    sh_mtm(Tn, tmp, u0x);
    etn[i] = 0.5 * (Ct[3] + u0x[3] - Ct[1] - u0x[1]);
  }

  double d[12];
  for (int i = 0; i < SH_NN; i++) sh_cross(&v[6 * i + 3], &fn[3 * i], &d[3 * i]);

  double ety[9];
  for (int t = 0; t < 9; t++) {
    const double *pt = sh_ty_pt[t];
    double Uxi[6], Xxi[6];
    sh_grad(3, 3, pt, X, Xxi);
    sh_grad(6, 3, pt, v, Uxi);
    const double nl = nonlinear ? 1.0 : 0.0;
    if (t < 2) {
      ety[t] = Uxi[0] * Xxi[0] + Uxi[2] * Xxi[2] + Uxi[4] * Xxi[4] +
               nl * 0.5 * (Uxi[0] * Uxi[0] + Uxi[2] * Uxi[2] + Uxi[4] * Uxi[4]);
    } else if (t < 4) {
      ety[t] = Uxi[1] * Xxi[1] + Uxi[3] * Xxi[3] + Uxi[5] * Xxi[5] +
               nl * 0.5 * (Uxi[1] * Uxi[1] + Uxi[3] * Uxi[3] + Uxi[5] * Uxi[5]);
    } else if (t < 5) {
      ety[t] = 0.5 * (Uxi[0] * Xxi[1] + Uxi[2] * Xxi[3] + Uxi[4] * Xxi[5] +
                      Uxi[1] * Xxi[0] + Uxi[3] * Xxi[2] + Uxi[5] * Xxi[4] +
                      nl * (Uxi[0] * Uxi[1] + Uxi[2] * Uxi[3] + Uxi[4] * Uxi[5]));
    } else {
      double d0[3], n0[3];
      sh_interp(3, 3, pt, d, d0);
      sh_interp(3, 3, pt, fn, n0);
      if (t < 7) { /* g23 */
        ety[t] = 0.5 * (Xxi[1] * d0[0] + Xxi[3] * d0[1] + Xxi[5] * d0[2] +
                        (n0[0] + nl * d0[0]) * Uxi[1] + (n0[1] + nl * d0[1]) * Uxi[3] +
                        (n0[2] + nl * d0[2]) * Uxi[5]);
      } else { /* g13 */
        ety[t] = 0.5 * (Xxi[0] * d0[0] + Xxi[2] * d0[1] + Xxi[4] * d0[2] +
                        (n0[0] + nl * d0[0]) * Uxi[0] + (n0[1] + nl * d0[1]) * Uxi[2] +
                        (n0[2] + nl * d0[2]) * Uxi[4]);
      }
    }
  }

  for (int qi = 0; qi < 4; qi++) {
    double pt[2] = {sh_gp[qi % 2], sh_gp[qi / 2]};
    double Xxi[6], n0[3], T[9], et, nxi[6];
    sh_grad(3, 3, pt, X, Xxi);
    sh_interp(3, 3, pt, fn, n0);
    sh_interp(1, 1, pt, etn, &et);
    sh_transform(Xxi, n0, T);
    sh_grad(3, 3, pt, fn, nxi);
    double Xd[9], Xdz[9], Xdinv[9], XdinvT[9], negz[9], XdinvzT[9];
    sh_frame(Xxi, n0, Xd);
    sh_frame(nxi, NULL, Xdz);
    double det = sh_inv3(Xd, Xdinv);
    sh_mm(Xdinv, Xdz, negz);
    for (int k = 0; k < 9; k++) negz[k] = -negz[k];
    sh_mm(Xdinv, T, XdinvT);
    sh_mm(negz, XdinvT, XdinvzT);
    double d0[3], d0xi[6], u0xi[6];
    sh_interp(3, 3, pt, d, d0);
    sh_grad(3, 3, pt, d, d0xi);
    sh_grad(6, 3, pt, v, u0xi);
    double u0d[9], d0d[9], tmp[9], tmp2[9], u1x[9];
    sh_frame(u0xi, d0, u0d);
    sh_frame(d0xi, NULL, d0d);
    sh_mm(d0d, XdinvT, tmp);
    sh_mm(u0d, XdinvzT, tmp2);
    for (int k = 0; k < 9; k++) tmp[k] += tmp2[k];
    sh_mtm(T, tmp, u1x);

    /* tying strain interpolation: G11/G13 vary along eta, G22/G23 along xi */
    double na[2], nb[2], da[2], db[2];
    sh_lagr(pt[0], na, da);
    sh_lagr(pt[1], nb, db);
    double g[6];
    g[0] = nb[0] * ety[0] + nb[1] * ety[1];
    g[3] = na[0] * ety[2] + na[1] * ety[3];
    g[1] = ety[4];
    g[4] = na[0] * ety[5] + na[1] * ety[6];
    g[2] = nb[0] * ety[7] + nb[1] * ety[8];
    g[5] = 0.0;
    /* e0ty = XdinvT^T G XdinvT (mat3x3SymmTransformTranspose,
       TACSElementAlgebra.h:1094) */
    double G[9] = {g[0], g[1], g[2], g[1], g[3], g[4], g[2], g[4], g[5]};
    double E0[9];
    sh_mm(G, XdinvT, tmp);
    sh_mtm(XdinvT, tmp, E0);
    double *eq = e[qi];
    eq[0] = E0[0];
    eq[1] = E0[4];
    eq[2] = 2.0 * E0[1];
    eq[3] = u1x[0];
    eq[4] = u1x[4];
    eq[5] = u1x[1] + u1x[3];
    if (nonlinear) {
      /* u0x = T^T [u0,xi | d0] Xdinv T (TacsShellComputeDispGrad,
         TACSShellUtilities.h:405-418) */
      double u0x[9];
      sh_mm(u0d, XdinvT, tmp);
      sh_mtm(T, tmp, u0x);
      eq[3] += u0x[0] * u1x[0] + u0x[3] * u1x[3] + u0x[6] * u1x[6];
      eq[4] += u0x[1] * u1x[1] + u0x[4] * u1x[4] + u0x[7] * u1x[7];
      eq[5] += u0x[0] * u1x[1] + u0x[3] * u1x[4] + u0x[6] * u1x[7] + u1x[0] * u0x[1] +
               u1x[3] * u0x[4] + u1x[6] * u0x[7];
    }
    eq[6] = 2.0 * E0[5];
    eq[7] = 2.0 * E0[2];
    eq[8] = drill ? et : 0.0;
    detw[qi] = det; /* quadrature weights are 1 */
  }
}

/* B_q[9][24] by columns of the linear strain map */
static void sh_bmat(const double X[12], int drill, double B[4][9][SH_NV],
                    double detw[4]) {
  double v[SH_NV] = {0}, e[4][9];
  for (int c = 0; c < SH_NV; c++) {
    v[c] = 1.0;
    orc_quad4_strains(X, v, drill, e, detw);
    v[c] = 0.0;
    for (int q = 0; q < 4; q++) {
      for (int k = 0; k < 9; k++) B[q][k][c] = e[q][k];
    }
  }
}

/* Plane-stress Q (evalTangentStiffness2D, TACSMaterialProperties.cpp:323)
   with G = E / (2 (1 + nu)) (:31). */
static void sh_q(double E, double nu, double Q[6]) {
  double D = E / (1.0 - nu * nu);
  Q[0] = D;
  Q[1] = nu * D;
  Q[2] = 0.0;
  Q[3] = D;
  Q[4] = 0.0;
  Q[5] = 0.5 * E / (1.0 + nu);
}

static void sh_symm3(const double A[6], const double x[3], double y[3]) {
  y[0] = A[0] * x[0] + A[1] * x[1] + A[2] * x[2];
  y[1] = A[1] * x[0] + A[3] * x[1] + A[4] * x[2];
  y[2] = A[2] * x[0] + A[4] * x[1] + A[5] * x[2];
}

static const double sh_kcorr = 5.0 / 6.0;
static const double sh_drill_reg = 0.1; /* TACSShellConstitutive.cpp:61 */

/* Stress s = Cs e of TACSIsoShellConstitutive::evalStress
   (TACSIsoShellConstitutive.cpp:153-189, tOffset = 0) through
   computeStress (TACSShellConstitutive.h:133): A = tQ, B = 0, D = t^3/12 Q,
   As = kcorr t G, drill = 0.5 * 0.1 * (As0 + As2). */
/* Thickness offset of TACSIsoShellConstitutive (tOffset, the mid-surface
   at t (1/2 - tOffset) below the top face), global to the oracle's shell
   functions; 0 unless orc_shell_set_offset says otherwise. */
static double sh_toff = 0.0;
void orc_shell_set_offset(double toff) { sh_toff = toff; }

/* computeStress with A = t Q, B = -tOffset t^2 Q, D = t^3 (1/12 + tOffset^2) Q
   (TACSIsoShellConstitutive::evalStress :153-189) */
void orc_shell_stress(double E, double nu, double t, const double e[9], double s[9]) {
  double Q[6], a[3], b[3];
  sh_q(E, nu, Q);
  sh_symm3(Q, e, a);
  sh_symm3(Q, &e[3], b);
  double I = t * t * t / 12.0;
  double As = sh_kcorr * t * Q[5];
  double Bc = -sh_toff * t * t, Dc = sh_toff * sh_toff * t * t * t;
  for (int k = 0; k < 3; k++) {
    s[k] = t * a[k] + Bc * b[k];
    s[3 + k] = I * b[k] + Dc * b[k] + Bc * a[k];
  }
  s[6] = As * e[6];
  s[7] = As * e[7];
  s[8] = 0.5 * sh_drill_reg * (As + As) * e[8];
}

/*
  Residual and Jacobian of one element (TACSShellElement::addResidual :295,
  addJacobian :435): mat += alpha * K (row-major 24x24, node-major dofs),
  res += K v.  Either output may be NULL.
*/
void orc_quad4_element(const double X[12], const double v[24], double E,
                       double nu, double t, double alpha, double *res,
                       double *mat) {
  double B[4][9][SH_NV], detw[4];
  sh_bmat(X, 1, B, detw);
  double K[SH_NV * SH_NV];
  memset(K, 0, sizeof(K));
  for (int q = 0; q < 4; q++) {
    double CB[9][SH_NV];
    for (int c = 0; c < SH_NV; c++) {
      double e[9], s[9];
      for (int k = 0; k < 9; k++) e[k] = B[q][k][c];
      orc_shell_stress(E, nu, t, e, s);
      for (int k = 0; k < 9; k++) CB[k][c] = s[k];
    }
    for (int r = 0; r < SH_NV; r++) {
      for (int c = 0; c < SH_NV; c++) {
        double acc = 0.0;
        for (int k = 0; k < 9; k++) acc += B[q][k][r] * CB[k][c];
        K[SH_NV * r + c] += detw[q] * acc;
      }
    }
  }
  if (mat) {
    for (int i = 0; i < SH_NV * SH_NV; i++) mat[i] += alpha * K[i];
  }
  if (res) {
    for (int r = 0; r < SH_NV; r++) {
      double acc = 0.0;
      for (int c = 0; c < SH_NV; c++) acc += K[SH_NV * r + c] * v[c];
      res[r] += acc;
    }
  }
}

/*
  Consistent mass of one element, added as alpha * M: the kinetic energy
  T = 1/2 sum_q w_q detXd_q (m0 |u0dot|^2 + 2 m1 u0dot.d0dot + m2 |d0dot|^2)
  of TACSShellElement::computeEnergies (TACSShellElement.h:272-284) with the
  mass moments of TACSIsoShellConstitutive::evalMassMoments
  (TACSIsoShellConstitutive.cpp:120-129): m0 = rho t, m1 = -rho t^2 tOffset,
  m2 = rho t^3 (tOffset^2 + 1/12).  u0 interpolates the node displacements, d0 the node
  directors d_a = q_a x fn_a (TACSLinearizedRotation::computeDirectorRates,
  TACSDirector.h:210-227); both are linear in the element variables, so
  M = sum_q w_q detXd_q (m0 Nu^T Nu + m1 (Nu^T Nd + Nd^T Nu) + m2 Nd^T Nd)
  (per quadrature point q) with the 3 x 24
  maps Nu_q, Nd_q built column by column from unit variable vectors.  This is
  the matrix TACSShellElement::getMatType(TACS_MASS_MATRIX) (:644) assembles
  through addJacobian's gamma terms (:605-617) and
  addDirectorJacobian (TACSDirector.h:369-429).
*/
void orc_quad4_mass(const double X[12], double rho, double t, double alpha,
                    double *mat) {
  double fn[12], zero[SH_NV] = {0}, e[4][9], detw[4];
  sh_node_normals(X, fn, NULL);
  orc_quad4_strains(X, zero, 1, e, detw);
  const double m0 = rho * t, m1 = -rho * t * t * sh_toff,
               m2 = rho * t * t * t * (sh_toff * sh_toff + 1.0 / 12.0);
  for (int q = 0; q < 4; q++) {
    double pt[2] = {sh_gp[q % 2], sh_gp[q / 2]};
    double Nu[3][SH_NV], Nd[3][SH_NV];
    for (int c = 0; c < SH_NV; c++) {
      double v[SH_NV] = {0}, dn[12];
      v[c] = 1.0;
      for (int a = 0; a < SH_NN; a++) {
        const double *qa = &v[6 * a + 3];
        sh_cross(qa, &fn[3 * a], &dn[3 * a]);
      }
      double u0[6], d0[3];
      sh_interp(6, 3, pt, v, u0);
      sh_interp(3, 3, pt, dn, d0);
      for (int k = 0; k < 3; k++) {
        Nu[k][c] = u0[k];
        Nd[k][c] = d0[k];
      }
    }
    for (int r = 0; r < SH_NV; r++) {
      for (int c = 0; c < SH_NV; c++) {
        double acc = 0.0;
        for (int k = 0; k < 3; k++)
          acc += m0 * Nu[k][r] * Nu[k][c] + m2 * Nd[k][r] * Nd[k][c] +
                 m1 * (Nu[k][r] * Nd[k][c] + Nd[k][r] * Nu[k][c]);
        mat[SH_NV * r + c] += alpha * detw[q] * acc;
      }
    }
  }
}

/*
  TACSShellNonlinearModel element (TACSQuad4NonlinearShell,
  TACSShellElementDefs.h:58-60) and the geometric stiffness of the linear
  element (TACSShellElement::getMatType(TACS_GEOMETRIC_STIFFNESS_MATRIX),
  TACSShellElement.h:661-908: the directional derivative, along the path
  p = the element variables, of the nonlinear model's tangent stiffness at the
  zero state).  With e(u) = L u + Q(u, u) the nonlinear strain (quadratic in
  u) at a quadrature point, s = C e and U = 1/2 sum_q w detXd e.s:
    B(u)[:, c] = de/du_c = L e_c + 2 Q(u, e_c),
    K_T(u)[r][c] = sum_q w detXd (B(u)[:, r] . C B(u)[:, c] + s(u) . H[:, r, c]),
    H[:, r, c] = 2 Q(e_r, e_c),
    G = dK_T(eps p)/deps at 0
      = sum_q w detXd (B1[:, r] . C L e_c + L e_r . C B1[:, c] + (C L p) . H[:, r, c]),
    B1[:, c] = 2 Q(p, e_c).
  The bilinear terms come from strain evaluations by polarisation,
  2 Q(a, b) = e(a + b) - e(a) - e(b), exact for the quadratic strain.
*/
static void sh_nl_bmat(const double X[12], const double u[SH_NV], double B[4][9][SH_NV]) {
  double eu[4][9], ec[4][9], euc[4][9], detw[4];
  double v[SH_NV];
  sh_strains(X, u, 1, 1, eu, detw);
  for (int c = 0; c < SH_NV; c++) {
    for (int k = 0; k < SH_NV; k++) v[k] = 0.0;
    v[c] = 1.0;
    sh_strains(X, v, 1, 1, ec, detw);
    double el[4][9];
    sh_strains(X, v, 1, 0, el, detw);
    for (int k = 0; k < SH_NV; k++) v[k] = u[k] + (k == c ? 1.0 : 0.0);
    sh_strains(X, v, 1, 1, euc, detw);
    for (int q = 0; q < 4; q++) {
      for (int k = 0; k < 9; k++) B[q][k][c] = el[q][k] + (euc[q][k] - eu[q][k] - ec[q][k]);
    }
  }
}

/* 2 Q(e_r, e_c) at the four points for all pairs: H[q][k][r][c] */
static void sh_nl_hess(const double X[12], double (*H)[9][SH_NV][SH_NV]) {
  static double ec[SH_NV][4][9];
  double detw[4], v[SH_NV] = {0}, erc[4][9];
  for (int c = 0; c < SH_NV; c++) {
    v[c] = 1.0;
    sh_strains(X, v, 1, 1, ec[c], detw);
    v[c] = 0.0;
  }
  for (int r = 0; r < SH_NV; r++) {
    for (int c = r; c < SH_NV; c++) {
      v[r] += 1.0;
      v[c] += 1.0;
      sh_strains(X, v, 1, 1, erc, detw);
      v[r] = 0.0;
      v[c] = 0.0;
      for (int q = 0; q < 4; q++) {
        for (int k = 0; k < 9; k++) {
          double h = erc[q][k] - ec[r][q][k] - ec[c][q][k];
          if (r == c) h = erc[q][k] - 2.0 * ec[r][q][k];
          H[q][k][r][c] = h;
          H[q][k][c][r] = h;
        }
      }
    }
  }
}

/* Residual and tangent stiffness (added, scaled by alpha) of the nonlinear
   model at the state u */
void orc_quad4_nl_element(const double X[12], const double u[24], double E, double nu,
                          double t, double alpha, double *res, double *mat) {
  static double B[4][9][SH_NV], H[4][9][SH_NV][SH_NV];
  double eu[4][9], detw[4];
  sh_strains(X, u, 1, 1, eu, detw);
  sh_nl_bmat(X, u, B);
  if (mat) sh_nl_hess(X, H);
  for (int q = 0; q < 4; q++) {
    double s[9], CB[9][SH_NV];
    orc_shell_stress(E, nu, t, eu[q], s);
    for (int c = 0; c < SH_NV; c++) {
      double b[9], sb[9];
      for (int k = 0; k < 9; k++) b[k] = B[q][k][c];
      orc_shell_stress(E, nu, t, b, sb);
      for (int k = 0; k < 9; k++) CB[k][c] = sb[k];
    }
    if (res) {
      for (int c = 0; c < SH_NV; c++) {
        double acc = 0.0;
        for (int k = 0; k < 9; k++) acc += B[q][k][c] * s[k];
        res[c] += detw[q] * acc;
      }
    }
    if (mat) {
      for (int r = 0; r < SH_NV; r++) {
        for (int c = 0; c < SH_NV; c++) {
          double acc = 0.0;
          for (int k = 0; k < 9; k++) acc += B[q][k][r] * CB[k][c] + s[k] * H[q][k][r][c];
          mat[SH_NV * r + c] += alpha * detw[q] * acc;
        }
      }
    }
  }
}

/* Geometric stiffness along the path p (added, scaled by alpha) */
void orc_quad4_geo(const double X[12], const double p[24], double E, double nu, double t,
                   double alpha, double *mat) {
  static double B0[4][9][SH_NV], Bp[4][9][SH_NV], H[4][9][SH_NV][SH_NV];
  double el[4][9], detw[4], zero[SH_NV] = {0};
  sh_nl_bmat(X, zero, B0);  /* L e_c */
  sh_nl_bmat(X, p, Bp);     /* L e_c + 2 Q(p, e_c) */
  sh_nl_hess(X, H);
  sh_strains(X, p, 1, 0, el, detw);  /* L p */
  for (int q = 0; q < 4; q++) {
    double sp[9], CB0[9][SH_NV];
    orc_shell_stress(E, nu, t, el[q], sp);
    for (int c = 0; c < SH_NV; c++) {
      double b[9], sb[9];
      for (int k = 0; k < 9; k++) b[k] = B0[q][k][c];
      orc_shell_stress(E, nu, t, b, sb);
      for (int k = 0; k < 9; k++) CB0[k][c] = sb[k];
    }
    for (int r = 0; r < SH_NV; r++) {
      for (int c = 0; c < SH_NV; c++) {
        double acc = 0.0;
        for (int k = 0; k < 9; k++) {
          const double b1r = Bp[q][k][r] - B0[q][k][r], b1c = Bp[q][k][c] - B0[q][k][c];
          acc += b1r * CB0[k][c] + CB0[k][r] * b1c + sp[k] * H[q][k][r][c];
        }
        mat[SH_NV * r + c] += alpha * detw[q] * acc;
      }
    }
  }
}

/* psi^T dR/dt of one element: TACSShellElement::addAdjResProduct (:917)
   -> TACSIsoShellConstitutive::addStressDVSens
   (TACSIsoShellConstitutive.cpp:229-251) with e = e(v), psi-strain = e(psi). */
double orc_quad4_adj_product(const double X[12], const double v[24],
                             const double psi[24], double E, double nu,
                             double t) {
  double eu[4][9], ep[4][9], detw[4], Q[6];
  orc_quad4_strains(X, v, 1, eu, detw);
  orc_quad4_strains(X, psi, 1, ep, detw);
  sh_q(E, nu, Q);
  /* addStressDVSens (:229-251): d/dt of A, B, D */
  double dI = (3.0 * sh_toff * sh_toff + 0.25) * t * t, dB = -2.0 * sh_toff * t, out = 0.0;
  for (int q = 0; q < 4; q++) {
    double a[3], b[3];
    sh_symm3(Q, ep[q], a);
    sh_symm3(Q, &ep[q][3], b);
    double val = a[0] * eu[q][0] + a[1] * eu[q][1] + a[2] * eu[q][2] +
                 dI * (b[0] * eu[q][3] + b[1] * eu[q][4] + b[2] * eu[q][5]) +
                 dB * (a[0] * eu[q][3] + a[1] * eu[q][4] + a[2] * eu[q][5] +
                       b[0] * eu[q][0] + b[1] * eu[q][1] + b[2] * eu[q][2]) +
                 sh_kcorr * Q[5] *
                     (ep[q][6] * eu[q][6] + ep[q][7] * eu[q][7] +
                      sh_drill_reg * ep[q][8] * eu[q][8]);
    out += detw[q] * val;
  }
  return out;
}

/* Von Mises of a plane-stress state (vonMisesFailure2D,
   TACSMaterialProperties.cpp:481) and its stress gradient (:487-500). */
static double sh_vm2(const double s[3], double ys, double sens[3]) {
  double f = sqrt(s[0] * s[0] + s[1] * s[1] - s[0] * s[1] + 3.0 * s[2] * s[2]);
  if (sens) {
    if (f != 0.0) {
      sens[0] = (s[0] - 0.5 * s[1]) / (f * ys);
      sens[1] = (s[1] - 0.5 * s[0]) / (f * ys);
      sens[2] = 3.0 * s[2] / (f * ys);
    } else {
      sens[0] = sens[1] = sens[2] = 0.0;
    }
  }
  return f / ys;
}

/*
  Failure index of TACSIsoShellConstitutive (evalFailure :254-296): KS
  (weight 100) of the top and bottom surface von Mises; with dfde/dfdt the
  strain gradient (evalFailureStrainSens :300-374) and the thickness
  derivative (addFailureDVSens :381-430).
*/
double orc_shell_failure(double E, double nu, double ys, double t,
                         const double e[9], double *dfde, double *dfdt) {
  const double ksw = 100.0;
  double Q[6], et[3], eb[3], st[3], sb[3], pt[3], pb[3];
  sh_q(E, nu, Q);
  double ht = (0.5 - sh_toff) * t, hb = -(0.5 + sh_toff) * t;
  for (int k = 0; k < 3; k++) {
    et[k] = e[k] + ht * e[3 + k];
    eb[k] = e[k] + hb * e[3 + k];
  }
  sh_symm3(Q, et, st);
  sh_symm3(Q, eb, sb);
  double top = sh_vm2(st, ys, pt), bot = sh_vm2(sb, ys, pb);
  double mx = top > bot ? top : bot;
  double sum = exp(ksw * (top - mx)) + exp(ksw * (bot - mx));
  double val = mx + log(sum) / ksw;
  if (dfde || dfdt) {
    double phit[3], phib[3];
    sh_symm3(Q, pt, phit);
    sh_symm3(Q, pb, phib);
    double ft = exp(ksw * (top - mx)) / sum, fb = exp(ksw * (bot - mx)) / sum;
    if (dfde) {
      for (int k = 0; k < 3; k++) {
        dfde[k] = ft * phit[k] + fb * phib[k];
        dfde[3 + k] = ft * ht * phit[k] + fb * hb * phib[k];
      }
      dfde[6] = dfde[7] = dfde[8] = 0.0;
    }
    if (dfdt) {
      *dfdt = ft * (0.5 - sh_toff) * (phit[0] * e[3] + phit[1] * e[4] + phit[2] * e[5]) -
              fb * (0.5 + sh_toff) * (phib[0] * e[3] + phib[1] * e[4] + phib[2] * e[5]);
    }
  }
  return val;
}

/* Failure index at quadrature point n with the drill strain set to zero
   (TACSShellElement::evalPointQuantity, TACS_FAILURE_INDEX); *detXd is the
   unweighted Jacobian determinant. */
double orc_quad4_failure(const double X[12], const double v[24], double E,
                         double nu, double ys, double t, int n, double *detXd) {
  double e[4][9], detw[4];
  orc_quad4_strains(X, v, 0, e, detw);
  if (detXd) *detXd = detw[n];
  return orc_shell_failure(E, nu, ys, t, e[n], NULL, NULL);
}

/* Uniform pressure load (TACSShellPressure::addResidual,
   TACSShellPressure.h:51-91): res += N^T f, f = -p detXd n(q) with n the
   interpolated (unnormalised) node normal. */
void orc_quad4_pressure(const double X[12], double p, double res[24]) {
  double fn[12];
  sh_node_normals(X, fn, NULL);
  for (int q = 0; q < 4; q++) {
    double pt[2] = {sh_gp[q % 2], sh_gp[q / 2]};
    double Xxi[6], n[3], Xd[9];
    sh_interp(3, 3, pt, fn, n);
    sh_grad(3, 3, pt, X, Xxi);
    sh_frame(Xxi, n, Xd);
    double det = sh_det3(Xd);
    double na[2], nb[2], da[2], db[2];
    sh_lagr(pt[0], na, da);
    sh_lagr(pt[1], nb, db);
    for (int j = 0; j < 2; j++) {
      for (int i = 0; i < 2; i++) {
        double N = na[i] * nb[j];
        for (int k = 0; k < 3; k++) res[6 * (i + 2 * j) + k] += N * (p * -det * n[k]);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* Mesh-level routines (4-node elements, 6 dof per node)               */

static void sh_gather(const int *c, const double *X, const double *u,
                      double Xe[12], double ue[24]) {
  for (int a = 0; a < SH_NN; a++) {
    for (int k = 0; k < 3; k++) Xe[3 * a + k] = X[3 * c[a] + k];
    if (ue) {
      for (int k = 0; k < 6; k++) ue[6 * a + k] = u ? u[6 * c[a] + k] : 0.0;
    }
  }
}

static long sh_find_slot(const int *rowp, const int *cols, int row, int col) {
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
  Element loop of TACSAssembler::assembleJacobian / assembleRes
  (src/TACSAssembler.cpp:4352-4383, :4190-4227) for Quad4 shells into 6x6
  BCSR with BCSRMat::addRowValues' add rule (src/bpmat/BCSRMat.cpp:
  1800-1840).  Accumulates into A / res (either may be NULL); returns the
  number of missing pattern entries.
*/
long orc_assemble_shell(int nelem, const int *conn, const double *X,
                        const double *u, const double *t_elem, double E,
                        double nu, const int *rowp, const int *cols, double *A,
                        double *res) {
  long missing = 0;
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[4 * el];
    double Xe[12], ue[24], re[24] = {0}, ke[SH_NV * SH_NV];
    sh_gather(c, X, u, Xe, ue);
    if (A) memset(ke, 0, sizeof(ke));
    orc_quad4_element(Xe, ue, E, nu, t_elem[el], 1.0, res ? re : NULL, A ? ke : NULL);
    if (res) {
      for (int a = 0; a < SH_NN; a++) {
        for (int k = 0; k < 6; k++) res[6 * c[a] + k] += re[6 * a + k];
      }
    }
    if (A) {
      for (int a = 0; a < SH_NN; a++) {
        for (int b = 0; b < SH_NN; b++) {
          long s = sh_find_slot(rowp, cols, c[a], c[b]);
          if (s < 0) {
            missing++;
            continue;
          }
          double *blk = &A[36 * s];
          for (int r = 0; r < 6; r++) {
            for (int q = 0; q < 6; q++) blk[6 * r + q] += ke[(6 * a + r) * SH_NV + 6 * b + q];
          }
        }
      }
    }
  }
  return missing;
}

/*
  kc K + mc M + gc G (G: geometric stiffness along the state u, NULL = 0)
  assembled element by element (TACSAssembler::assembleMatType /
  assembleMatCombo, src/TACSAssembler.cpp:4418-4560, over
  TACSShellElement::getMatType) into 6x6 BCSR; returns the number of missing
  pattern entries.
*/
long orc_assemble_shell_mat(int nelem, const int *conn, const double *X,
                            const double *u, const double *t_elem, double E, double nu,
                            double rho, double kc, double mc, double gc, const int *rowp,
                            const int *cols, double *A) {
  long missing = 0;
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[4 * el];
    double Xe[12], ue[24], ke[SH_NV * SH_NV], zero[SH_NV] = {0};
    sh_gather(c, X, u, Xe, ue);
    memset(ke, 0, sizeof(ke));
    orc_quad4_element(Xe, zero, E, nu, t_elem[el], kc, NULL, ke);
    orc_quad4_mass(Xe, rho, t_elem[el], mc, ke);
    if (gc != 0.0) orc_quad4_geo(Xe, ue, E, nu, t_elem[el], gc, ke);
    for (int a = 0; a < SH_NN; a++) {
      for (int b = 0; b < SH_NN; b++) {
        long s = sh_find_slot(rowp, cols, c[a], c[b]);
        if (s < 0) {
          missing++;
          continue;
        }
        double *blk = &A[36 * s];
        for (int r = 0; r < 6; r++) {
          for (int q = 0; q < 6; q++) blk[6 * r + q] += ke[(6 * a + r) * SH_NV + 6 * b + q];
        }
      }
    }
  }
  return missing;
}

/* Pressure aux elements over the listed elements (one pressure each). */
void orc_shell_add_pressure(int naux, const int *aux_elem, const double *p,
                            const int *conn, const double *X, double *res) {
  for (int i = 0; i < naux; i++) {
    const int *c = &conn[4 * aux_elem[i]];
    double Xe[12], re[24] = {0};
    sh_gather(c, X, NULL, Xe, NULL);
    orc_quad4_pressure(Xe, p[i], re);
    for (int a = 0; a < SH_NN; a++) {
      for (int k = 0; k < 6; k++) res[6 * c[a] + k] += re[6 * a + k];
    }
  }
}

/* Compliance sum_e sum_q w detXd e.s (TACSCompliance::elementWiseEval,
   src/functions/TACSCompliance.cpp:69, strain energy density of
   TACSShellElement::evalPointQuantity: drill strain excluded). */
double orc_shell_compliance(int nelem, const int *conn, const double *X,
                            const double *u, const double *t_elem, double E,
                            double nu) {
  double total = 0.0;
  for (int el = 0; el < nelem; el++) {
    double Xe[12], ue[24], e[4][9], detw[4];
    sh_gather(&conn[4 * el], X, u, Xe, ue);
    orc_quad4_strains(Xe, ue, 0, e, detw);
    for (int q = 0; q < 4; q++) {
      double s[9], w = 0.0;
      orc_shell_stress(E, nu, t_elem[el], e[q], s);
      for (int k = 0; k < 9; k++) w += e[q][k] * s[k];
      total += detw[q] * w;
    }
  }
  return total;
}

/* Continuous KS aggregate of the failure index (TACSKSFailure::
   elementWiseEval src/functions/TACSKSFailure.cpp:125-175, getFunctionValue
   :74): out = {KS, max, sum}. */
void orc_shell_ks_failure(int nelem, const int *conn, const double *X,
                          const double *u, const double *t_elem, double E,
                          double nu, double ys, double ksWeight, double alpha,
                          double safety, double out[3]) {
  double maxf = -1e20, sum = 0.0;
  for (int pass = 0; pass < 2; pass++) {
    for (int el = 0; el < nelem; el++) {
      double Xe[12], ue[24], e[4][9], detw[4];
      sh_gather(&conn[4 * el], X, u, Xe, ue);
      orc_quad4_strains(Xe, ue, 0, e, detw);
      for (int q = 0; q < 4; q++) {
        double f = safety * orc_shell_failure(E, nu, ys, t_elem[el], e[q], NULL, NULL);
        if (pass == 0) {
          if (f > maxf) maxf = f;
        } else {
          sum += detw[q] * exp(ksWeight * (f - maxf));
        }
      }
    }
  }
  out[0] = maxf + log(sum / alpha) / ksWeight;
  out[1] = maxf;
  out[2] = sum;
}

/* dKS/du (TACSKSFailure::getElementSVSens :221-283 ->
   TACSShellElement::addPointQuantitySVSens -> evalFailureStrainSens and the
   transposed strain map) and dKS/dt_e (addElementDVSens ->
   addPointQuantityDVSens -> addFailureDVSens).  Either output may be NULL. */
void orc_shell_ks_sens(int nelem, const int *conn, const double *X,
                       const double *u, const double *t_elem, double E,
                       double nu, double ys, double ksWeight, double safety,
                       double maxFail, double ksSum, double *dfdu,
                       double *dfdt) {
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[4 * el];
    double Xe[12], ue[24], e[4][9], detw[4], B[4][9][SH_NV];
    sh_gather(c, X, u, Xe, ue);
    orc_quad4_strains(Xe, ue, 0, e, detw);
    if (dfdu) sh_bmat(Xe, 0, B, detw);
    for (int q = 0; q < 4; q++) {
      double de[9], dt;
      double f = safety * orc_shell_failure(E, nu, ys, t_elem[el], e[q], de, &dt);
      double coef = exp(ksWeight * (f - maxFail)) / ksSum * detw[q] * safety;
      if (dfdt) dfdt[el] += coef * dt;
      if (dfdu) {
        for (int a = 0; a < SH_NV; a++) {
          double acc = 0.0;
          for (int k = 0; k < 9; k++) acc += B[q][k][a] * de[k];
          dfdu[6 * c[a / 6] + a % 6] += coef * acc;
        }
      }
    }
  }
}

/* dfdx[e] += scale * psi_e^T dR_e/dt_e over the mesh
   (TACSAssembler::addAdjointResProducts src/TACSAssembler.cpp:5033). */
void orc_shell_adj_res_product(int nelem, const int *conn, const double *X,
                               const double *u, const double *psi,
                               const double *t_elem, double E, double nu,
                               double scale, double *dfdx) {
  for (int el = 0; el < nelem; el++) {
    double Xe[12], ue[24], pe[24];
    sh_gather(&conn[4 * el], X, u, Xe, ue);
    sh_gather(&conn[4 * el], X, psi, Xe, pe);
    dfdx[el] += scale * orc_quad4_adj_product(Xe, ue, pe, E, nu, t_elem[el]);
  }
}

/* dC/du of the compliance (TACSShellElement::addPointQuantitySVSens,
   TACS_STRAIN_ENERGY_DENSITY: 2 s with e[8] = 0 through the transposed
   strain map) and dC/dt_e at fixed u (addPointQuantityDVSens ->
   TACSIsoShellConstitutive::addStressDVSens(e, e), :229).  Either output
   may be NULL; both accumulate. */
void orc_shell_compliance_sens(int nelem, const int *conn, const double *X,
                               const double *u, const double *t_elem, double E,
                               double nu, double *dfdu, double *dfdt) {
  double Q[6];
  sh_q(E, nu, Q);
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[4 * el];
    double Xe[12], ue[24], e[4][9], detw[4], B[4][9][SH_NV];
    sh_gather(c, X, u, Xe, ue);
    orc_quad4_strains(Xe, ue, 0, e, detw);
    if (dfdu) sh_bmat(Xe, 1, B, detw);
    double t = t_elem[el];
    for (int q = 0; q < 4; q++) {
      double s[9];
      orc_shell_stress(E, nu, t, e[q], s);
      if (dfdu) {
        for (int a = 0; a < SH_NV; a++) {
          double acc = 0.0;
          for (int k = 0; k < 9; k++) acc += B[q][k][a] * 2.0 * s[k];
          dfdu[6 * c[a / 6] + a % 6] += detw[q] * acc;
        }
      }
      if (dfdt) {
        double a3[3], b3[3];
        sh_symm3(Q, e[q], a3);
        sh_symm3(Q, &e[q][3], b3);
        const double dI = (3.0 * sh_toff * sh_toff + 0.25) * t * t;
        dfdt[el] += detw[q] * (a3[0] * e[q][0] + a3[1] * e[q][1] + a3[2] * e[q][2] +
                               dI * (b3[0] * e[q][3] + b3[1] * e[q][4] + b3[2] * e[q][5]) -
                               4.0 * sh_toff * t *
                                   (a3[0] * e[q][3] + a3[1] * e[q][4] + a3[2] * e[q][5]) +
                               sh_kcorr * Q[5] * (e[q][6] * e[q][6] + e[q][7] * e[q][7]));
      }
    }
  }
}

/* ------------------------------------------------------------------------
   Node-coordinate sensitivities.  The reference's TACSShellElement has no
   analytic node derivatives: it inherits TACSElement::addAdjResXptProduct
   and addPointQuantityXptSens (src/elements/TACSElement.cpp:270, :746), which
   difference the residual / point quantity in each coordinate (forward steps
   of 1e-7 in the real build, complex steps of 1e-30 in the complex build).
   The restatement differences the same functions with a sixth-order central
   stencil (steps of 1e-3 times the element size; truncation and rounding
   both near 1e-12 relative), pinned against the complex-step values of the
   reference's complex build (tests/golden/elements.npz).
   ------------------------------------------------------------------------ */
typedef double (*sh_xfun)(const double X[12], const void *ctx);

static void sh_xgrad(sh_xfun f, const void *ctx, const double X[12], double g[12]) {
  static const double cw[3] = {45.0, -9.0, 1.0};
  double L = 0.0;
  for (int a = 0; a < SH_NN; a++) {
    for (int b = a + 1; b < SH_NN; b++) {
      double d2 = 0.0;
      for (int k = 0; k < 3; k++) d2 += (X[3 * a + k] - X[3 * b + k]) * (X[3 * a + k] - X[3 * b + k]);
      if (d2 > L) L = d2;
    }
  }
  const double h = 1e-3 * sqrt(L);
  for (int k = 0; k < 12; k++) {
    double Xp[12], acc = 0.0;
    memcpy(Xp, X, sizeof(Xp));
    for (int j = 0; j < 3; j++) {
      Xp[k] = X[k] + (j + 1) * h;
      double fp = f(Xp, ctx);
      Xp[k] = X[k] - (j + 1) * h;
      double fm = f(Xp, ctx);
      acc += cw[j] * (fp - fm);
    }
    g[k] = acc / (60.0 * h);
  }
}

typedef struct {
  const double *v, *psi;
  double E, nu, ys, t, p;
  const double *dfddet, *dfdq;
  int quantity;
} sh_xctx;

static double sh_x_adj(const double X[12], const void *c) {
  const sh_xctx *x = (const sh_xctx *)c;
  double res[SH_NV] = {0}, r = 0.0;
  orc_quad4_element(X, x->v, x->E, x->nu, x->t, 1.0, res, NULL);
  for (int i = 0; i < SH_NV; i++) r += x->psi[i] * res[i];
  return r;
}

static double sh_x_pres(const double X[12], const void *c) {
  const sh_xctx *x = (const sh_xctx *)c;
  double res[SH_NV] = {0}, r = 0.0;
  orc_quad4_pressure(X, x->p, res);
  for (int i = 0; i < SH_NV; i++) r += x->psi[i] * res[i];
  return r;
}

/* sum_q dfddet_q detXd_q + dfdq_q q_q with q the strain energy density e.s
   (quantity 2, TACS_STRAIN_ENERGY_DENSITY) or the failure index (quantity 1),
   both with the drill strain dropped (TACSShellElement::evalPointQuantity) */
static double sh_x_point(const double X[12], const void *c) {
  const sh_xctx *x = (const sh_xctx *)c;
  double e[4][9], detw[4], r = 0.0;
  orc_quad4_strains(X, x->v, 0, e, detw);
  for (int q = 0; q < 4; q++) {
    double val;
    if (x->quantity == 1) {
      val = orc_shell_failure(x->E, x->nu, x->ys, x->t, e[q], NULL, NULL);
    } else {
      double s[9];
      orc_shell_stress(x->E, x->nu, x->t, e[q], s);
      val = 0.0;
      for (int k = 0; k < 9; k++) val += e[q][k] * s[k];
    }
    r += x->dfddet[q] * detw[q] + x->dfdq[q] * val;
  }
  return r;
}

/* scale psi^T dR/dX of one element (TACSElement::addAdjResXptProduct) */
void orc_quad4_adj_xpt(const double X[12], const double v[24], const double psi[24],
                       double E, double nu, double t, double scale, double out[12]) {
  sh_xctx c = {v, psi, E, nu, 0.0, t, 0.0, NULL, NULL, 0};
  sh_xgrad(sh_x_adj, &c, X, out);
  for (int k = 0; k < 12; k++) out[k] *= scale;
}

/* scale psi^T dF/dX of the element's pressure load (TACSShellPressure) */
void orc_quad4_pressure_xpt(const double X[12], double p, const double psi[24], double scale,
                            double out[12]) {
  sh_xctx c = {NULL, psi, 0.0, 0.0, 0.0, 0.0, p, NULL, NULL, 0};
  sh_xgrad(sh_x_pres, &c, X, out);
  for (int k = 0; k < 12; k++) out[k] *= scale;
}

/* d/dX sum_q (dfddet_q detXd_q + dfdq_q q_q) (TACSElement::addPointQuantityXptSens
   with scale 1, summed over the quadrature points) */
void orc_quad4_point_xpt(int quantity, const double X[12], const double v[24], double E,
                         double nu, double ys, double t, const double dfddet[4],
                         const double dfdq[4], double out[12]) {
  sh_xctx c = {v, NULL, E, nu, ys, t, 0.0, dfddet, dfdq, quantity};
  sh_xgrad(sh_x_point, &c, X, out);
}

static void sh_scatter_x(const int *c, const double g[12], double *out) {
  for (int a = 0; a < SH_NN; a++) {
    for (int k = 0; k < 3; k++) out[3 * c[a] + k] += g[3 * a + k];
  }
}

/*
  Mesh sums (TACSAssembler::addAdjointResXptSensProducts src/TACSAssembler.cpp:
  5136 and addXptSens :4858 over the element routines above):
    mode 0: scale psi^T dR/dX (element residuals),
    mode 1: compliance, sum_q w detXd e.s (TACSCompliance::getElementXptSens),
    mode 2: continuous KS failure with the aggregate's max and sum
            (TACSKSFailure::getElementXptSens, KS_CONTINUOUS: dfddetXd =
            w exp(k (f - max)) / (k sum), dfdq = safety w detXd exp(..) / sum).
  out (3 per node) is added to.
*/
void orc_shell_xpt_sens(int mode, int nelem, const int *conn, const double *X,
                        const double *u, const double *psi, const double *t_elem, double E,
                        double nu, double ys, double ksWeight, double maxFail, double ksSum,
                        double safety, double scale, double *out) {
  for (int el = 0; el < nelem; el++) {
    const int *c = &conn[4 * el];
    double Xe[12], ue[24], pe[24], g[12];
    sh_gather(c, X, u, Xe, ue);
    if (mode == 0) {
      sh_gather(c, X, psi, Xe, pe);
      orc_quad4_adj_xpt(Xe, ue, pe, E, nu, t_elem[el], scale, g);
    } else {
      double e[4][9], detw[4], dfddet[4], dfdq[4];
      orc_quad4_strains(Xe, ue, 0, e, detw);
      for (int q = 0; q < 4; q++) {
        if (mode == 1) {
          double s[9], U = 0.0;
          orc_shell_stress(E, nu, t_elem[el], e[q], s);
          for (int k = 0; k < 9; k++) U += e[q][k] * s[k];
          dfddet[q] = scale * U;
          dfdq[q] = scale * detw[q];
        } else {
          double f = safety * orc_shell_failure(E, nu, ys, t_elem[el], e[q], NULL, NULL);
          double ef = exp(ksWeight * (f - maxFail)) / ksSum;
          dfddet[q] = scale * ef / ksWeight;
          dfdq[q] = scale * safety * detw[q] * ef;
        }
      }
      orc_quad4_point_xpt(mode == 1 ? 2 : 1, Xe, ue, E, nu, ys, t_elem[el], dfddet, dfdq, g);
    }
    sh_scatter_x(c, g, out);
  }
}

/* scale psi^T dF/dX of pressure aux elements (one pressure each) */
void orc_shell_pressure_xpt(int naux, const int *aux_elem, const double *p, const int *conn,
                            const double *X, const double *psi, double scale, double *out) {
  for (int i = 0; i < naux; i++) {
    const int *c = &conn[4 * aux_elem[i]];
    double Xe[12], pe[24], g[12];
    sh_gather(c, X, psi, Xe, pe);
    orc_quad4_pressure_xpt(Xe, p[i], pe, scale, g);
    sh_scatter_x(c, g, out);
  }
}
