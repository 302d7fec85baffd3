// Element-level driver over the reference's own sources (compiled from
// /root/reference by oracle/build_ref.sh into oracle/_ref/libtacs_ref.so).
// TEST INFRASTRUCTURE: it only produces reference values for the oracle's
// checks and the committed fixtures under tests/golden/.
#include "TACSElement3D.h"
#include "TACSHexaBasis.h"
#include "TACSIsoShellConstitutive.h"
#include "TACSLinearElasticity.h"
#include "TACSMaterialProperties.h"
#include "TACSShellElementDefs.h"
#include "TACSSolidConstitutive.h"
#include "TACSTraction3D.h"
#include "TACSPressure3D.h"
#include "BCSRMatImpl.h"

extern "C" {

// Residual and Jacobian (alpha, beta, gamma) of one 8- or 27-node solid element.
void ref_solid_element(int order, const double* X, const double* u, double E, double nu,
                       double rho, double t, double alpha, double gamma, double* res,
                       double* mat) {
  TACSMaterialProperties* props = new TACSMaterialProperties(rho, 921.0, E, nu, 270e6, 0.0, 0.0);
  TACSSolidConstitutive* con = new TACSSolidConstitutive(props, t, 0);
  TACSLinearElasticity3D* model = new TACSLinearElasticity3D(con, TACS_LINEAR_STRAIN);
  TACSElementBasis* basis;
  if (order == 2)
    basis = new TACSLinearHexaBasis();
  else
    basis = new TACSQuadraticHexaBasis();
  TACSElement3D* elem = new TACSElement3D(model, basis);
  elem->incref();
  int nv = elem->getNumVariables();
  double* zero = new double[nv]();
  for (int i = 0; i < nv; i++) res[i] = 0.0;
  for (int i = 0; i < nv * nv; i++) mat[i] = 0.0;
  elem->addJacobian(0, 0.0, alpha, 0.0, gamma, X, u, zero, zero, res, mat);
  delete[] zero;
  elem->decref();
}

// Matrix-free product of one 8- or 27-node solid element
// (TACSElement3D::getMatVecProductData + addMatVecProduct) for matType
// 0 (Jacobian), 1 (stiffness) or 2 (mass) with coefficients alpha, gamma.
void ref_solid_matvec(int order, int matType, const double* X, const double* x, double E,
                      double nu, double rho, double t, double alpha, double gamma, double* y) {
  TACSMaterialProperties* props = new TACSMaterialProperties(rho, 921.0, E, nu, 270e6, 0.0, 0.0);
  TACSSolidConstitutive* con = new TACSSolidConstitutive(props, t, 0);
  TACSLinearElasticity3D* model = new TACSLinearElasticity3D(con, TACS_LINEAR_STRAIN);
  TACSElementBasis* basis;
  if (order == 2)
    basis = new TACSLinearHexaBasis();
  else
    basis = new TACSQuadraticHexaBasis();
  TACSElement3D* elem = new TACSElement3D(model, basis);
  elem->incref();
  ElementMatrixType mt = matType == 0 ? TACS_JACOBIAN_MATRIX
                                      : (matType == 1 ? TACS_STIFFNESS_MATRIX : TACS_MASS_MATRIX);
  int nv = elem->getNumVariables();
  int dsize = 0, tsize = 0;
  elem->getMatVecDataSizes(mt, 0, &dsize, &tsize);
  double* data = new double[dsize]();
  double* temp = new double[tsize]();
  double* zero = new double[nv]();
  elem->getMatVecProductData(mt, 0, 0.0, alpha, 0.0, gamma, X, zero, zero, zero, data);
  for (int i = 0; i < nv; i++) y[i] = 0.0;
  elem->addMatVecProduct(mt, 0, data, temp, x, y);
  delete[] data;
  delete[] temp;
  delete[] zero;
  elem->decref();
}

// Element matrix of one 8- or 27-node solid element for matType 1 (stiffness),
// 2 (mass) or 3 (geometric stiffness at u) (TACSElement3D::getMatType).
void ref_solid_mat_type(int order, int matType, const double* X, const double* u, double E,
                        double nu, double rho, double t, double* mat) {
  TACSMaterialProperties* props = new TACSMaterialProperties(rho, 921.0, E, nu, 270e6, 0.0, 0.0);
  TACSSolidConstitutive* con = new TACSSolidConstitutive(props, t, 0);
  TACSLinearElasticity3D* model = new TACSLinearElasticity3D(con, TACS_LINEAR_STRAIN);
  TACSElementBasis* basis;
  if (order == 2)
    basis = new TACSLinearHexaBasis();
  else
    basis = new TACSQuadraticHexaBasis();
  TACSElement3D* elem = new TACSElement3D(model, basis);
  elem->incref();
  ElementMatrixType mt = matType == 1 ? TACS_STIFFNESS_MATRIX
                                      : (matType == 2 ? TACS_MASS_MATRIX
                                                      : TACS_GEOMETRIC_STIFFNESS_MATRIX);
  elem->getMatType(mt, 0, 0.0, X, u, mat);
  elem->decref();
}

static TACSElement3D* make_solid(int order, double E, double nu, double t) {
  TACSMaterialProperties* props = new TACSMaterialProperties(2700.0, 921.0, E, nu, 270e6, 0.0, 0.0);
  TACSSolidConstitutive* con = new TACSSolidConstitutive(props, t, 0);
  TACSLinearElasticity3D* model = new TACSLinearElasticity3D(con, TACS_LINEAR_STRAIN);
  TACSElementBasis* basis;
  if (order == 2)
    basis = new TACSLinearHexaBasis();
  else
    basis = new TACSQuadraticHexaBasis();
  return new TACSElement3D(model, basis);
}

// psi^T dR/dX of one solid element (TACSElement3D::addAdjResXptProduct).
void ref_solid_adj_xpt(int order, const double* X, const double* u, const double* psi, double E,
                       double nu, double t, double* out) {
  TACSElement3D* elem = make_solid(order, E, nu, t);
  elem->incref();
  int nv = elem->getNumVariables();
  double* zero = new double[nv]();
  for (int i = 0; i < nv; i++) out[i] = 0.0;
  elem->addAdjResXptProduct(0, 0.0, 1.0, psi, X, u, zero, zero, out);
  delete[] zero;
  elem->decref();
}

// d/dX of the element compliance sum_q w detXd U0, through the element's
// point-quantity interface as the compliance function drives it.
void ref_solid_compliance_xpt(int order, const double* X, const double* u, double E, double nu,
                              double t, double* out) {
  TACSElement3D* elem = make_solid(order, E, nu, t);
  elem->incref();
  int nv = elem->getNumVariables();
  double* zero = new double[nv]();
  for (int i = 0; i < nv; i++) out[i] = 0.0;
  for (int i = 0; i < elem->getNumQuadraturePoints(); i++) {
    double pt[3];
    double w = elem->getQuadraturePoint(i, pt);
    TacsScalar U0 = 0.0, detXd = 0.0;
    elem->evalPointQuantity(0, TACS_STRAIN_ENERGY_DENSITY, 0.0, i, pt, X, u, zero, zero, &detXd,
                            &U0);
    TacsScalar dfdq = w * detXd;
    elem->addPointQuantityXptSens(0, TACS_STRAIN_ENERGY_DENSITY, 0.0, 1.0, i, pt, X, u, zero,
                                  zero, w * U0, &dfdq, out);
  }
  delete[] zero;
  elem->decref();
}

// Traction residual of one solid element face (TACSTraction3D, coordinate
// components).
void ref_solid_traction(int order, int face, const double* X, const double* trac, double* res) {
  TACSElementBasis* basis;
  if (order == 2)
    basis = new TACSLinearHexaBasis();
  else
    basis = new TACSQuadraticHexaBasis();
  basis->incref();
  TACSTraction3D* tr = new TACSTraction3D(3, face, basis, trac, 1);
  tr->incref();
  int nv = 3 * basis->getNumNodes();
  double* zero = new double[nv]();
  for (int i = 0; i < nv; i++) res[i] = 0.0;
  tr->addResidual(0, 0.0, X, zero, zero, zero, res);
  delete[] zero;
  tr->decref();
  basis->decref();
}

// Pressure residual of one solid element face (TACSPressure3D).
void ref_solid_pressure(int order, int face, const double* X, double p, double* res) {
  TACSElementBasis* basis;
  if (order == 2)
    basis = new TACSLinearHexaBasis();
  else
    basis = new TACSQuadraticHexaBasis();
  basis->incref();
  TACSPressure3D* pr = new TACSPressure3D(3, face, basis, p);
  pr->incref();
  int nv = 3 * basis->getNumNodes();
  double* zero = new double[nv]();
  for (int i = 0; i < nv; i++) res[i] = 0.0;
  pr->addResidual(0, 0.0, X, zero, zero, zero, res);
  delete[] zero;
  pr->decref();
  basis->decref();
}

static TACSElement* make_quad4(double E, double nu, double ys, double t) {
  TACSMaterialProperties* props = new TACSMaterialProperties(2700.0, 921.0, E, nu, ys, 0.0, 0.0);
  TACSIsoShellConstitutive* con = new TACSIsoShellConstitutive(props, t, 0);
  TACSShellTransform* transform = new TACSShellNaturalTransform();
  return new TACSQuad4Shell(transform, con);
}

// Residual and Jacobian of one TACSQuad4Shell (natural transform,
// TACSIsoShellConstitutive with thickness t).
void ref_quad4_shell(const double* X, const double* vars, double E, double nu, double ys,
                     double t, double alpha, double* res, double* mat) {
  TACSElement* elem = make_quad4(E, nu, ys, t);
  elem->incref();
  double zero[24] = {0};
  for (int i = 0; i < 24; i++) res[i] = 0.0;
  for (int i = 0; i < 576; i++) mat[i] = 0.0;
  elem->addJacobian(0, 0.0, alpha, 0.0, 0.0, X, vars, zero, zero, res, mat);
  elem->decref();
}

// Element matrix of one TACSQuad4Shell by type (TACSElement::getMatType):
// 1 (stiffness) or 2 (mass) at the state vars.
void ref_quad4_mat_type(int matType, const double* X, const double* vars, double E, double nu,
                        double t, double* mat) {
  TACSElement* elem = make_quad4(E, nu, 270e6, t);
  elem->incref();
  elem->getMatType((ElementMatrixType)matType, 0, 0.0, X, vars, mat);
  elem->decref();
}

// Residual, Jacobian and failure index (point n) of one TACSQuad4Shell with
// the reference-axis transform (TACSShellRefAxisTransform).
void ref_quad4_shell_axis(const double* X, const double* vars, double E, double nu, double ys,
                          double t, const double* axis, double* res, double* mat, double* fail) {
  TACSMaterialProperties* props = new TACSMaterialProperties(2700.0, 921.0, E, nu, ys, 0.0, 0.0);
  TACSIsoShellConstitutive* con = new TACSIsoShellConstitutive(props, t, 0);
  TACSShellTransform* transform = new TACSShellRefAxisTransform(axis);
  TACSElement* elem = new TACSQuad4Shell(transform, con);
  elem->incref();
  double zero[24] = {0};
  for (int i = 0; i < 24; i++) res[i] = 0.0;
  for (int i = 0; i < 576; i++) mat[i] = 0.0;
  elem->addJacobian(0, 0.0, 1.0, 0.0, 0.0, X, vars, zero, zero, res, mat);
  for (int n = 0; n < 4; n++) {
    double pt[3], det = 0.0;
    elem->getQuadraturePoint(n, pt);
    elem->evalPointQuantity(0, TACS_FAILURE_INDEX, 0.0, n, pt, X, vars, zero, zero, &det,
                            &fail[n]);
  }
  elem->decref();
}

// One TACSQuad4Shell whose TACSIsoShellConstitutive has the offset toff:
// residual and Jacobian, mass matrix (getMatType), psi^T dR/dt, and per
// quadrature point the failure index and its thickness derivative.
void ref_quad4_shell_offset(const double* X, const double* vars, const double* psi, double E,
                            double nu, double ys, double t, double toff, double* res,
                            double* mat, double* mass, double* adj, double* fail,
                            double* dfail) {
  TACSMaterialProperties* props = new TACSMaterialProperties(2700.0, 921.0, E, nu, ys, 0.0, 0.0);
  TACSIsoShellConstitutive* con = new TACSIsoShellConstitutive(props, t, 0, 0.0, 1e20, toff);
  TACSElement* elem = new TACSQuad4Shell(new TACSShellNaturalTransform(), con);
  elem->incref();
  double zero[24] = {0};
  for (int i = 0; i < 24; i++) res[i] = 0.0;
  for (int i = 0; i < 576; i++) mat[i] = 0.0;
  elem->addJacobian(0, 0.0, 1.0, 0.0, 0.0, X, vars, zero, zero, res, mat);
  elem->getMatType(TACS_MASS_MATRIX, 0, 0.0, X, vars, mass);
  adj[0] = 0.0;
  elem->addAdjResProduct(0, 0.0, 1.0, psi, X, vars, zero, zero, 1, adj);
  for (int n = 0; n < 4; n++) {
    double pt[3], det = 0.0, one = 1.0;
    elem->getQuadraturePoint(n, pt);
    elem->evalPointQuantity(0, TACS_FAILURE_INDEX, 0.0, n, pt, X, vars, zero, zero, &det,
                            &fail[n]);
    dfail[n] = 0.0;
    elem->addPointQuantityDVSens(0, TACS_FAILURE_INDEX, 0.0, 1.0, n, pt, X, vars, zero, zero,
                                 &one, 1, &dfail[n]);
  }
  elem->decref();
}

// Jacobian (alpha = 1) of one TACSQuad4NonlinearShell at the state vars
// (TACSShellNonlinearModel), with the residual.
void ref_quad4_nonlinear_shell(const double* X, const double* vars, double E, double nu,
                               double t, double* res, double* mat) {
  TACSMaterialProperties* props = new TACSMaterialProperties(2700.0, 921.0, E, nu, 270e6, 0.0, 0.0);
  TACSIsoShellConstitutive* con = new TACSIsoShellConstitutive(props, t, 0);
  TACSElement* elem = new TACSQuad4NonlinearShell(new TACSShellNaturalTransform(), con);
  elem->incref();
  double zero[24] = {0};
  for (int i = 0; i < 24; i++) res[i] = 0.0;
  for (int i = 0; i < 576; i++) mat[i] = 0.0;
  elem->addJacobian(0, 0.0, 1.0, 0.0, 0.0, X, vars, zero, zero, res, mat);
  elem->decref();
}

// psi^T dR/dt of one TACSQuad4Shell (thickness DV, number 0).
double ref_quad4_adj_product(const double* X, const double* vars, const double* psi, double E,
                             double nu, double ys, double t) {
  TACSElement* elem = make_quad4(E, nu, ys, t);
  elem->incref();
  double zero[24] = {0};
  double dfdx[1] = {0.0};
  elem->addAdjResProduct(0, 0.0, 1.0, psi, X, vars, zero, zero, 1, dfdx);
  elem->decref();
  return dfdx[0];
}

// Failure index (KS of the top/bottom von Mises, ksWeight 100 inside the
// constitutive model) at quadrature point n of one TACSQuad4Shell.
double ref_quad4_failure(const double* X, const double* vars, double E, double nu, double ys,
                         double t, int n, double* detXd) {
  TACSElement* elem = make_quad4(E, nu, ys, t);
  elem->incref();
  double zero[24] = {0};
  double pt[3];
  elem->getQuadraturePoint(n, pt);
  double q = 0.0;
  elem->evalPointQuantity(0, TACS_FAILURE_INDEX, 0.0, n, pt, X, vars, zero, zero, detXd, &q);
  elem->decref();
  return q;
}

// Pressure load residual of one TACSQuad4Shell (TACSShellPressure).
void ref_quad4_pressure(const double* X, double p, double* res) {
  TACSElement* elem = make_quad4(70e9, 0.3, 270e6, 0.01);
  elem->incref();
  TACSElement* pr = elem->createElementPressure(0, p);
  pr->incref();
  double zero[24] = {0};
  for (int i = 0; i < 24; i++) res[i] = 0.0;
  pr->addResidual(0, 0.0, X, zero, zero, zero, res);
  pr->decref();
  elem->decref();
}

// Speed of the reference's hot loops, for the oracle's speed ratio against
// them on identical inputs: reps products y = A x through BCSRMatVecMult3 /
// BCSRMatVecMult6 (src/bpmat/BCSRMatMult3.cpp:28, BCSRMatMult6.cpp:82) over
// the caller's arrays (borrowed: handed back before the data object goes).
void ref_bsr_mult(int b, int nrows, int* rowp, int* cols, double* A, double* x, double* y,
                  int reps) {
  BCSRMatData* d = new BCSRMatData(b, nrows, nrows);
  d->incref();
  d->rowp = rowp;
  d->cols = cols;
  d->A = A;
  for (int r = 0; r < reps; r++) {
    if (b == 3)
      BCSRMatVecMult3(d, x, y);
    else
      BCSRMatVecMult6(d, x, y);
  }
  d->rowp = NULL;
  d->cols = NULL;
  d->A = NULL;
  d->decref();
}

// Jacobians of nelem solid elements through one TACSElement3D (the body of
// the reference's element loop, TACSAssembler.cpp:4352-4383 ->
// TACSElement3D::addJacobian): Xe holds 3 npe coordinates per element, mat
// receives the last element's matrix (earlier ones are overwritten).
void ref_solid_jacobian_loop(int order, int nelem, const double* Xe, double E, double nu,
                             double t, double* res, double* mat) {
  TACSElement3D* elem = make_solid(order, E, nu, t);
  elem->incref();
  const int nv = elem->getNumVariables(), npe = nv / 3;
  double* zero = new double[nv]();
  for (int e = 0; e < nelem; e++) {
    for (int i = 0; i < nv; i++) res[i] = 0.0;
    for (int i = 0; i < nv * nv; i++) mat[i] = 0.0;
    elem->addJacobian(0, 0.0, 1.0, 0.0, 0.0, &Xe[(size_t)3 * npe * e], zero, zero, zero, res,
                      mat);
  }
  delete[] zero;
  elem->decref();
}
// Jacobians of nelem TACSQuad4Shell elements through one element object
// (X: 12 coordinates per element); mat receives the last element's matrix.
void ref_quad4_jacobian_loop(int nelem, const double* X, double E, double nu, double t,
                             double* res, double* mat) {
  TACSElement* elem = make_quad4(E, nu, 270e6, t);
  elem->incref();
  double zero[24] = {0};
  for (int e = 0; e < nelem; e++) {
    for (int i = 0; i < 24; i++) res[i] = 0.0;
    for (int i = 0; i < 576; i++) mat[i] = 0.0;
    elem->addJacobian(0, 0.0, 1.0, 0.0, 0.0, &X[(size_t)12 * e], zero, zero, zero, res, mat);
  }
  elem->decref();
}
}  // extern "C"
