// Element-level node-coordinate sensitivities of TACSQuad4Shell from the
// reference sources built with TACS_USE_COMPLEX (oracle/build_ref.sh): the
// reference's default TACSElement::addAdjResXptProduct and
// addPointQuantityXptSens (src/elements/TACSElement.cpp:270, :746) then take
// complex steps of 1e-30, i.e. the exact derivatives to rounding.  Inputs and
// outputs are real arrays.  Test infrastructure: writes fixtures only.
#include <complex>
#include <cstring>

#include "TACSElementTypes.h"
#include "TACSIsoShellConstitutive.h"
#include "TACSMaterialProperties.h"
#include "TACSShellElementDefs.h"

typedef std::complex<double> cplx;

static TACSElement* make_quad4_c(double E, double nu, double ys, double t) {
  TACSMaterialProperties* props = new TACSMaterialProperties(2700.0, 921.0, E, nu, ys, 0.0, 0.0);
  TACSIsoShellConstitutive* con = new TACSIsoShellConstitutive(props, t, 0);
  TACSShellTransform* transform = new TACSShellNaturalTransform();
  return new TACSQuad4Shell(transform, con);
}

static void to_c(int n, const double* a, cplx* b) {
  for (int i = 0; i < n; i++) b[i] = cplx(a[i], 0.0);
}

extern "C" {

// out[12] = scale psi^T dR/dX of one element (TACSElement::addAdjResXptProduct)
void refc_quad4_adj_xpt(const double* X, const double* vars, const double* psi, double E,
                        double nu, double t, double scale, double* out) {
  TACSElement* elem = make_quad4_c(E, nu, 270e6, t);
  elem->incref();
  cplx Xc[12], vc[24], pc[24], zero[24], f[12];
  to_c(12, X, Xc);
  to_c(24, vars, vc);
  to_c(24, psi, pc);
  for (int i = 0; i < 24; i++) zero[i] = 0.0;
  for (int i = 0; i < 12; i++) f[i] = 0.0;
  elem->addAdjResXptProduct(0, 0.0, cplx(scale, 0.0), pc, Xc, vc, zero, zero, f);
  for (int i = 0; i < 12; i++) out[i] = f[i].real();
  elem->decref();
}

// Pressure load of one element (TACSShellPressure through createElementPressure):
// out[12] = scale psi^T dF/dX
void refc_quad4_pressure_xpt(const double* X, double p, const double* psi, double scale,
                             double* out) {
  TACSElement* elem = make_quad4_c(70e9, 0.3, 270e6, 0.01);
  elem->incref();
  TACSElement* pr = elem->createElementPressure(0, cplx(p, 0.0));
  pr->incref();
  cplx Xc[12], pc[24], zero[24], f[12];
  to_c(12, X, Xc);
  to_c(24, psi, pc);
  for (int i = 0; i < 24; i++) zero[i] = 0.0;
  for (int i = 0; i < 12; i++) f[i] = 0.0;
  pr->addAdjResXptProduct(0, 0.0, cplx(scale, 0.0), pc, Xc, zero, zero, zero, f);
  for (int i = 0; i < 12; i++) out[i] = f[i].real();
  pr->decref();
  elem->decref();
}

// d/dX of sum_q w_q (a_q detXd_q + b_q q_q) for one point quantity (the
// element loops of TACSCompliance::getElementXptSens
// src/functions/TACSCompliance.cpp:129 and TACSKSFailure::getElementXptSens
// src/functions/TACSKSFailure.cpp:289 call addPointQuantityXptSens with
// dfddetXd and dfdq per point); quantity 2 = TACS_STRAIN_ENERGY_DENSITY,
// 1 = TACS_FAILURE_INDEX.  dfddetXd[4], dfdq[4] per quadrature point.
void refc_quad4_point_xpt(int quantity, const double* X, const double* vars, double E, double nu,
                          double ys, double t, const double* dfddetXd, const double* dfdq,
                          double* out) {
  TACSElement* elem = make_quad4_c(E, nu, ys, t);
  elem->incref();
  cplx Xc[12], vc[24], zero[24], f[12];
  to_c(12, X, Xc);
  to_c(24, vars, vc);
  for (int i = 0; i < 24; i++) zero[i] = 0.0;
  for (int i = 0; i < 12; i++) f[i] = 0.0;
  const int type = quantity == 1 ? TACS_FAILURE_INDEX : TACS_STRAIN_ENERGY_DENSITY;
  for (int n = 0; n < elem->getNumQuadraturePoints(); n++) {
    double pt[3];
    elem->getQuadraturePoint(n, pt);
    cplx dq = cplx(dfdq[n], 0.0);
    elem->addPointQuantityXptSens(0, type, 0.0, cplx(1.0, 0.0), n, pt, Xc, vc, zero, zero,
                                  cplx(dfddetXd[n], 0.0), &dq, f);
  }
  for (int i = 0; i < 12; i++) out[i] = f[i].real();
  elem->decref();
}

// Point quantity and detXd (real parts) at quadrature point n.
double refc_quad4_point(int quantity, const double* X, const double* vars, double E, double nu,
                        double ys, double t, int n, double* detXd) {
  TACSElement* elem = make_quad4_c(E, nu, ys, t);
  elem->incref();
  cplx Xc[12], vc[24], zero[24], q = 0.0, det = 0.0;
  to_c(12, X, Xc);
  to_c(24, vars, vc);
  for (int i = 0; i < 24; i++) zero[i] = 0.0;
  double pt[3];
  elem->getQuadraturePoint(n, pt);
  const int type = quantity == 1 ? TACS_FAILURE_INDEX : TACS_STRAIN_ENERGY_DENSITY;
  elem->evalPointQuantity(0, type, 0.0, n, pt, Xc, vc, zero, zero, &det, &q);
  *detXd = det.real();
  elem->decref();
  return q.real();
}
}  // extern "C"
