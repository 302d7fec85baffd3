// A C++ driver of the kind written against the reference's headers (cf.
// src/TACSAssembler.h, src/bpmat/KSM.h, src/elements/TACSElement3D.h), built
// against include/TACSHip.h and tacs_amd/lib/libtacs_hip.so instead: an
// n x n x n cube of 8-node hexahedra (E = 70e9, nu = 0.3), x = 0 clamped, a
// 1e6 Pa z-traction on x = 1; assembleJacobian, GMRES (non-flexible and
// flexible), compliance, its adjoint and the thickness gradient.  The only
// line that differs from a reference driver is the preconditioner:
// TACSBlockJacobiPc (the reference builds TACSAdditiveSchwarz / TACSSchurPc).
// Prints one JSON line; tests/test_gpu_cpp_driver.py checks it.
#include <math.h>
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>

#include "TACSHip.h"

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  const int n = argc > 1 ? atoi(argv[1]) : 4;
  const int nn = n + 1, nnodes = nn * nn * nn, nelems = n * n * n;
  auto id = [&](int i, int j, int k) { return (i * nn + j) * nn + k; };

  TACSMaterialProperties* props = new TACSMaterialProperties(2700.0, 921.0, 70e9, 0.3, 270e6);
  TACSSolidConstitutive* stiff = new TACSSolidConstitutive(props, 1.0, 0);  // one shared DV
  TACSLinearElasticity3D* model = new TACSLinearElasticity3D(stiff, TACS_LINEAR_STRAIN);
  TACSLinearHexaBasis* basis = new TACSLinearHexaBasis();
  TACSElement3D* elem = new TACSElement3D(model, basis);

  TACSAssembler* assembler = new TACSAssembler(MPI_COMM_WORLD, 3, nnodes, nelems);
  assembler->incref();
  int* ptr = new int[nelems + 1];
  int* conn = new int[8 * nelems];
  TACSElement** elems = new TACSElement*[nelems];
  int e = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++, e++) {
        ptr[e] = 8 * e;
        int* c = &conn[8 * e];  // local order x fastest (TACSHexaBasis.cpp:137)
        c[0] = id(i, j, k), c[1] = id(i + 1, j, k), c[2] = id(i, j + 1, k);
        c[3] = id(i + 1, j + 1, k), c[4] = id(i, j, k + 1), c[5] = id(i + 1, j, k + 1);
        c[6] = id(i, j + 1, k + 1), c[7] = id(i + 1, j + 1, k + 1);
        elems[e] = elem;
      }
  ptr[nelems] = 8 * nelems;
  int fail = assembler->setElementConnectivity(ptr, conn);
  fail |= assembler->setElements(elems);
  int* bcn = new int[nn * nn];
  for (int j = 0; j < nn; j++)
    for (int k = 0; k < nn; k++) bcn[j * nn + k] = id(0, j, k);
  fail |= assembler->addBCs(nn * nn, bcn);
  fail |= assembler->initialize();
  if (fail) {
    printf("{\"error\": \"set-up failed\"}\n");
    return 1;
  }

  TACSBVec* X = assembler->createNodeVec();
  X->incref();
  TacsScalar* Xa;
  X->getArray(&Xa);
  for (int i = 0; i < nn; i++)
    for (int j = 0; j < nn; j++)
      for (int k = 0; k < nn; k++) {
        TacsScalar* x = &Xa[3 * id(i, j, k)];
        x[0] = (double)i / n, x[1] = (double)j / n, x[2] = (double)k / n;
      }
  assembler->setNodes(X);

  // traction aux elements on the x = 1 face (face 1 = +x)
  TACSAuxElements* aux = new TACSAuxElements();
  TacsScalar trac[3] = {0.0, 0.0, 1e6};
  TACSElement* tr = elem->createElementTraction(1, trac);
  for (int j = 0; j < n; j++)
    for (int k = 0; k < n; k++) aux->addElement(((n - 1) * n + j) * n + k, tr);
  assembler->setAuxElements(aux);

  TACSParallelMat* mat = assembler->createMat();
  mat->incref();
  TACSBVec *res = assembler->createVec(), *u = assembler->createVec();
  TACSBVec* u2 = assembler->createVec();
  res->incref(), u->incref(), u2->incref();
  assembler->zeroVariables();
  assembler->assembleJacobian(1.0, 0.0, 0.0, res, mat);
  TACSBlockJacobiPc* pc = new TACSBlockJacobiPc(mat);
  pc->incref();
  pc->factor();

  GMRES* ksm = new GMRES(mat, pc, 30, 100);
  ksm->incref();
  ksm->setTolerances(1e-12, 1e-30);
  int flag = ksm->solve(res, u);
  int its = ksm->getIterCount();
  GMRES* fksm = new GMRES(mat, pc, 30, 100, 1);  // flexible
  fksm->incref();
  fksm->setTolerances(1e-12, 1e-30);
  int fflag = fksm->solve(res, u2);
  int fits = fksm->getIterCount();

  u->scale(-1.0), u2->scale(-1.0);
  u2->axpy(-1.0, u);
  const double diff = u2->norm() / u->norm();
  assembler->setVariables(u);

  TACSFunction* comp = new TACSCompliance(assembler);
  comp->incref();
  TacsScalar c = 0.0;
  assembler->evalFunctions(1, &comp, &c);
  TACSBVec* dfdu = assembler->createVec();
  TACSBVec* psi = assembler->createVec();
  TACSBVec* dfdx = assembler->createDesignVec();
  dfdu->incref(), psi->incref(), dfdx->incref();
  assembler->addSVSens(1.0, 0.0, 0.0, 1, &comp, &dfdu);
  assembler->applyBCs(dfdu);
  int aflag = ksm->solve(dfdu, psi);
  assembler->addDVSens(1.0, 1, &comp, &dfdx);
  assembler->addAdjointResProducts(-1.0, 1, &psi, &dfdx);
  TacsScalar* g;
  dfdx->getArray(&g);
  // the error convention: a second connectivity after initialize() is refused
  int late = assembler->setElementConnectivity(ptr, conn);

  printf("{\"n\": %d, \"compliance\": %.17g, \"flag\": %d, \"iterations\": %d, "
         "\"flexible_flag\": %d, \"flexible_iterations\": %d, \"flexible_vs_plain\": %.3e, "
         "\"adjoint_flag\": %d, \"dcdt\": %.17g, \"late_connectivity_rejected\": %d, "
         "\"assembler_refs\": %d}\n",
         n, c, flag, its, fflag, fits, diff, aflag, g[0], late != 0, assembler->refcount());
  fflush(stdout);

  TACSObject* objs[] = {dfdx, psi, dfdu, comp, fksm, ksm, pc, u2, u, res, mat, X};
  for (TACSObject* o : objs) o->decref();
  assembler->decref();
  delete[] ptr;
  delete[] conn;
  delete[] elems;
  delete[] bcn;
  MPI_Finalize();
  return 0;
}
