// Object-level C ABI over the classes of include/TACSHip.h (declared in
// include/tacs_hip.h).  Every function mirrors one method bound by the
// reference's Cython layer (tacs/cpp_headers/TACS.pxd, tacs/*.pyx).
#include <stdio.h>

#include <vector>

#include "../../include/TACSHip.h"

int tacs_set_error(const char* fmt, ...);
const char* tacs_get_error();
void tacs_set_stream_impl(hipStream_t s);
hipStream_t tacs_stream();

// Every entry point starts with the error (and the halt of launches that a
// failed allocation, copy or launch sets) cleared: tacs_last_error()
// describes the most recent failing call.
#define OBJ(T, h) (dynamic_cast<T*>((TACSObject*)(h)))
#define REQUIRE(T, v, h)                                           \
  tacs_clear_error();                                              \
  T* v = OBJ(T, h);                                                \
  if (!v) return tacs_set_error("%s: expected a %s handle", __func__, #T)
#define REQUIRE_NULL(T, v, h)                                      \
  tacs_clear_error();                                              \
  T* v = OBJ(T, h);                                                \
  if (!v) {                                                        \
    tacs_set_error("%s: expected a %s handle", __func__, #T);      \
    return NULL;                                                   \
  }
// device work needs the arrays initialize() uploads
#define REQUIRE_INIT(a)                                                              \
  if (!(a)->initialized)                                                             \
  return tacs_set_error("%s: the assembler is not initialized", __func__)
#define REQUIRE_INIT_NULL(a)                                                         \
  if (!(a)->initialized) {                                                           \
    tacs_set_error("%s: the assembler is not initialized", __func__);                \
    return NULL;                                                                     \
  }

static int last_error_code() {
  const char* e = tacs_get_error();
  return (e && e[0]) ? 1 : 0;
}
void tacs_clear_error();
int tacs_halted();
static void clear_error() { tacs_clear_error(); }
// An object whose construction failed (a device allocation or upload) is
// destroyed and NULL returned, so no half-built object reaches a caller.
template <class T>
static tacs_obj built(T* o) {
  if (o && (tacs_halted() || last_error_code())) {
    o->incref();
    o->decref();
    return NULL;
  }
  return (TACSObject*)o;
}

extern "C" {

const char* tacs_last_error(void) { return tacs_get_error(); }
void tacs_incref(tacs_obj o) {
  if (o) ((TACSObject*)o)->incref();
}
void tacs_decref(tacs_obj o) {
  if (o) ((TACSObject*)o)->decref();
}
int tacs_refcount(tacs_obj o) { return o ? ((TACSObject*)o)->refcount() : 0; }
void tacs_set_stream(hipStream_t s) { tacs_set_stream_impl(s); }
hipStream_t tacs_get_stream(void) { return tacs_stream(); }
int tacs_device_synchronize(void) { return (int)hipStreamSynchronize(tacs_stream()); }

tacs_obj tacs_material_create(double rho, double cp, double E, double nu, double ys, double alpha,
                              double kappa) {
  return (TACSObject*)new TACSMaterialProperties(rho, cp, E, nu, ys, alpha, kappa);
}
tacs_obj tacs_solid_constitutive_create(tacs_obj props, double t, int tNum, double tlb,
                                        double tub) {
  REQUIRE_NULL(TACSMaterialProperties, p, props);
  return (TACSObject*)new TACSSolidConstitutive(p, t, tNum, tlb, tub);
}
tacs_obj tacs_linear_hexa_basis_create(void) { return (TACSObject*)new TACSLinearHexaBasis(); }
tacs_obj tacs_quadratic_hexa_basis_create(void) {
  return (TACSObject*)new TACSQuadraticHexaBasis();
}
tacs_obj tacs_linear_elasticity3d_create(tacs_obj stiff, int strain_type) {
  REQUIRE_NULL(TACSSolidConstitutive, s, stiff);
  return (TACSObject*)new TACSLinearElasticity3D(s, (ElementStrainType)strain_type);
}
tacs_obj tacs_element3d_create(tacs_obj model, tacs_obj basis) {
  REQUIRE_NULL(TACSElementModel, m, model);
  REQUIRE_NULL(TACSElementBasis, b, basis);
  return (TACSObject*)new TACSElement3D(m, b);
}
tacs_obj tacs_traction3d_create(int vpn, int face, tacs_obj basis, const double* trac, int coord) {
  REQUIRE_NULL(TACSElementBasis, b, basis);
  if (face < 0 || face > 5) {
    tacs_set_error("tacs_traction3d_create: face %d out of range", face);
    return NULL;
  }
  return (TACSObject*)new TACSTraction3D(vpn, face, b, trac, coord);
}
tacs_obj tacs_pressure3d_create(int vpn, int face, tacs_obj basis, double p) {
  REQUIRE_NULL(TACSElementBasis, b, basis);
  if (face < 0 || face > 5) {
    tacs_set_error("tacs_pressure3d_create: face %d out of range", face);
    return NULL;
  }
  return (TACSObject*)new TACSPressure3D(vpn, face, b, p);
}
tacs_obj tacs_iso_shell_constitutive_create(tacs_obj props, double t, int tNum, double tlb,
                                            double tub, double tOffset, double kcorr) {
  REQUIRE_NULL(TACSMaterialProperties, p, props);
  return (TACSObject*)new TACSIsoShellConstitutive(p, t, tNum, tlb, tub, tOffset, kcorr);
}
tacs_obj tacs_shell_natural_transform_create(void) {
  return (TACSObject*)new TACSShellNaturalTransform();
}
tacs_obj tacs_shell_ref_axis_transform_create(const double* axis) {
  if (!axis) {
    tacs_set_error("%s: axis is NULL", __func__);
    return NULL;
  }
  return (TACSObject*)new TACSShellRefAxisTransform(axis);
}
tacs_obj tacs_quad4_shell_create(tacs_obj transform, tacs_obj con) {
  REQUIRE_NULL(TACSShellTransform, tr, transform);
  REQUIRE_NULL(TACSShellConstitutive, c, con);
  return (TACSObject*)new TACSQuad4Shell(tr, c);
}
tacs_obj tacs_element_create_pressure(tacs_obj element, int face, double p) {
  REQUIRE_NULL(TACSQuad4Shell, e, element);
  return (TACSObject*)e->createElementPressure(face, p);
}
tacs_obj tacs_aux_elements_create(int size) { return (TACSObject*)new TACSAuxElements(size); }
int tacs_aux_elements_add(tacs_obj aux, int elem, tacs_obj element) {
  REQUIRE(TACSAuxElements, a, aux);
  REQUIRE(TACSElement, e, element);
  a->addElement(elem, e);
  return 0;
}

tacs_obj tacs_assembler_create(int vpn, int nowned, int nelems, int ndep) {
  clear_error();
  TACSAssembler* a = new TACSAssembler(TACS_COMM_SELF, vpn, nowned, nelems, ndep);
  return (TACSObject*)a;
}
int tacs_assembler_set_element_connectivity(tacs_obj h, const int* ptr, const int* conn) {
  REQUIRE(TACSAssembler, a, h);
  return a->setElementConnectivity(ptr, conn);
}
int tacs_assembler_set_elements(tacs_obj h, tacs_obj* elems) {
  REQUIRE(TACSAssembler, a, h);
  TACSElement** e = new TACSElement*[a->numElements];
  for (int i = 0; i < a->numElements; i++) {
    e[i] = OBJ(TACSElement, elems[i]);
    if (!e[i]) {
      delete[] e;
      return tacs_set_error("setElements: entry %d is not an element", i);
    }
  }
  int r = a->setElements(e);
  delete[] e;
  return r;
}
int tacs_assembler_add_bcs(tacs_obj h, int nnodes, const int* nodes, int nbcs, const int* vars,
                           const double* vals) {
  REQUIRE(TACSAssembler, a, h);
  return a->addBCs(nnodes, nodes, nbcs, vars, vals);
}
int tacs_assembler_initialize(tacs_obj h) {
  REQUIRE(TACSAssembler, a, h);
  clear_error();
  return a->initialize();
}
int tacs_assembler_set_element_precision(tacs_obj h, int prec) {
  REQUIRE(TACSAssembler, a, h);
  return a->setElementPrecision(prec);
}
int tacs_assembler_set_jacobian_kernel(tacs_obj h, int kernel) {
  REQUIRE(TACSAssembler, a, h);
  return a->setJacobianKernel(kernel);
}
int tacs_assembler_jacobian_kernel(tacs_obj h) {
  REQUIRE(TACSAssembler, a, h);
  return a->jacobianKernelInUse();
}
int tacs_assembler_autotune(tacs_obj h, tacs_obj mat, int reps, double* ms, int maxv,
                            int* nchoices) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE(TACSParallelMat, A, mat);
  const int nc = a->autotune(A, reps, ms, maxv);
  if (nchoices) *nchoices = last_error_code() ? 0 : nc;
  return last_error_code();
}
int tacs_assembler_get_jacobian_variant(tacs_obj h, int* variant) {
  REQUIRE(TACSAssembler, a, h);
  if (!variant) return tacs_set_error("%s: variant pointer", __func__);
  *variant = a->jacobianVariant();
  return 0;
}
int tacs_assembler_set_jacobian_variant(tacs_obj h, int variant) {
  REQUIRE(TACSAssembler, a, h);
  const int nv = tacs_hip_solid_jacobian_nvariants(a->order);
  if (variant < -1 || variant > nv)
    return tacs_set_error("%s: variant -1..%d (%d: element blocks)", __func__, nv, nv);
  a->jacVar = variant;
  return 0;
}
int tacs_assembler_set_aux_elements(tacs_obj h, tacs_obj aux) {
  REQUIRE(TACSAssembler, a, h);
  a->setAuxElements(aux ? OBJ(TACSAuxElements, aux) : NULL);
  return 0;
}
int tacs_assembler_get_info(tacs_obj h, int* vpn, int* nn, int* ne, int* ndv) {
  REQUIRE(TACSAssembler, a, h);
  if (vpn) *vpn = a->varsPerNode;
  if (nn) *nn = a->numNodes;
  if (ne) *ne = a->numElements;
  if (ndv) *ndv = a->numDesignVars;
  return 0;
}
tacs_obj tacs_assembler_create_vec(tacs_obj h) {
  REQUIRE_NULL(TACSAssembler, a, h);
  return built(a->createVec());
}
tacs_obj tacs_assembler_create_node_vec(tacs_obj h) {
  REQUIRE_NULL(TACSAssembler, a, h);
  return built(a->createNodeVec());
}
tacs_obj tacs_assembler_create_design_vec(tacs_obj h) {
  REQUIRE_NULL(TACSAssembler, a, h);
  return built(a->createDesignVec());
}
tacs_obj tacs_assembler_create_mat(tacs_obj h) {
  REQUIRE_NULL(TACSAssembler, a, h);
  return built(a->createMat());
}

#define ASM_VEC_OP(name, call)                      \
  int name(tacs_obj h, tacs_obj v) {                \
    REQUIRE(TACSAssembler, a, h);                   \
    REQUIRE_INIT(a);                                \
    REQUIRE(TACSBVec, x, v);                        \
    clear_error();                                  \
    a->call(x);                                     \
    return last_error_code();                       \
  }
ASM_VEC_OP(tacs_assembler_set_nodes, setNodes)
ASM_VEC_OP(tacs_assembler_get_nodes, getNodes)
ASM_VEC_OP(tacs_assembler_set_variables, setVariables)
ASM_VEC_OP(tacs_assembler_get_variables, getVariables)
ASM_VEC_OP(tacs_assembler_set_design_vars, setDesignVars)
ASM_VEC_OP(tacs_assembler_get_design_vars, getDesignVars)

int tacs_assembler_zero_variables(tacs_obj h) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  a->zeroVariables();
  return last_error_code();
}
int tacs_assembler_apply_bcs_vec(tacs_obj h, tacs_obj v) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE(TACSBVec, x, v);
  a->applyBCs((TACSVec*)x);
  return last_error_code();
}
int tacs_assembler_apply_bcs_mat(tacs_obj h, tacs_obj m) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  REQUIRE(TACSParallelMat, x, m);
  a->applyBCs((TACSMat*)x);
  return last_error_code();
}
int tacs_assembler_assemble_res(tacs_obj h, tacs_obj res, double lambda, int bcs) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  REQUIRE(TACSBVec, r, res);
  clear_error();
  a->assembleRes(r, lambda, bcs != 0);
  return last_error_code();
}
int tacs_assembler_assemble_jacobian(tacs_obj h, double alpha, double beta, double gamma,
                                     tacs_obj res, tacs_obj mat, int mat_or, double lambda,
                                     int bcs) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  REQUIRE(TACSParallelMat, m, mat);
  TACSBVec* r = res ? OBJ(TACSBVec, res) : NULL;
  clear_error();
  a->assembleJacobian(alpha, beta, gamma, r, m, (MatrixOrientation)mat_or, lambda, bcs != 0);
  return last_error_code();
}
int tacs_assembler_eval_functions(tacs_obj h, int n, tacs_obj* funcs, double* vals) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  TACSFunction* f[64];
  if (n > 64) return tacs_set_error("evalFunctions: at most 64 functions");
  for (int i = 0; i < n; i++) f[i] = OBJ(TACSFunction, funcs[i]);
  clear_error();
  a->evalFunctions(n, f, vals);
  return last_error_code();
}
int tacs_assembler_add_dv_sens(tacs_obj h, double coef, int n, tacs_obj* funcs, tacs_obj* dfdx) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  TACSFunction* f[64];
  TACSBVec* d[64];
  if (n > 64) return tacs_set_error("addDVSens: at most 64 functions");
  for (int i = 0; i < n; i++) {
    f[i] = OBJ(TACSFunction, funcs[i]);
    d[i] = OBJ(TACSBVec, dfdx[i]);
  }
  clear_error();
  a->addDVSens(coef, n, f, d);
  return last_error_code();
}
int tacs_assembler_add_sv_sens(tacs_obj h, double alpha, double beta, double gamma, int n,
                               tacs_obj* funcs, tacs_obj* dfdu) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  TACSFunction* f[64];
  TACSBVec* d[64];
  if (n > 64) return tacs_set_error("addSVSens: at most 64 functions");
  for (int i = 0; i < n; i++) {
    f[i] = OBJ(TACSFunction, funcs[i]);
    d[i] = OBJ(TACSBVec, dfdu[i]);
  }
  clear_error();
  a->addSVSens(alpha, beta, gamma, n, f, d);
  return last_error_code();
}
int tacs_assembler_add_adjoint_res_products(tacs_obj h, double scale, int n, tacs_obj* adjoint,
                                            tacs_obj* dfdx, double lambda) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  TACSBVec* ad[64];
  TACSBVec* d[64];
  if (n > 64) return tacs_set_error("addAdjointResProducts: at most 64 adjoints");
  for (int i = 0; i < n; i++) {
    ad[i] = OBJ(TACSBVec, adjoint[i]);
    d[i] = OBJ(TACSBVec, dfdx[i]);
  }
  clear_error();
  a->addAdjointResProducts(scale, n, ad, d, lambda);
  return last_error_code();
}

int tacs_assembler_add_xpt_sens(tacs_obj h, double coef, int n, tacs_obj* funcs,
                                tacs_obj* dfdX) {
  clear_error();
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  std::vector<TACSFunction*> f(n);
  std::vector<TACSBVec*> v(n);
  for (int k = 0; k < n; k++) {
    f[k] = OBJ(TACSFunction, funcs[k]);
    v[k] = OBJ(TACSBVec, dfdX[k]);
    if (!f[k] || !v[k]) return tacs_set_error("addXptSens: bad handle %d", k);
  }
  a->addXptSens(coef, n, f.data(), v.data());
  return last_error_code();
}
int tacs_assembler_add_adjoint_res_xpt_sens_products(tacs_obj h, double scale, int n,
                                                     tacs_obj* adjoint, tacs_obj* dfdX,
                                                     double lambda) {
  clear_error();
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  std::vector<TACSBVec*> psi(n), v(n);
  for (int k = 0; k < n; k++) {
    psi[k] = OBJ(TACSBVec, adjoint[k]);
    v[k] = OBJ(TACSBVec, dfdX[k]);
    if (!psi[k] || !v[k]) return tacs_set_error("addAdjointResXptSensProducts: bad handle %d", k);
  }
  a->addAdjointResXptSensProducts(scale, n, psi.data(), v.data(), lambda);
  return last_error_code();
}

int tacs_vec_get_size(tacs_obj v, int* size, int* bsize) {
  REQUIRE(TACSBVec, x, v);
  if (size) *size = (int)x->n;
  if (bsize) *bsize = x->bsize;
  return 0;
}
double* tacs_vec_get_array(tacs_obj v) {
  REQUIRE_NULL(TACSBVec, x, v);
  double* h = NULL;
  x->getArray(&h);
  return h;
}
double* tacs_vec_device_array(tacs_obj v) {
  REQUIRE_NULL(TACSBVec, x, v);
  return x->devWrite();
}
int tacs_vec_copy_from_device(tacs_obj v, const double* src, size_t n) {
  REQUIRE(TACSBVec, x, v);
  if (n != x->n) return tacs_set_error("copy_from_device: size %zu != %zu", n, x->n);
  hipError_t e = hipMemcpyAsync(x->devWrite(), src, n * sizeof(double),
                                hipMemcpyDeviceToDevice, tacs_stream());
  return e == hipSuccess ? 0 : tacs_set_error("copy_from_device: %s", hipGetErrorString(e));
}
int tacs_vec_copy_to_device(tacs_obj v, double* dst, size_t n) {
  REQUIRE(TACSBVec, x, v);
  if (n != x->n) return tacs_set_error("copy_to_device: size %zu != %zu", n, x->n);
  hipError_t e = hipMemcpyAsync(dst, x->devRead(), n * sizeof(double), hipMemcpyDeviceToDevice,
                                tacs_stream());
  return e == hipSuccess ? 0 : tacs_set_error("copy_to_device: %s", hipGetErrorString(e));
}
double tacs_vec_norm(tacs_obj v) {
  clear_error();
  TACSBVec* x = OBJ(TACSBVec, v);
  return x ? x->norm() : 0.0;
}
double tacs_vec_dot(tacs_obj v, tacs_obj w) {
  clear_error();
  TACSBVec* x = OBJ(TACSBVec, v);
  TACSBVec* y = OBJ(TACSBVec, w);
  return (x && y) ? x->dot(y) : 0.0;
}
int tacs_vec_axpy(tacs_obj v, double alpha, tacs_obj w) {
  REQUIRE(TACSBVec, y, v);
  REQUIRE(TACSBVec, x, w);
  y->axpy(alpha, x);
  return last_error_code();
}
int tacs_vec_axpby(tacs_obj v, double alpha, double beta, tacs_obj w) {
  REQUIRE(TACSBVec, y, v);
  REQUIRE(TACSBVec, x, w);
  y->axpby(alpha, beta, x);
  return last_error_code();
}
int tacs_vec_scale(tacs_obj v, double alpha) {
  REQUIRE(TACSBVec, y, v);
  y->scale(alpha);
  return last_error_code();
}
int tacs_vec_copy_values(tacs_obj v, tacs_obj w) {
  REQUIRE(TACSBVec, y, v);
  REQUIRE(TACSBVec, x, w);
  if (x->n != y->n) return tacs_set_error("copyValues: size mismatch");
  y->copyValues(x);
  return last_error_code();
}
int tacs_vec_zero_entries(tacs_obj v) {
  REQUIRE(TACSBVec, y, v);
  y->zeroEntries();
  return last_error_code();
}
int tacs_vec_set(tacs_obj v, double val) {
  REQUIRE(TACSBVec, y, v);
  y->set(val);
  return last_error_code();
}
int tacs_vec_set_rand(tacs_obj v, double lo, double hi) {
  REQUIRE(TACSBVec, y, v);
  y->setRand(lo, hi);
  return last_error_code();
}

int tacs_mat_mult(tacs_obj m, tacs_obj x, tacs_obj y) {
  REQUIRE(TACSMat, A, m);
  REQUIRE(TACSBVec, xv, x);
  REQUIRE(TACSBVec, yv, y);
  clear_error();
  A->mult(xv, yv);
  return last_error_code();
}
int tacs_assembler_assemble_mat_type(tacs_obj h, int matType, tacs_obj mat, int matOr,
                                     double loadScale, int applyBCs) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  REQUIRE(TACSMat, A, mat);
  clear_error();
  a->assembleMatType((ElementMatrixType)matType, A, (MatrixOrientation)matOr, loadScale,
                     applyBCs != 0);
  return last_error_code();
}
int tacs_assembler_assemble_mat_combo(tacs_obj h, int nmats, const int* matTypes,
                                      const double* scale, tacs_obj mat, int matOr,
                                      double loadScale, int applyBCs) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  REQUIRE(TACSMat, A, mat);
  if (nmats < 1 || nmats > 8 || !matTypes || !scale)
    return tacs_set_error("assembleMatCombo: 1 to 8 matrix types with their scales");
  ElementMatrixType t[8];
  TacsScalar sc[8];
  for (int i = 0; i < nmats; i++) {
    t[i] = (ElementMatrixType)matTypes[i];
    sc[i] = scale[i];
  }
  clear_error();
  a->assembleMatCombo(t, sc, nmats, A, (MatrixOrientation)matOr, loadScale, applyBCs != 0);
  return last_error_code();
}
tacs_obj tacs_matfree_create(tacs_obj assembler) {
  REQUIRE_NULL(TACSAssembler, a, assembler);
  REQUIRE_INIT_NULL(a);
  return built(new TACSMatrixFreeMat(a));
}
int tacs_matfree_assemble_data(tacs_obj m, int matType, double alpha, double beta,
                               double gamma) {
  REQUIRE(TACSMatrixFreeMat, F, m);
  clear_error();
  F->assembleMatrixFreeData((ElementMatrixType)matType, alpha, beta, gamma);
  return last_error_code();
}
int tacs_matfree_mult_transpose(tacs_obj m, tacs_obj x, tacs_obj y) {
  REQUIRE(TACSMatrixFreeMat, F, m);
  REQUIRE(TACSBVec, xv, x);
  REQUIRE(TACSBVec, yv, y);
  clear_error();
  F->multTranspose(xv, yv);
  return last_error_code();
}
int tacs_matfree_mult_device(tacs_obj m, const double* x, double* y) {
  REQUIRE(TACSMatrixFreeMat, F, m);
  clear_error();
  F->mult64(x, y);
  return last_error_code();
}
int tacs_assembler_get_matrix_free_data_size(tacs_obj h, int matType, int* data_size,
                                             int* temp_size) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  clear_error();
  a->getMatrixFreeDataSize((ElementMatrixType)matType, data_size, temp_size);
  return last_error_code();
}
int tacs_assembler_assemble_matrix_free_data(tacs_obj h, int matType, double alpha, double beta,
                                             double gamma, double* data) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  clear_error();
  a->assembleMatrixFreeData((ElementMatrixType)matType, alpha, beta, gamma, data);
  return last_error_code();
}
int tacs_assembler_add_matrix_free_vec_product(tacs_obj h, int matType, const double* data,
                                               double* temp, tacs_obj x, tacs_obj y,
                                               int matOr, double lambda, int applyBCs) {
  REQUIRE(TACSAssembler, a, h);
  REQUIRE_INIT(a);
  REQUIRE(TACSBVec, xv, x);
  REQUIRE(TACSBVec, yv, y);
  clear_error();
  a->addMatrixFreeVecProduct((ElementMatrixType)matType, data, temp, xv, yv,
                             (MatrixOrientation)matOr, lambda, applyBCs != 0);
  return last_error_code();
}
int tacs_mat_mult_prec(tacs_obj m, int prec_vals, int prec_vecs, const void* x, void* y) {
  REQUIRE(TACSParallelMat, A, m);
  clear_error();
  if (prec_vals == 64 && prec_vecs == 64)
    A->mult64((const double*)x, (double*)y);
  else if (prec_vals == 32 && prec_vecs == 64)
    A->multMixed((const double*)x, (double*)y);
  else if (prec_vals == 32 && prec_vecs == 32)
    A->mult32((const float*)x, (float*)y);
  else
    return tacs_set_error("tacs_mat_mult_prec: precisions %d/%d", prec_vals, prec_vecs);
  return last_error_code();
}
int tacs_mat_prepare_products(tacs_obj m, int prec) {
  REQUIRE(TACSParallelMat, A, m);
  if (prec != 64 && prec != 32) return tacs_set_error("%s: prec 64 or 32", __func__);
  if ((prec == 64 ? A->sellVer64 : A->sellVer32) != A->version) A->sellUpdate(prec);
  return last_error_code();
}
int tacs_mat_cheb_step(tacs_obj m, tacs_obj pc, int prec_vals, const double* z, double* zout,
                       const double* rhs, double* d, double c1, double c2) {
  REQUIRE(TACSParallelMat, A, m);
  REQUIRE(TACSBlockJacobiPc, p, pc);
  if ((prec_vals != 64 && prec_vals != 32 && prec_vals != 16) || z == zout || p->mat != A)
    return tacs_set_error("%s: prec 64, 32 or 16, zout != z, the matrix's own preconditioner",
                          __func__);
  A->chebStep(prec_vals, prec_vals == 64 ? (const void*)p->d_Dinv : (const void*)p->d_Dinv32, z,
              zout, rhs, d, c1, c2);
  return last_error_code();
}
int tacs_mat_cheb_step3(tacs_obj m, tacs_obj pc, int prec_vals, const double* z, double* zprev,
                        const double* rhs, double c1, double c2) {
  REQUIRE(TACSParallelMat, A, m);
  REQUIRE(TACSBlockJacobiPc, p, pc);
  if ((prec_vals != 64 && prec_vals != 32 && prec_vals != 16) || !z || !zprev || z == zprev ||
      p->mat != A)
    return tacs_set_error("%s: prec 64, 32 or 16, zprev != z, the matrix's own preconditioner",
                          __func__);
  A->chebStep3(prec_vals, prec_vals == 64 ? (const void*)p->d_Dinv : (const void*)p->d_Dinv32, z,
               zprev, rhs, c1, c2);
  return last_error_code();
}
int tacs_mat_autotune(tacs_obj m, int reps, double* ms, int maxv, int* nvariants) {
  REQUIRE(TACSParallelMat, A, m);
  const int nv = A->autotune(reps, ms, maxv);
  if (nvariants) *nvariants = last_error_code() ? 0 : nv;
  return last_error_code();
}
int tacs_mat_get_sell_variant(tacs_obj m, int kind, int* variant) {
  REQUIRE(TACSParallelMat, A, m);
  if (kind < 0 || kind > 2 || !variant) return tacs_set_error("%s: kind 0, 1 or 2", __func__);
  *variant = A->sellVar[kind] >= 0 ? A->sellVar[kind]
                                   : tacs_tune_variant(TACS_TUNE_SELL_SPMV, 10 * A->bsize + kind,
                                                       A->nrows);
  return 0;
}
int tacs_mat_set_sell_variant(tacs_obj m, int kind, int variant) {
  REQUIRE(TACSParallelMat, A, m);
  if (kind < 0 || kind > 2 || variant < -1 || variant >= tacs_hip_sell_spmv_nvariants(A->bsize))
    return tacs_set_error("%s: kind 0..2, variant -1..%d", __func__,
                          tacs_hip_sell_spmv_nvariants(A->bsize) - 1);
  A->sellVar[kind] = variant;
  return 0;
}
int tacs_mat_autotune_spmm(tacs_obj m, int nrhs, int prec, int reps, double* ms, int maxv,
                           int* nvariants) {
  REQUIRE(TACSParallelMat, A, m);
  const int nv = A->autotuneSpMM(nrhs, prec, reps, ms, maxv);
  if (nvariants) *nvariants = last_error_code() ? 0 : nv;
  return last_error_code();
}
int tacs_mat_zero_entries(tacs_obj m) {
  REQUIRE(TACSParallelMat, A, m);
  A->zeroEntries();
  return last_error_code();
}
int tacs_mat_get_info(tacs_obj m, int* bsize, int* nrows, long* nnzb) {
  REQUIRE(TACSParallelMat, A, m);
  if (bsize) *bsize = A->bsize;
  if (nrows) *nrows = A->nrows;
  if (nnzb) *nnzb = A->nnzb;
  return 0;
}
int tacs_mat_get_arrays(tacs_obj m, const int** rowp, const int** cols, const double** Av) {
  REQUIRE(TACSParallelMat, A, m);
  double* h = NULL;
  A->getArrays(NULL, NULL, rowp, cols, &h);
  if (Av) *Av = h;
  return last_error_code();
}
int tacs_mat_device_arrays(tacs_obj m, const int** rowp, const int** cols, const double** Av) {
  REQUIRE(TACSParallelMat, A, m);
  if (rowp) *rowp = A->d_rowp;
  if (cols) *cols = A->d_cols;
  if (Av) *Av = A->d_A;
  return 0;
}

tacs_obj tacs_ksm_print_stdout_create(const char* descript, int rank, int freq) {
  return (TACSObject*)new KSMPrintStdout(descript, rank, freq);
}
tacs_obj tacs_ksm_print_file_create(const char* filename, const char* descript, int rank,
                                    int freq) {
  clear_error();
  KSMPrintFile* f = new KSMPrintFile(filename, descript, rank, freq);
  if (rank == 0 && !f->fp) {
    delete f;
    return NULL;
  }
  return (TACSObject*)f;
}
int tacs_ksm_set_monitor(tacs_obj ksm, tacs_obj monitor) {
  KSMPrint* m = NULL;
  if (monitor) {
    m = OBJ(KSMPrint, monitor);
    if (!m) return tacs_set_error("%s: expected a KSMPrint handle", __func__);
  }
  if (TACSKsm* k = OBJ(TACSKsm, ksm)) {
    k->setMonitor(m);
    return 0;
  }
  REQUIRE(TACSBatchedGMRES, b, ksm);
  b->setMonitor(m);
  return 0;
}

tacs_obj tacs_block_jacobi_pc_create(tacs_obj mat) {
  REQUIRE_NULL(TACSParallelMat, A, mat);
  return built(new TACSBlockJacobiPc(A));
}
int tacs_pc_factor(tacs_obj pc) {
  REQUIRE(TACSPc, p, pc);
  clear_error();
  p->factor();
  return last_error_code();
}
int tacs_pc_set_chebyshev(tacs_obj pc, int degree, double ratio, int values) {
  REQUIRE(TACSBlockJacobiPc, p, pc);
  return p->setChebyshev(degree, ratio, values);
}
int tacs_pc_get_chebyshev(tacs_obj pc, int* degree, double* ratio, double* lmax) {
  REQUIRE(TACSBlockJacobiPc, p, pc);
  if (degree) *degree = p->chebDegree;
  if (ratio) *ratio = p->chebRatio;
  if (lmax) *lmax = p->chebMax;
  return 0;
}
int tacs_pc_apply(tacs_obj pc, tacs_obj x, tacs_obj y) {
  REQUIRE(TACSPc, p, pc);
  REQUIRE(TACSBVec, xv, x);
  REQUIRE(TACSBVec, yv, y);
  p->applyFactor(xv, yv);
  return last_error_code();
}
tacs_obj tacs_gmres_create(tacs_obj mat, tacs_obj pc, int m, int nrestart, int flexible) {
  clear_error();
  // the operator: an assembled TACSParallelMat or a TACSMatrixFreeMat
  TACSParallelMat* A = OBJ(TACSParallelMat, mat);
  TACSMatrixFreeMat* F = A ? NULL : OBJ(TACSMatrixFreeMat, mat);
  if (!A && !F) {
    tacs_set_error("GMRES: expected a TACSParallelMat or TACSMatrixFreeMat operator");
    return NULL;
  }
  TACSPc* p = pc ? OBJ(TACSPc, pc) : NULL;
  if (m < 1 || m > 512) {
    tacs_set_error("GMRES: subspace size must be in [1, 512]");
    return NULL;
  }
  TACSBlockJacobiPc* bj = p ? dynamic_cast<TACSBlockJacobiPc*>(p) : NULL;
  const size_t n = A ? (size_t)A->nrows * A->bsize : F->length();
  if (bj && (size_t)bj->mat->nrows * bj->mat->bsize != n) {
    tacs_set_error("GMRES: preconditioner size does not match the operator");
    return NULL;
  }
  return built(new GMRES(A ? (TACSMat*)A : (TACSMat*)F, p, m, nrestart, flexible));
}
int tacs_gmres_set_cycle_graph(tacs_obj ksm, int on) {
  REQUIRE(GMRES, g, ksm);
  g->setCycleGraph(on);
  return 0;
}
int tacs_gmres_set_ortho_type(tacs_obj ksm, int ortho) {
  REQUIRE(GMRES, g, ksm);
  g->setOrthoType((enum OrthoType)ortho);
  return 0;
}
tacs_obj tacs_mixed_gmres_create(tacs_obj mat, tacs_obj pc, int m, int nrestart) {
  REQUIRE_NULL(TACSParallelMat, A, mat);
  REQUIRE_NULL(TACSBlockJacobiPc, p, pc);
  if (m < 1 || m > 512) {
    tacs_set_error("mixed GMRES: subspace size must be in [1, 512]");
    return NULL;
  }
  return built(new TACSMixedGMRES(A, p, m, nrestart));
}
int tacs_mixed_gmres_set_deflation(tacs_obj ksm, int k) {
  clear_error();
  REQUIRE(TACSMixedGMRES, g, ksm);
  return g->setDeflation(k);
}
tacs_obj tacs_batched_gmres_create(tacs_obj mat, tacs_obj pc, int m, int nrestart, int nrhs,
                                   int prec) {
  clear_error();
  REQUIRE_NULL(TACSParallelMat, A, mat);
  REQUIRE_NULL(TACSBlockJacobiPc, p, pc);
  if (A->bsize != 3 || m < 1 || m > 512 || nrhs < 1 || nrhs > 64 || (prec != 32 && prec != 64) ||
      nrestart < 0) {
    tacs_set_error("batched GMRES: need 3x3 blocks, 1 <= m <= 512, 1 <= nrhs <= 64, "
                   "prec 32 or 64, nrestart >= 0");
    return NULL;
  }
  return built(new TACSBatchedGMRES(A, p, m, nrestart, nrhs, prec));
}
tacs_obj tacs_batched_gmres_create_segments(tacs_obj mat, tacs_obj pc, int m, int nrestart,
                                            int nseg, const int* node_off, int prec) {
  clear_error();
  REQUIRE_NULL(TACSParallelMat, A, mat);
  REQUIRE_NULL(TACSBlockJacobiPc, p, pc);
  if (A->bsize != 3 || m < 1 || m > 512 || nseg < 1 || nseg > 64 || !node_off ||
      (prec != 32 && prec != 64) || nrestart < 0) {
    tacs_set_error("segmented batched GMRES: need 3x3 blocks, 1 <= m <= 512, 1 <= nseg <= 64, "
                   "node offsets, prec 32 or 64, nrestart >= 0");
    return NULL;
  }
  return built(new TACSBatchedGMRES(A, p, m, nrestart, nseg, node_off, prec));
}
int tacs_batched_gmres_solve_segments(tacs_obj ksm, tacs_obj b, tacs_obj x, int zero_guess) {
  clear_error();
  TACSBatchedGMRES* k = OBJ(TACSBatchedGMRES, ksm);
  TACSBVec* bv = OBJ(TACSBVec, b);
  TACSBVec* xv = OBJ(TACSBVec, x);
  if (!k || !bv || !xv) {
    tacs_set_error("%s: expected a TACSBatchedGMRES handle and two TACSBVec", __func__);
    return -1;
  }
  return k->solveSegments(bv, xv, zero_guess);
}
int tacs_batched_gmres_set_tolerances(tacs_obj ksm, double rtol, double atol) {
  REQUIRE(TACSBatchedGMRES, k, ksm);
  k->setTolerances(rtol, atol);
  return 0;
}
int tacs_batched_gmres_set_deflation(tacs_obj ksm, int k) {
  clear_error();
  REQUIRE(TACSBatchedGMRES, g, ksm);
  return g->setDeflation(k);
}
int tacs_batched_gmres_set_delayed_reorth(tacs_obj ksm, int delayed) {
  REQUIRE(TACSBatchedGMRES, g, ksm);
  g->delayed = delayed ? 1 : 0;
  return 0;
}
int tacs_batched_gmres_set_recycle_space(tacs_obj ksm, tacs_obj src) {
  clear_error();
  TACSBatchedGMRES* k = OBJ(TACSBatchedGMRES, ksm);
  TACSMixedGMRES* g = OBJ(TACSMixedGMRES, src);
  if (!k || !g) {
    tacs_set_error("%s: expected a batched GMRES and a mixed GMRES handle", __func__);
    return -1;
  }
  return k->setRecycleSpace(g);
}
int tacs_batched_gmres_solve(tacs_obj ksm, int nrhs, tacs_obj* b, tacs_obj* x, int zero_guess) {
  clear_error();
  TACSBatchedGMRES* k = OBJ(TACSBatchedGMRES, ksm);
  if (!k) {
    tacs_set_error("%s: expected a TACSBatchedGMRES handle", __func__);
    return -1;
  }
  if (nrhs < 1 || nrhs > k->R) {
    tacs_set_error("%s: nrhs = %d outside [1, %d]", __func__, nrhs, k->R);
    return -1;
  }
  std::vector<TACSBVec*> bv(nrhs), xv(nrhs);
  for (int j = 0; j < nrhs; j++) {
    bv[j] = OBJ(TACSBVec, b[j]);
    xv[j] = OBJ(TACSBVec, x[j]);
    if (!bv[j] || !xv[j]) {
      tacs_set_error("%s: vector %d is not a TACSBVec", __func__, j);
      return -1;
    }
  }
  return k->solve(nrhs, bv.data(), xv.data(), zero_guess);
}
int tacs_batched_gmres_solve_block(tacs_obj ksm, int nrhs, const double* B, double* X,
                                   int zero_guess) {
  clear_error();
  TACSBatchedGMRES* k = OBJ(TACSBatchedGMRES, ksm);
  if (!k) {
    tacs_set_error("%s: expected a TACSBatchedGMRES handle", __func__);
    return -1;
  }
  return k->solveBlock(nrhs, B, X, zero_guess);
}
int tacs_batched_gmres_get_stats(tacs_obj ksm, int nrhs, int* iters, int* flags, double* r0,
                                 double* rn) {
  REQUIRE(TACSBatchedGMRES, k, ksm);
  for (int j = 0; j < nrhs && j < (int)k->iters.size(); j++) {
    if (iters) iters[j] = k->iters[j];
    if (flags) flags[j] = k->flags[j];
    if (r0) r0[j] = k->r0[j];
    if (rn) rn[j] = k->rn[j];
  }
  return k->cycles;
}
int tacs_ksm_set_tolerances(tacs_obj ksm, double rtol, double atol) {
  REQUIRE(TACSKsm, k, ksm);
  k->setTolerances(rtol, atol);
  return 0;
}
int tacs_ksm_solve(tacs_obj ksm, tacs_obj b, tacs_obj x, int zero_guess) {
  clear_error();
  TACSKsm* k = OBJ(TACSKsm, ksm);
  TACSBVec* bv = OBJ(TACSBVec, b);
  TACSBVec* xv = OBJ(TACSBVec, x);
  if (!k || !bv || !xv) {
    tacs_set_error("tacs_ksm_solve: bad handle");
    return -1;
  }
  return k->solve(bv, xv, zero_guess);
}
int tacs_ksm_get_iter_count(tacs_obj ksm) {
  TACSKsm* k = OBJ(TACSKsm, ksm);
  return k ? k->getIterCount() : -1;
}
double tacs_ksm_get_res_norm(tacs_obj ksm) {
  TACSKsm* k = OBJ(TACSKsm, ksm);
  return k ? k->getResidualNorm() : -1.0;
}
int tacs_mixed_gmres_get_outer_count(tacs_obj ksm) {
  TACSMixedGMRES* k = OBJ(TACSMixedGMRES, ksm);
  return k ? k->outerCount : -1;
}

tacs_obj tacs_compliance_create(tacs_obj assembler) {
  REQUIRE_NULL(TACSAssembler, a, assembler);
  return (TACSObject*)new TACSCompliance(a);
}
tacs_obj tacs_ksfailure_create(tacs_obj assembler, double w, double alpha, double sf) {
  REQUIRE_NULL(TACSAssembler, a, assembler);
  return (TACSObject*)new TACSKSFailure(a, w, alpha, sf);
}
double tacs_function_get_value(tacs_obj f) {
  TACSFunction* fn = OBJ(TACSFunction, f);
  return fn ? fn->getFunctionValue() : 0.0;
}
double tacs_ksfailure_get_max_failure(tacs_obj f) {
  TACSKSFailure* k = OBJ(TACSKSFailure, f);
  return k ? k->maxFail : 0.0;
}

}  // extern "C"
