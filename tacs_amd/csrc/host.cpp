// Host side of the MI355X TACS path: the C++ classes of include/TACSHip.h.
// They keep the reference's public semantics (src/TACSAssembler.h,
// src/bpmat/TACSBVec.h, src/bpmat/TACSParallelMat.h, src/bpmat/KSM.h,
// src/functions/*.h) while all data lives on the GPU and every loop runs as a
// kernel of tacs_amd/csrc/*.hip.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "../../include/TACSHip.h"

// ---------------------------------------------------------------------------
// Library state: stream and last error
// ---------------------------------------------------------------------------
static hipStream_t g_stream = 0;
static thread_local std::string g_err;
// Set when an allocation, copy or kernel launch failed: every later copy and
// launch of this thread is skipped (and the solvers leave their loops) until
// the error is cleared, so a failed hipMalloc never turns into a kernel
// working through a NULL buffer.  The C ABI clears it at the start of each
// call; C++ callers clear it with tacs_clear_error().
static thread_local int g_halt = 0;

// Host threads for independent per-system work (at most 16, the share of a
// GPU's host cores; TACS_HOST_THREADS overrides).
static int host_threads() {
  static int n = [] {
    const char* e = getenv("TACS_HOST_THREADS");
    int v = e ? atoi(e) : (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(v, 16));
  }();
  return n;
}

int tacs_set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  fprintf(stderr, "TACS error: %s\n", buf);
  return 1;
}
const char* tacs_get_error() { return g_err.c_str(); }
void tacs_clear_error() {
  g_err.clear();
  g_halt = 0;
}
int tacs_halted() { return g_halt; }
void tacs_set_stream_impl(hipStream_t s) { g_stream = s; }
hipStream_t tacs_stream() { return g_stream; }

#define HCHECK(expr)                                                                     \
  do {                                                                                   \
    if (g_halt) break;                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) {                                                              \
      tacs_set_error("%s failed at %s:%d: %s", #expr, __FILE__, __LINE__,                \
                     hipGetErrorString(_e));                                             \
      g_halt = 1;                                                                        \
    }                                                                                    \
  } while (0)
#define KCHECK(expr)                                                                     \
  do {                                                                                   \
    if (g_halt) break;                                                                   \
    int _e = (expr);                                                                     \
    if (_e != 0) {                                                                       \
      tacs_set_error("%s failed at %s:%d: code %d", #expr, __FILE__, __LINE__, _e);      \
      g_halt = 1;                                                                        \
    }                                                                                    \
  } while (0)

// Device allocation; NULL (with the error set and later launches halted)
// when it fails.
template <typename T>
static T* dalloc(size_t n) {
  T* p = NULL;
  if (n == 0) n = 1;
  HCHECK(hipMalloc((void**)&p, n * sizeof(T)));
  if (g_halt) return NULL;
  return p;
}
// Frees whatever it is given, halted or not.
static void dfree(void* p) {
  if (p) (void)hipFree(p);
}
static void hfree(void* p) {
  if (p) (void)hipHostFree(p);
}
template <typename T>
static void h2d(T* d, const T* h, size_t n) {
  if (n) HCHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, g_stream));
}
template <typename T>
static void d2h(T* h, const T* d, size_t n) {
  if (n) HCHECK(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, g_stream));
  HCHECK(hipStreamSynchronize(g_stream));
}
static double read_scalar(const double* d) {
  double v = 0.0;
  d2h(&v, d, 1);
  return v;
}

// ---------------------------------------------------------------------------
// TACSObject
// ---------------------------------------------------------------------------
TACSObject::TACSObject() : ref_count(0) {}
TACSObject::~TACSObject() {}
void TACSObject::incref() { ref_count++; }
void TACSObject::decref() {
  ref_count--;
  if (ref_count == 0) delete this;
}
const char* TACSObject::getObjectName() { return "TACSObject"; }

// ---------------------------------------------------------------------------
// Materials and elements (descriptors; the element family is compiled into
// the kernels)
// ---------------------------------------------------------------------------
TACSMaterialProperties::TACSMaterialProperties(TacsScalar _rho, TacsScalar _cp, TacsScalar _E,
                                               TacsScalar _nu, TacsScalar _ys, TacsScalar _alpha,
                                               TacsScalar _kappa)
    : rho(_rho), specific_heat(_cp), E(_E), nu(_nu), ys(_ys), alpha(_alpha), kappa(_kappa) {
  G = 0.5 * E / (1.0 + nu);  // TACSMaterialProperties.cpp:31
}

// Isotropic C[21] (TACSMaterialProperties::evalTangentStiffness3D :270)
void TACSMaterialProperties::evalTangentStiffness3D(TacsScalar C[]) {
  TacsScalar D = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  for (int i = 0; i < 21; i++) C[i] = 0.0;
  C[0] = C[6] = C[11] = (1.0 - nu) * D;
  C[1] = C[2] = C[7] = nu * D;
  C[15] = C[18] = C[20] = G;
}

TACSSolidConstitutive::TACSSolidConstitutive(TACSMaterialProperties* p, TacsScalar _t, int _tNum,
                                             TacsScalar _tlb, TacsScalar _tub)
    : props(p), t(_t), tlb(_tlb), tub(_tub), tNum(_tNum) {
  if (props) props->incref();
}
TACSSolidConstitutive::~TACSSolidConstitutive() {
  if (props) props->decref();
}

TACSLinearElasticity3D::TACSLinearElasticity3D(TACSSolidConstitutive* s, ElementStrainType st)
    : stiff(s), strain_type(st) {
  stiff->incref();
}
TACSLinearElasticity3D::~TACSLinearElasticity3D() { stiff->decref(); }

TACSElement3D::TACSElement3D(TACSElementModel* m, TACSElementBasis* b) : model(m), basis(b) {
  model->incref();
  basis->incref();
}
TACSElement3D::~TACSElement3D() {
  model->decref();
  basis->decref();
}

TACSIsoShellConstitutive::TACSIsoShellConstitutive(TACSMaterialProperties* p, TacsScalar _t,
                                                   int _tNum, TacsScalar _tlb, TacsScalar _tub,
                                                   TacsScalar _tOffset, TacsScalar _kcorr)
    : props(p), t(_t), tlb(_tlb), tub(_tub), tOffset(_tOffset), kcorr(_kcorr), tNum(_tNum) {
  if (props) props->incref();
}
TACSIsoShellConstitutive::~TACSIsoShellConstitutive() {
  if (props) props->decref();
}

TACSShellRefAxisTransform::TACSShellRefAxisTransform(const TacsScalar _axis[]) {
  const double nrm = sqrt(_axis[0] * _axis[0] + _axis[1] * _axis[1] + _axis[2] * _axis[2]);
  const double inv = nrm != 0.0 ? 1.0 / nrm : 0.0;
  for (int k = 0; k < 3; k++) axis[k] = inv * _axis[k];
}
void TACSShellRefAxisTransform::getRefAxis(TacsScalar _axis[]) {
  for (int k = 0; k < 3; k++) _axis[k] = axis[k];
}
TACSQuad4Shell::TACSQuad4Shell(TACSShellTransform* tr, TACSShellConstitutive* c)
    : transform(tr), con(c) {
  transform->incref();
  con->incref();
}
TACSQuad4Shell::~TACSQuad4Shell() {
  transform->decref();
  con->decref();
}
TACSElement* TACSQuad4Shell::createElementPressure(int, TacsScalar p) {
  return new TACSShellPressure(p);
}

TACSTraction3D::TACSTraction3D(int vpn, int face, TACSElementBasis* b, const TacsScalar t[],
                               int coord)
    : varsPerNode(vpn), faceIndex(face), coordComponent(coord), basis(b) {
  basis->incref();
  int n = coord ? vpn : 3 * vpn;
  for (int i = 0; i < 24; i++) trac[i] = (i < n) ? t[i] : 0.0;
}
TACSTraction3D::~TACSTraction3D() { basis->decref(); }
TACSPressure3D::TACSPressure3D(int vpn, int face, TACSElementBasis* b, TacsScalar _p)
    : varsPerNode(vpn), faceIndex(face), basis(b), p(_p) {
  basis->incref();
}
TACSPressure3D::~TACSPressure3D() { basis->decref(); }
TACSElement* TACSElement3D::createElementTraction(int face, const TacsScalar t[]) {
  return new TACSTraction3D(getVarsPerNode(), face, basis, t, 1);
}
TACSElement* TACSElement3D::createElementPressure(int face, TacsScalar p) {
  return new TACSPressure3D(getVarsPerNode(), face, basis, p);
}

TACSAuxElements::TACSAuxElements(int size) : n(0), cap(size > 0 ? size : 16), version(0) {
  nums = new int[cap];
  elems = new TACSElement*[cap];
}
TACSAuxElements::~TACSAuxElements() {
  for (int i = 0; i < n; i++) elems[i]->decref();
  delete[] nums;
  delete[] elems;
}
void TACSAuxElements::addElement(int num, TACSElement* e) {
  if (n >= cap) {
    int nc = 2 * cap;
    int* nn = new int[nc];
    TACSElement** ne = new TACSElement*[nc];
    memcpy(nn, nums, n * sizeof(int));
    memcpy(ne, elems, n * sizeof(TACSElement*));
    delete[] nums;
    delete[] elems;
    nums = nn;
    elems = ne;
    cap = nc;
  }
  e->incref();
  nums[n] = num;
  elems[n] = e;
  n++;
  version++;
}
void TACSAuxElements::clear() {
  for (int i = 0; i < n; i++) elems[i]->decref();
  n = 0;
  version++;
}

// ---------------------------------------------------------------------------
// TACSBcMap (KSM.cpp:76)
// ---------------------------------------------------------------------------
TACSBcMap::TACSBcMap(int _bsize, int num) : bsize(_bsize), nbcs(0) {
  max_size = num > 0 ? num : 32;
  nodes = new int[max_size];
  vars = new int[max_size];
  values = new TacsScalar[bsize * max_size];
}
TACSBcMap::~TACSBcMap() {
  delete[] nodes;
  delete[] vars;
  delete[] values;
}
void TACSBcMap::addBC(int node, int nvals, const int* bc_vars, const TacsScalar* bc_vals) {
  if (nbcs >= max_size) {
    int ms = 2 * max_size;
    int* tn = new int[ms];
    int* tv = new int[ms];
    TacsScalar* tval = new TacsScalar[bsize * ms];
    memcpy(tn, nodes, nbcs * sizeof(int));
    memcpy(tv, vars, nbcs * sizeof(int));
    memcpy(tval, values, bsize * nbcs * sizeof(TacsScalar));
    delete[] nodes;
    delete[] vars;
    delete[] values;
    nodes = tn;
    vars = tv;
    values = tval;
    max_size = ms;
  }
  nodes[nbcs] = node;
  vars[nbcs] = 0;
  for (int k = 0; k < bsize; k++) values[bsize * nbcs + k] = 0.0;
  if (bc_vars && bc_vals) {
    for (int i = 0; i < nvals; i++) {
      if (bc_vars[i] >= 0 && bc_vars[i] < bsize) {
        vars[nbcs] |= (1 << bc_vars[i]);
        values[bsize * nbcs + bc_vars[i]] = bc_vals[i];
      }
    }
  } else if (bc_vars) {
    for (int i = 0; i < nvals; i++)
      if (bc_vars[i] >= 0 && bc_vars[i] < bsize) vars[nbcs] |= (1 << bc_vars[i]);
  } else {
    for (int i = 0; i < nvals && i < bsize; i++) vars[nbcs] |= (1 << i);
  }
  nbcs++;
}
int TACSBcMap::getBCs(const int** n, const int** v, TacsScalar** vals) {
  if (n) *n = nodes;
  if (v) *v = vars;
  if (vals) *vals = values;
  return nbcs;
}

// ---------------------------------------------------------------------------
// TACSBVec: device-resident block vector with a lazily synchronised host copy
// ---------------------------------------------------------------------------
enum { VEC_DEVICE = 0, VEC_BOTH = 1, VEC_HOST = 2 };

static void* g_vec_scratch = NULL;  // reduction scratch shared by vector ops
static double* g_vec_scalar = NULL;
static void* vec_scratch() {
  if (!g_vec_scratch || !g_vec_scalar) {
    size_t b = tacs_hip_red_scratch_bytes();
    void* rs = dalloc<char>(b);
    double* sc = dalloc<double>(8);
    HCHECK(hipMemset(rs, 0, b));
    if (g_halt) {  // both or neither
      dfree(rs);
      dfree(sc);
      return NULL;
    }
    g_vec_scratch = rs;
    g_vec_scalar = sc;
  }
  return g_vec_scratch;
}

TACSBVec::TACSBVec(int _nnodes, int _bsize) : nnodes(_nnodes), bsize(_bsize) {
  n = (size_t)nnodes * bsize;
  d_x = dalloc<double>(n);
  HCHECK(hipMemsetAsync(d_x, 0, (n ? n : 1) * sizeof(double), g_stream));
  h_x = NULL;
  state = VEC_DEVICE;
}
TACSBVec::~TACSBVec() {
  dfree(d_x);
  hfree(h_x);
}
void TACSBVec::getSize(int* size) { *size = (int)n; }
int TACSBVec::getArray(TacsScalar** array) {
  if (!h_x) {
    HCHECK(hipHostMalloc((void**)&h_x, (n ? n : 1) * sizeof(double), 0));
    if (state == VEC_BOTH) state = VEC_DEVICE;
  }
  if (state == VEC_DEVICE) d2h(h_x, d_x, n);
  state = VEC_HOST;
  if (array) *array = h_x;
  return (int)n;
}
// A host mirror handed out by getArray stays writable (the reference's
// getArray returns the vector's own array), so while the device copy has not
// been written since, every device read uploads the mirror: writes through a
// held array reach the device however many reads came between them.
const double* TACSBVec::devRead() {
  if (state == VEC_HOST || state == VEC_BOTH) {
    h2d(d_x, h_x, n);
    state = VEC_BOTH;
  }
  return d_x;
}
// in-place device updates (axpy, scale) read the values too
double* TACSBVec::devWrite() {
  if (state == VEC_HOST || state == VEC_BOTH) h2d(d_x, h_x, n);
  state = VEC_DEVICE;
  return d_x;
}
TacsScalar TACSBVec::norm() {
  void* rs = vec_scratch();
  const double* x = devRead();
  KCHECK(tacs_hip_dot(64, n, NULL, NULL, (void*)x, x, rs, g_vec_scalar, 1, NULL, g_stream));
  return read_scalar(g_vec_scalar);
}
void TACSBVec::scale(TacsScalar alpha) {
  KCHECK(tacs_hip_scale(64, n, devWrite(), NULL, alpha, 0, g_stream));
}
TacsScalar TACSBVec::dot(TACSVec* tx) {
  TACSBVec* x = dynamic_cast<TACSBVec*>(tx);
  void* rs = vec_scratch();
  const double* a = devRead();
  const double* b = x->devRead();
  KCHECK(tacs_hip_dot(64, n, NULL, NULL, (void*)a, b, rs, g_vec_scalar, 0, NULL, g_stream));
  return read_scalar(g_vec_scalar);
}
void TACSBVec::axpy(TacsScalar alpha, TACSVec* tx) {
  TACSBVec* x = dynamic_cast<TACSBVec*>(tx);
  const double* xd = x->devRead();
  KCHECK(tacs_hip_axpby(64, 64, n, alpha, xd, 1.0, devWrite(), g_stream));
}
void TACSBVec::axpby(TacsScalar alpha, TacsScalar beta, TACSVec* tx) {
  TACSBVec* x = dynamic_cast<TACSBVec*>(tx);
  const double* xd = x->devRead();
  KCHECK(tacs_hip_axpby(64, 64, n, alpha, xd, beta, devWrite(), g_stream));
}
void TACSBVec::copyValues(TACSVec* tx) {
  TACSBVec* x = dynamic_cast<TACSBVec*>(tx);
  const double* xd = x->devRead();
  if (n) HCHECK(hipMemcpyAsync(devWrite(), xd, n * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
}
void TACSBVec::zeroEntries() { KCHECK(tacs_hip_fill(64, n, devWrite(), 0.0, g_stream)); }
void TACSBVec::set(TacsScalar v) { KCHECK(tacs_hip_fill(64, n, devWrite(), v, g_stream)); }
void TACSBVec::setRand(double lower, double upper) {
  double* h;
  getArray(&h);
  for (size_t i = 0; i < n; i++) h[i] = lower + (upper - lower) * (rand() / (double)RAND_MAX);
}
void TACSBVec::applyBCs(TACSBcMap* map, TACSVec* tvec, TacsScalar lambda) {
  // TACSBVec::applyBCs src/bpmat/TACSBVec.cpp:546 (entries applied in order)
  const int *nodes, *vars;
  TacsScalar* vals;
  int nbc = map->getBCs(&nodes, &vars, &vals);
  if (nbc == 0) return;
  TACSBVec* u = tvec ? dynamic_cast<TACSBVec*>(tvec) : NULL;
  double* h;
  double* hu = NULL;
  if (u) {
    u->getArray(&hu);
  }
  getArray(&h);
  for (int i = 0; i < nbc; i++) {
    for (int k = 0; k < bsize; k++) {
      if (vars[i] & (1 << k)) {
        size_t idx = (size_t)bsize * nodes[i] + k;
        h[idx] = hu ? hu[idx] - lambda * vals[bsize * i + k] : 0.0;
      }
    }
  }
}
void TACSBVec::setBCs(TACSBcMap* map, TacsScalar lambda) {
  const int *nodes, *vars;
  TacsScalar* vals;
  int nbc = map->getBCs(&nodes, &vars, &vals);
  if (nbc == 0) return;
  double* h;
  getArray(&h);
  for (int i = 0; i < nbc; i++)
    for (int k = 0; k < bsize; k++)
      if (vars[i] & (1 << k)) h[(size_t)bsize * nodes[i] + k] = lambda * vals[bsize * i + k];
}

// ---------------------------------------------------------------------------
// TACSParallelMat (one rank = the [Aloc] block-CSR matrix)
// ---------------------------------------------------------------------------
TACSParallelMat::TACSParallelMat(int _bsize, int _nrows, const int* rowp, const int* cols)
    : bsize(_bsize), nrows(_nrows), version(0) {
  nnzb = rowp[nrows];
  h_rowp = new int[nrows + 1];
  h_cols = new int[nnzb > 0 ? nnzb : 1];
  memcpy(h_rowp, rowp, (nrows + 1) * sizeof(int));
  memcpy(h_cols, cols, nnzb * sizeof(int));
  d_rowp = dalloc<int>(nrows + 1);
  d_cols = dalloc<int>(nnzb);
  h2d(d_rowp, h_rowp, nrows + 1);
  h2d(d_cols, h_cols, nnzb);
  size_t nv = (size_t)nnzb * bsize * bsize;
  d_A = dalloc<double>(nv);
  HCHECK(hipMemsetAsync(d_A, 0, (nv ? nv : 1) * sizeof(double), g_stream));
  d_A32 = NULL;
  A32_current = 0;
  h_A = NULL;
  d_sw = NULL;
  d_soff = NULL;
  d_scols = NULL;
  d_s64 = NULL;
  d_s32 = NULL;
  d_s16 = NULL;
  sellVer64 = sellVer32 = sellVer16 = -1;
  nslots = 0;
  sellVar[0] = sellVar[1] = sellVar[2] = -1;
}
TACSParallelMat::~TACSParallelMat() {
  dfree(d_rowp);
  dfree(d_cols);
  dfree(d_A);
  dfree(d_A32);
  void* sell[] = {d_sw, d_soff, d_scols, d_s64, d_s32, d_s16};
  for (void* q : sell) dfree(q);
  delete[] h_rowp;
  delete[] h_cols;
  hfree(h_A);
}
TACSVec* TACSParallelMat::createVec() { return new TACSBVec(nrows, bsize); }
void TACSParallelMat::mult(TACSVec* tx, TACSVec* ty) {
  TACSBVec* x = dynamic_cast<TACSBVec*>(tx);
  TACSBVec* y = dynamic_cast<TACSBVec*>(ty);
  const double* xd = x->devRead();
  mult64(xd, y->devWrite());
}
// Sliced ELL pattern on first use (host widths and offsets from the row
// pointer, columns with the first fill); values refreshed when the matrix
// version moved.
void TACSParallelMat::sellUpdate(int prec) {
  const int ns = (nrows + 63) / 64;
  const size_t b2 = (size_t)bsize * bsize;
  int* cs_fill = NULL;
  if (!d_sw) {
    std::vector<int> w(ns > 0 ? ns : 1);
    std::vector<long> off((size_t)ns + 1);
    nslots = tacs_hip_sell_slots(nrows, h_rowp, w.data(), off.data());
    int* sw = dalloc<int>(ns);
    long* soff = dalloc<long>((size_t)ns + 1);
    int* scols = dalloc<int>(nslots);
    h2d(sw, w.data(), ns);
    h2d(soff, off.data(), (size_t)ns + 1);
    HCHECK(hipStreamSynchronize(g_stream));
    if (g_halt) {  // the pattern arrays exist together or not at all
      dfree(sw);
      dfree(soff);
      dfree(scols);
      return;
    }
    d_sw = sw;
    d_soff = soff;
    d_scols = scols;
    cs_fill = d_scols;
  }
  // the requested copy, and the other precision's copy too when it exists
  // and is stale (the recycling solver reads both after every assembly):
  // one pass over the BCSR values writes both
  if (prec == 64 && !d_s64) d_s64 = dalloc<double>(b2 * nslots);
  if (prec == 32 && !d_s32) d_s32 = dalloc<float>(b2 * nslots);
  if (prec == 16 && !d_s16) d_s16 = dalloc<unsigned short>(b2 * nslots);
  if (g_halt) return;
  double* v64 = (d_s64 && sellVer64 != version) ? d_s64 : NULL;
  float* v32 = (d_s32 && sellVer32 != version) ? d_s32 : NULL;
  unsigned short* v16 = (d_s16 && sellVer16 != version) ? d_s16 : NULL;
  KCHECK(tacs_hip_sell_fill3(bsize, nrows, d_rowp, d_cols, d_sw, d_soff, d_A, v64, v32, v16,
                             cs_fill, g_stream));
  if (v64) sellVer64 = version;
  if (v32) sellVer32 = version;
  if (v16) sellVer16 = version;
}
// The products run over the sliced ELL copies (3x3 and 6x6 blocks), in the
// matrix's launch variant (autotune) or the tuning table's.
static int sell_kind_variant(const TACSParallelMat* m, int kind) {
  return m->sellVar[kind] >= 0 ? m->sellVar[kind]
                               : tacs_tune_variant(TACS_TUNE_SELL_SPMV, 10 * m->bsize + kind,
                                                   m->nrows);
}
void TACSParallelMat::mult64(const double* x, double* y) {
  if (sellVer64 != version) sellUpdate(64);
  KCHECK(tacs_hip_sell_spmv_variant(sell_kind_variant(this, 0), bsize, 64, 64, nrows, d_sw, d_soff,
                                    d_scols, d_s64, x, y, g_stream));
}
void TACSParallelMat::multMixed(const double* x, double* y) {
  if (sellVer32 != version) sellUpdate(32);
  KCHECK(tacs_hip_sell_spmv_variant(sell_kind_variant(this, 1), bsize, 32, 64, nrows, d_sw, d_soff,
                                    d_scols, d_s32, x, y, g_stream));
}
void TACSParallelMat::mult32(const float* x, float* y) {
  if (sellVer32 != version) sellUpdate(32);
  KCHECK(tacs_hip_sell_spmv_variant(sell_kind_variant(this, 2), bsize, 32, 32, nrows, d_sw, d_soff,
                                    d_scols, d_s32, x, y, g_stream));
}
// One Chebyshev step fused into the product with the values of precision pr
// (zout = z + d after d = c1 d + c2 D^{-1} (rhs - A z)).
void TACSParallelMat::chebStep(int pr, const void* Dinv, const double* z, double* zout,
                               const double* rhs, double* d, double c1, double c2) {
  const int kind = pr == 64 ? 0 : 1;
  if (pr == 64 && sellVer64 != version) sellUpdate(64);
  if (pr == 32 && sellVer32 != version) sellUpdate(32);
  if (pr == 16 && sellVer16 != version) sellUpdate(16);
  const void* vs = pr == 64 ? (const void*)d_s64 : pr == 32 ? (const void*)d_s32 : (const void*)d_s16;
  KCHECK(tacs_hip_sell_cheb_step(sell_kind_variant(this, kind), bsize, pr, nrows, d_sw, d_soff,
                                 d_scols, vs, z, zout, rhs, Dinv, d, c1, c2, g_stream));
}
// The same step in three-term form: zprev <- (1 + c1) z - c1 zprev
// + c2 D^{-1} (rhs - A z).
void TACSParallelMat::chebStep3(int pr, const void* Dinv, const double* z, double* zprev,
                                const double* rhs, double c1, double c2) {
  const int kind = pr == 64 ? 0 : 1;
  if (pr == 64 && sellVer64 != version) sellUpdate(64);
  if (pr == 32 && sellVer32 != version) sellUpdate(32);
  if (pr == 16 && sellVer16 != version) sellUpdate(16);
  const void* vs = pr == 64 ? (const void*)d_s64 : pr == 32 ? (const void*)d_s32 : (const void*)d_s16;
  KCHECK(tacs_hip_sell_cheb_step3(sell_kind_variant(this, kind), bsize, pr, nrows, d_sw, d_soff,
                                  d_scols, vs, z, zprev, rhs, Dinv, c1, c2, g_stream));
}
// Average time of reps launches of fn on the library stream (one untimed
// launch first), in ms; negative after a failure.
template <typename F>
static double time_launches(int reps, F fn) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1.0;
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return -1.0;
  }
  int err = fn();
  HCHECK(hipEventRecord(e0, g_stream));
  for (int r = 0; r < reps && !err; r++) err = fn();
  HCHECK(hipEventRecord(e1, g_stream));
  HCHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HCHECK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (err) {
    tacs_set_error("autotune: launch failed (%d)", err);
    return -1.0;
  }
  return g_halt ? -1.0 : (double)ms / reps;
}
int TACSParallelMat::autotune(int reps, double* ms, int maxv) {
  const int nv = tacs_hip_sell_spmv_nvariants(bsize);
  if (reps < 1 || !ms || maxv < nv) return tacs_set_error("autotune: reps >= 1 and room for %d variants", nv);
  if (sellVer64 != version) sellUpdate(64);
  if (!d_s32) sellUpdate(32);
  else if (sellVer32 != version) sellUpdate(32);
  const size_t n = (size_t)nrows * bsize;
  double* x = dalloc<double>(n);
  double* y = dalloc<double>(n);
  float* x32 = dalloc<float>(n);
  float* y32 = dalloc<float>(n);
  if (!g_halt) {
    KCHECK(tacs_hip_fill(64, n, x, 1.0, g_stream));
    KCHECK(tacs_hip_fill(32, n, x32, 1.0, g_stream));
  }
  for (int kind = 0; kind < 3 && !g_halt; kind++) {
    int best = 0;
    for (int v = 0; v < nv && !g_halt; v++) {
      const int pv = kind == 0 ? 64 : 32, pw = kind == 2 ? 32 : 64;
      const void* vs = kind == 0 ? (const void*)d_s64 : (const void*)d_s32;
      const void* xa = kind == 2 ? (const void*)x32 : (const void*)x;
      void* ya = kind == 2 ? (void*)y32 : (void*)y;
      ms[kind * maxv + v] = time_launches(reps, [&]() {
        return tacs_hip_sell_spmv_variant(v, bsize, pv, pw, nrows, d_sw, d_soff, d_scols, vs, xa,
                                          ya, g_stream);
      });
      if (ms[kind * maxv + v] < ms[kind * maxv + best]) best = v;
    }
    if (!g_halt) sellVar[kind] = best;
  }
  dfree(x);
  dfree(y);
  dfree(x32);
  dfree(y32);
  return g_halt ? 1 : nv;
}
int TACSParallelMat::autotuneSpMM(int nrhs, int prec, int reps, double* ms, int maxv) {
  if (bsize != 3) return tacs_set_error("autotuneSpMM: 3x3 blocks only");
  if (nrhs < 1 || nrhs > 64 || (prec != 64 && prec != 32))
    return tacs_set_error("autotuneSpMM: 1 <= nrhs <= 64, prec 64 or 32");
  const int nv = tacs_hip_bsr3_spmm_nvariants(nrhs);
  if (reps < 1 || !ms || maxv < nv) return tacs_set_error("autotuneSpMM: reps >= 1 and room for %d variants", nv);
  const void* vals = prec == 64 ? (const void*)d_A : (const void*)getValues32(1);
  const size_t n = (size_t)nrows * 3 * nrhs;
  double* X = dalloc<double>(n);
  double* Y = dalloc<double>(n);
  if (!g_halt) KCHECK(tacs_hip_fill(64, n, X, 1.0, g_stream));
  int best = 0;
  for (int v = 0; v < nv && !g_halt; v++) {
    ms[v] = time_launches(reps, [&]() {
      return tacs_hip_bsr3_spmm_variant(v, prec, nrows, nrhs, d_rowp, d_cols, vals, X, Y, g_stream);
    });
    if (ms[v] < ms[best]) best = v;
  }
  dfree(X);
  dfree(Y);
  if (g_halt) return 1;
  tacs_tune_set(TACS_TUNE_SPMM, (prec == 32 ? 100 : 0) + nrhs, tacs_tune_size_class(nrows), best);
  return nv;
}
void TACSParallelMat::zeroEntries() {
  KCHECK(tacs_hip_fill(64, (size_t)nnzb * bsize * bsize, d_A, 0.0, g_stream));
  A32_current = 0;
  version++;
}
void TACSParallelMat::applyBCs(TACSBcMap* bcmap) {
  const int *nodes, *vars;
  int nbc = bcmap->getBCs(&nodes, &vars, NULL);
  if (nbc == 0) return;
  int* dn = dalloc<int>(nbc);
  int* dv = dalloc<int>(nbc);
  h2d(dn, nodes, nbc);
  h2d(dv, vars, nbc);
  KCHECK(tacs_hip_bsr_zero_rows(bsize, nbc, dn, dv, d_rowp, d_cols, d_A, 1, g_stream));
  A32_current = 0;
  version++;
  HCHECK(hipStreamSynchronize(g_stream));
  dfree(dn);
  dfree(dv);
  A32_current = 0;
}
int TACSParallelMat::getSize(int* nr, int* nc) {
  *nr = nrows * bsize;
  *nc = nrows * bsize;
  return 0;
}
void TACSParallelMat::getArrays(int* b, int* nr, const int** rowp, const int** cols,
                                TacsScalar** A) {
  size_t nv = (size_t)nnzb * bsize * bsize;
  if (!h_A) HCHECK(hipHostMalloc((void**)&h_A, (nv ? nv : 1) * sizeof(double), 0));
  d2h(h_A, d_A, nv);
  if (b) *b = bsize;
  if (nr) *nr = nrows;
  if (rowp) *rowp = h_rowp;
  if (cols) *cols = h_cols;
  if (A) *A = h_A;
}
float* TACSParallelMat::getValues32(int refresh) {
  size_t nv = (size_t)nnzb * bsize * bsize;
  if (!d_A32) {
    d_A32 = dalloc<float>(nv);
    A32_current = 0;
  }
  if (refresh && !A32_current) {
    KCHECK(tacs_hip_convert_f64_f32(nv, d_A, d_A32, g_stream));
    A32_current = 1;
  }
  return d_A32;
}

// ---------------------------------------------------------------------------
// Block-Jacobi preconditioner
// ---------------------------------------------------------------------------
TACSBlockJacobiPc::TACSBlockJacobiPc(TACSParallelMat* m) : mat(m) {
  mat->incref();
  d_diag = dalloc<int>(mat->nrows);
  KCHECK(tacs_hip_bsr_find_diag(mat->nrows, mat->d_rowp, mat->d_cols, d_diag, g_stream));
  size_t nb = (size_t)mat->nrows * mat->bsize * mat->bsize;
  d_Dinv = dalloc<double>(nb);
  d_Dinv32 = dalloc<float>(nb);
}
TACSBlockJacobiPc::~TACSBlockJacobiPc() {
  dfree(d_diag);
  dfree(d_Dinv);
  dfree(d_Dinv32);
  mat->decref();
}
void TACSBlockJacobiPc::factor() {
  int* st = dalloc<int>(1);
  HCHECK(hipMemsetAsync(st, 0, sizeof(int), g_stream));
  KCHECK(tacs_hip_bsr_diag_inverse(mat->bsize, mat->nrows, d_diag, mat->d_A, d_Dinv, d_Dinv32, st,
                                   g_stream));
  int hs = 0;
  d2h(&hs, st, 1);
  dfree(st);
  if (hs) tacs_set_error("TACSBlockJacobiPc::factor: singular or missing diagonal block (%d)", hs);
  if (chebDegree > 1 && !g_halt) chebMax = 1.05 * lanczosMax(30);
}
// Largest eigenvalue of D^{-1} A (FP32-rounded values, as the solvers apply
// the polynomial): D^{-1} A is self-adjoint in the A inner product, so
// Lanczos in that inner product (A v carried along: one product and one
// block-Jacobi application per step) gives Ritz values from below that reach
// the top of the spectrum in a few tens of steps (Golub & Van Loan, Matrix
// Computations, 4th ed., sec. 10.1); the largest eigenvalue of the
// tridiagonal matrix comes from Sturm-sequence bisection.  A hashed start
// vector makes it reproducible.
double TACSBlockJacobiPc::lanczosMax(int steps) {
  const size_t n = (size_t)mat->nrows * mat->bsize;
  double* buf = dalloc<double>(6 * n + 2);
  if (g_halt) {
    dfree(buf);
    return 0.0;
  }
  double *v = buf, *u = buf + n, *vp = buf + 2 * n, *up = buf + 3 * n, *w = buf + 4 * n,
         *aw = buf + 5 * n, *dot = buf + 6 * n;
  auto dotp = [&](const double* x, const double* y) {
    KCHECK(tacs_hip_mdot_noalloc(64, n, 1, x, n, y, dot, 0, g_stream));
    double h = 0.0;
    d2h(&h, dot, 1);
    return h;
  };
  // the start vector vanishes on the Dirichlet (identity) rows: the
  // iteration then stays in the free dofs, where A is symmetric positive
  // definite (the identity rows only add the eigenvalue 1)
  KCHECK(tacs_hip_fill_random(n, v, 12345, g_stream));
  KCHECK(tacs_hip_bsr_zero_identity_rows(mat->bsize, mat->nrows, mat->d_rowp, mat->d_cols,
                                         mat->d_A, v, g_stream));
  mat->multMixed(v, u);
  double g2 = dotp(v, u);
  std::vector<double> al, be;  // T: diagonal al, off-diagonal be
  if (g2 > 0.0) {
    const double g = 1.0 / std::sqrt(g2);
    KCHECK(tacs_hip_scale(64, n, v, NULL, g, 0, g_stream));
    KCHECK(tacs_hip_scale(64, n, u, NULL, g, 0, g_stream));
    KCHECK(tacs_hip_fill(64, n, vp, 0.0, g_stream));
    KCHECK(tacs_hip_fill(64, n, up, 0.0, g_stream));
    double beta = 0.0;
    for (int k = 0; k < steps && !g_halt; k++) {
      KCHECK(tacs_hip_bdiag_apply(mat->bsize, 32, mat->nrows, 1, d_Dinv32, u, w, g_stream));
      const double alpha = dotp(u, w);
      mat->multMixed(w, aw);
      // w -= alpha v + beta v_prev, and A w alike
      KCHECK(tacs_hip_axpby(64, 64, n, -alpha, v, 1.0, w, g_stream));
      KCHECK(tacs_hip_axpby(64, 64, n, -beta, vp, 1.0, w, g_stream));
      KCHECK(tacs_hip_axpby(64, 64, n, -alpha, u, 1.0, aw, g_stream));
      KCHECK(tacs_hip_axpby(64, 64, n, -beta, up, 1.0, aw, g_stream));
      al.push_back(alpha);
      const double b2 = dotp(w, aw);
      if (!(b2 > 1e-30 * alpha * alpha) || k == steps - 1) break;
      beta = std::sqrt(b2);
      be.push_back(beta);
      std::swap(vp, v);
      std::swap(up, u);
      // v = w / beta, u = A w / beta (into the old v_prev, u_prev storage)
      HCHECK(hipMemcpyAsync(v, w, n * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
      HCHECK(hipMemcpyAsync(u, aw, n * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
      KCHECK(tacs_hip_scale(64, n, v, NULL, 1.0 / beta, 0, g_stream));
      KCHECK(tacs_hip_scale(64, n, u, NULL, 1.0 / beta, 0, g_stream));
    }
  }
  dfree(buf);
  const int m = (int)al.size();
  if (m == 0 || g_halt) return 0.0;
  // largest eigenvalue of T by bisection on the Sturm count
  double lo = al[0], hi = al[0];
  for (int i = 0; i < m; i++) {
    const double r = (i > 0 ? std::fabs(be[i - 1]) : 0.0) + (i < m - 1 ? std::fabs(be[i]) : 0.0);
    lo = std::min(lo, al[i] - r);
    hi = std::max(hi, al[i] + r);
  }
  auto count_below = [&](double x) {  // eigenvalues of T less than x
    int c = 0;
    double q = al[0] - x;
    if (q < 0) c++;
    for (int i = 1; i < m; i++) {
      if (q == 0.0) q = 1e-300;
      q = al[i] - x - be[i - 1] * be[i - 1] / q;
      if (q < 0) c++;
    }
    return c;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-12 * std::fabs(hi); it++) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) >= m) hi = mid;  // every eigenvalue below mid
    else lo = mid;
  }
  return hi;
}
int TACSBlockJacobiPc::setChebyshev(int degree, double ratio, int values) {
  if (degree < 1 || degree > 32 || !(ratio > 1.0) || (values != 32 && values != 16))
    return tacs_set_error("setChebyshev: 1 <= degree <= 32, ratio > 1, values 32 or 16");
  chebDegree = degree;
  chebRatio = ratio;
  chebValues = values;
  chebMax = 0.0;
  return 0;
}
void TACSBlockJacobiPc::applyFactor(TACSVec* tx, TACSVec* ty) {
  TACSBVec* x = dynamic_cast<TACSBVec*>(tx);
  TACSBVec* y = dynamic_cast<TACSBVec*>(ty);
  const double* xd = x->devRead();
  KCHECK(tacs_hip_bsr_diag_apply(mat->bsize, mat->nrows, d_Dinv, xd, y->devWrite(), g_stream));
}
void TACSBlockJacobiPc::applyFactor32(const float* x, float* y) {
  KCHECK(tacs_hip_bsr_diag_apply_f32(mat->bsize, mat->nrows, d_Dinv32, x, y, g_stream));
}

// ---------------------------------------------------------------------------
// GMRES (FP64), restating GMRES::solve src/bpmat/KSM.cpp:785-958 with the
// Arnoldi scalars kept on the device.
// ---------------------------------------------------------------------------
// scalar layout inside d_scal for subspace m
struct GmresScal {
  size_t H, res, Qc, Qs, misc, tmp, total;
  GmresScal(int m) {
    H = 0;
    res = H + (size_t)(m + 1) * m;
    Qc = res + (m + 1);
    Qs = Qc + m;
    misc = Qs + m;
    tmp = misc + 16;
    total = tmp + (m + 1);
  }
};
enum { MISC_ST = 0, MISC_NRM = 1, MISC_INV = 2, MISC_NINV = 3, MISC_ONE = 4 };

// KSMPrintStdout / KSMPrintFile (src/bpmat/KSM.cpp:239, :288)
KSMPrintStdout::KSMPrintStdout(const char* d, int r, int f)
    : descript(d ? d : ""), rank(r), freq(f < 1 ? 1 : f) {}
void KSMPrintStdout::printResidual(int iter, TacsScalar res) {
  if ((iter == 0 || iter % freq == 0) && rank == 0) {
    printf("%s[%3d]: %15.8e\n", descript.c_str(), iter, res);
    fflush(stdout);
  }
}
void KSMPrintStdout::print(const char* cstr) {
  if (rank == 0) {
    printf("%s", cstr);
    fflush(stdout);
  }
}
KSMPrintFile::KSMPrintFile(const char* filename, const char* d, int r, int f)
    : fp(NULL), descript(d ? d : ""), rank(r), freq(f < 1 ? 1 : f) {
  if (rank == 0 && filename) {
    fp = fopen(filename, "w");
    if (!fp) tacs_set_error("KSMPrintFile: cannot open %s", filename);
  }
}
KSMPrintFile::~KSMPrintFile() {
  if (fp) fclose(fp);
}
void KSMPrintFile::printResidual(int iter, TacsScalar res) {
  if ((iter == 0 || iter % freq == 0) && fp) {
    fprintf(fp, "%s[%3d]: %15.8e\n", descript.c_str(), iter, res);
    fflush(fp);
  }
}
void KSMPrintFile::print(const char* cstr) {
  if (fp) {
    fprintf(fp, "%s", cstr);
    fflush(fp);
  }
}

GMRES::GMRES(TACSMat* m, TACSPc* p, int _m, int _nrestart, int isFlexible)
    : rtol(1e-8), atol(1e-30), msub(_m), nrestart(_nrestart >= 0 ? _nrestart : 0),
      ortho(MODIFIED_GRAM_SCHMIDT), useGraph(1), gstream(NULL), gexec(NULL), gexec_ortho(-1),
      d_hist(NULL), h_hist(NULL), d_Z(NULL) {
  mat = dynamic_cast<TACSParallelMat*>(m);
  mfree = mat ? NULL : dynamic_cast<TACSMatrixFreeMat*>(m);
  pc = p ? dynamic_cast<TACSBlockJacobiPc*>(p) : NULL;
  // the C ABI checks the operator type and the preconditioner's size
  m->incref();
  if (pc) pc->incref();
  n = mat ? (size_t)mat->nrows * mat->bsize : mfree->length();
  flexible = (isFlexible && pc) ? 1 : 0;
  if (flexible) d_Z = dalloc<double>((size_t)msub * n);
  d_W = dalloc<double>((size_t)(msub + 1) * n);
  d_work = dalloc<double>(n);
  GmresScal L(msub);
  d_scal = dalloc<double>(L.total);
  HCHECK(hipMemsetAsync(d_scal, 0, L.total * sizeof(double), g_stream));
  size_t rb = tacs_hip_red_scratch_bytes();
  HCHECK(hipMalloc(&d_red, rb));
  HCHECK(hipMemsetAsync(d_red, 0, rb, g_stream));
  HCHECK(hipHostMalloc((void**)&h_pin, 64 * sizeof(double), 0));
}
GMRES::~GMRES() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (gstream) (void)hipStreamDestroy(gstream);
  dfree(d_hist);
  hfree(h_hist);
  dfree(d_Z);
  dfree(d_W);
  dfree(d_work);
  dfree(d_scal);
  dfree(d_red);
  hfree(h_pin);
  if (mat)
    mat->decref();
  else
    mfree->decref();
  if (pc) pc->decref();
}
void GMRES::opMult(const double* x, double* y) {
  if (mat)
    mat->mult64(x, y);
  else
    mfree->mult64(x, y);
}
void GMRES::setTolerances(double r, double a) {
  rtol = r;
  atol = a;
}
void GMRES::setOrthoType(enum OrthoType o) { ortho = o; }
void GMRES::setCycleGraph(int on) { useGraph = on ? 1 : 0; }

static void cgs2(int prec, size_t n, int k, const void* W, void* q, double* h, double* tmp,
                 void* red);

// One Arnoldi step of GMRES::solve (src/bpmat/KSM.cpp:850-900): w = A M^{-1}
// v_i, Gram-Schmidt against v_0..v_i, h_{i+1,i} = |w|, v_{i+1} = w / |w|, the
// Givens update of column i; |residual estimate| goes to st.
void GMRES::arnoldiStep(int i, double* st) {
  GmresScal L(msub);
  double* H = d_scal + L.H;
  double* res = d_scal + L.res;
  double* Qc = d_scal + L.Qc;
  double* Qs = d_scal + L.Qs;
  double* misc = d_scal + L.misc;
  double* tmp = d_scal + L.tmp;
  const int bs = pc ? pc->mat->bsize : 0, nr = pc ? pc->mat->nrows : 0;
  auto W = [&](int k) { return d_W + (size_t)k * n; };
  double* q = W(i + 1);
  if (pc) {
    // flexible: keep z_i = M^{-1} v_i for the update (FGMRES)
    double* z = flexible ? d_Z + (size_t)i * n : d_work;
    KCHECK(tacs_hip_bsr_diag_apply(bs, nr, pc->d_Dinv, W(i), z, g_stream));
    opMult(z, q);
  } else {
    opMult(W(i), q);
  }
  double* h = H + (size_t)i * (msub + 1);
  if (ortho == MODIFIED_GRAM_SCHMIDT) {
    // h_j = W_j . q, q -= h_j W_j for j = 0..i (ModifiedGramSchmidt
    // KSM.cpp:526); each subtraction is fused into the next dot.
    KCHECK(tacs_hip_dot(64, n, NULL, NULL, q, W(0), d_red, &h[0], 0, NULL, g_stream));
    for (int j = 1; j <= i; j++)
      KCHECK(tacs_hip_dot(64, n, W(j - 1), &h[j - 1], q, W(j), d_red, &h[j], 0, NULL, g_stream));
    KCHECK(tacs_hip_dot(64, n, W(i), &h[i], q, q, d_red, &h[i + 1], 1, misc + MISC_INV, g_stream));
  } else {
    // classical Gram-Schmidt, twice (ClassicalGramSchmidt KSM.cpp:515)
    cgs2(64, n, i + 1, d_W, q, h, tmp, d_red);
    KCHECK(tacs_hip_dot(64, n, NULL, NULL, q, q, d_red, &h[i + 1], 1, misc + MISC_INV, g_stream));
  }
  KCHECK(tacs_hip_scale(64, n, q, misc + MISC_INV, 0.0, 0, g_stream));
  KCHECK(tacs_hip_givens(i, msub, H, Qc, Qs, res, st, g_stream));
}

// Captures the msub Arnoldi steps of a cycle as one graph on the solver's own
// stream (the library stream is switched to it for the capture); step i's
// estimate goes to d_hist[i].  The graph reads the matrix's sliced ELL copy,
// the block-Jacobi inverse and the solver's own arrays, whose addresses stay
// fixed for the solver's life.  Returns 0 on success.
int GMRES::captureCycle() {
  if (!gstream) HCHECK(hipStreamCreateWithFlags(&gstream, hipStreamNonBlocking));
  if (!d_hist) d_hist = dalloc<double>(msub);
  if (!h_hist) HCHECK(hipHostMalloc((void**)&h_hist, msub * sizeof(double), 0));
  if (g_halt) return 1;
  if (gexec) {
    (void)hipGraphExecDestroy(gexec);
    gexec = NULL;
  }
  HCHECK(hipStreamSynchronize(g_stream));
  hipStream_t saved = g_stream;
  g_stream = gstream;
  hipGraph_t graph = NULL;
  hipError_t e = hipStreamBeginCapture(gstream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int i = 0; i < msub; i++) arnoldiStep(i, d_hist + i);
    e = hipStreamEndCapture(gstream, &graph);
  }
  g_stream = saved;
  if (e == hipSuccess && !g_halt) e = hipGraphInstantiate(&gexec, graph, NULL, NULL, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess || g_halt) {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = NULL;
    tacs_set_error("GMRES: capture of the Arnoldi cycle failed: %s", hipGetErrorString(e));
    g_halt = 1;
    return 1;
  }
  gexec_ortho = ortho;
  return 0;
}


// Classical Gram-Schmidt, twice, of q against the k vectors of W (stride n),
// in chunks of 32 vectors per multi-dot: h[0..k) = W^T q (both passes
// summed), q -= W h.  prec: 64 or 32 (storage of W and q).
static void cgs2(int prec, size_t n, int k, const void* W, void* q, double* h, double* tmp,
                 void* red) {
  const size_t esz = prec == 64 ? 8 : 4;
  for (int pass = 0; pass < 2; pass++) {
    double* c = pass == 0 ? h : tmp;
    for (int j0 = 0; j0 < k; j0 += 32) {
      int kk = std::min(32, k - j0);
      KCHECK(tacs_hip_mdot(prec, n, kk, (const char*)W + esz * n * j0, n, q, red, c + j0, 0,
                           g_stream));
    }
    for (int j0 = 0; j0 < k; j0 += 32) {
      int kk = std::min(32, k - j0);
      KCHECK(tacs_hip_cgs_update(prec, n, kk, (const char*)W + esz * n * j0, n, c + j0, NULL, q,
                                 g_stream));
    }
  }
  KCHECK(tacs_hip_axpby(64, 64, k, 1.0, tmp, 1.0, h, g_stream));
}

static double pin_read(double* h_pin, const double* d) {
  HCHECK(hipMemcpyAsync(h_pin, d, sizeof(double), hipMemcpyDeviceToHost, g_stream));
  HCHECK(hipStreamSynchronize(g_stream));
  return h_pin[0];
}

int GMRES::solve(TACSVec* tb, TACSVec* tx, int zero_guess) {
  TACSBVec* b = dynamic_cast<TACSBVec*>(tb);
  TACSBVec* x = dynamic_cast<TACSBVec*>(tx);
  const int bs = pc ? pc->mat->bsize : 0, nr = pc ? pc->mat->nrows : 0;
  GmresScal L(msub);
  double* H = d_scal + L.H;
  double* res = d_scal + L.res;
  double* Qc = d_scal + L.Qc;
  double* Qs = d_scal + L.Qs;
  double* misc = d_scal + L.misc;
  double* tmp = d_scal + L.tmp;
  auto W = [&](int i) { return d_W + (size_t)i * n; };
  const double* bd = b->devRead();
  double* xd = x->devWrite();
  double rhs_norm = 0.0;
  int solve_flag = 0;
  iterCount = 0;
  // the captured cycle: assembled operator only, product layout current
  // before the capture and before every launch
  const bool graph = useGraph && mat != NULL;
  if (graph) {
    if (mat->sellVer64 != mat->version) mat->sellUpdate(64);
    if ((!gexec || gexec_ortho != ortho) && captureCycle() != 0) return 0;
  }

  for (int count = 0; count < nrestart + 1; count++) {
    if (g_halt) return 0;  // a copy or launch failed: no further work
    if (zero_guess && count == 0) {
      KCHECK(tacs_hip_fill(64, n, xd, 0.0, g_stream));
      HCHECK(hipMemcpyAsync(W(0), bd, n * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
      KCHECK(tacs_hip_dot(64, n, NULL, NULL, W(0), W(0), d_red, res, 1, misc + MISC_INV, g_stream));
      KCHECK(tacs_hip_scale(64, n, W(0), misc + MISC_INV, 0.0, 0, g_stream));
    } else {
      opMult(xd, W(0));
      KCHECK(tacs_hip_axpby(64, 64, n, -1.0, bd, 1.0, W(0), g_stream));
      KCHECK(tacs_hip_dot(64, n, NULL, NULL, W(0), W(0), d_red, res, 1, misc + MISC_INV, g_stream));
      KCHECK(tacs_hip_scalar_op(misc + MISC_INV, -1.0, misc + MISC_NINV, g_stream));
      KCHECK(tacs_hip_scale(64, n, W(0), misc + MISC_NINV, 0.0, 0, g_stream));
    }
    double r0 = pin_read(h_pin, res);
    if (monitor) monitor->printResidual(0, fabs(r0));
    if (count == 0) {
      rhs_norm = r0;
      resNorm = rhs_norm;
    }
    int niters = 0;
    if (r0 < atol) {
      solve_flag = 1;
      break;
    }
    if (graph) {
      // the whole cycle as one graph launch, then the first step that stops
      HCHECK(hipGraphLaunch(gexec, g_stream));
      d2h(h_hist, d_hist, msub);
      if (g_halt) return 0;
      for (int i = 0; i < msub; i++) {
        niters++;
        resNorm = h_hist[i];
        if (monitor) monitor->printResidual(i + 1, resNorm);
        if (resNorm < atol || resNorm < rtol * rhs_norm || !std::isfinite(resNorm)) {
          solve_flag = std::isfinite(resNorm) ? 1 : 0;
          break;
        }
      }
    } else {
      for (int i = 0; i < msub; i++) {
        if (g_halt) return 0;  // a copy or launch failed: no further work
        arnoldiStep(i, misc + MISC_ST);
        niters++;
        resNorm = pin_read(h_pin, misc + MISC_ST);
        if (monitor) monitor->printResidual(i + 1, resNorm);
        if (resNorm < atol || resNorm < rtol * rhs_norm || !std::isfinite(resNorm)) {
          solve_flag = std::isfinite(resNorm) ? 1 : 0;
          break;
        }
      }
    }
    iterCount += niters;
    if (niters > 0) {
      KCHECK(tacs_hip_backsolve(niters, msub, H, res, g_stream));
      if (pc && flexible) {
        KCHECK(tacs_hip_lincomb(64, 64, n, niters, d_Z, n, res, xd, 1, g_stream));
      } else if (pc) {
        KCHECK(tacs_hip_lincomb(64, 64, n, niters, d_W, n, res, d_work, 0, g_stream));
        KCHECK(tacs_hip_bsr_diag_apply(bs, nr, pc->d_Dinv, d_work, W(0), g_stream));
        KCHECK(tacs_hip_axpby(64, 64, n, 1.0, W(0), 1.0, xd, g_stream));
      } else {
        KCHECK(tacs_hip_lincomb(64, 64, n, niters, d_W, n, res, xd, 1, g_stream));
      }
    }
    if (solve_flag || !std::isfinite(resNorm)) break;
  }
  HCHECK(hipStreamSynchronize(g_stream));
  return solve_flag;
}

// ---------------------------------------------------------------------------
// Mixed precision restarted GMRES: every cycle starts from the FP64 residual
// r = b - A x (FP64 matrix), runs up to m Arnoldi steps in FP32 (FP32 matrix,
// block-Jacobi and basis; classical Gram-Schmidt twice with FP64 dot
// products; FP64 Hessenberg and Givens), and adds the FP32 correction to the
// FP64 solution.  This is iterative refinement with one GMRES(m) cycle as the
// inner solver, so the convergence follows FP64 GMRES(m) while the
// bandwidth-bound work runs on FP32 data.
// ---------------------------------------------------------------------------
TACSMixedGMRES::TACSMixedGMRES(TACSParallelMat* m, TACSBlockJacobiPc* p, int _m, int _nrestart)
    : rtol(1e-8), atol(1e-30), msub(_m), nrestart(_nrestart), outerCount(0), ndefl(0),
      recycleDim(0), recycleVersion(-1), mat(m), pc(p), rec(NULL) {
  mat->incref();
  if (pc) pc->incref();
  mat->getValues32(0);  // allocate the FP32 copy now; assembly fills it
  size_t n = (size_t)mat->nrows * mat->bsize;
  d_W = dalloc<float>((size_t)(msub + 1) * n);
  d_work32 = dalloc<float>(n);
  d_d32 = dalloc<float>(n);
  d_r = dalloc<double>(n);
  GmresScal L(msub);
  d_scal = dalloc<double>(L.total);
  HCHECK(hipMemsetAsync(d_scal, 0, L.total * sizeof(double), g_stream));
  size_t rb = tacs_hip_red_scratch_bytes();
  HCHECK(hipMalloc(&d_red, rb));
  HCHECK(hipMemsetAsync(d_red, 0, rb, g_stream));
  HCHECK(hipHostMalloc((void**)&h_pin, (size_t)(msub + 64) * sizeof(double), 0));
}
TACSMixedGMRES::~TACSMixedGMRES() {
  if (rec) rec->decref();
  dfree(d_W);
  dfree(d_work32);
  dfree(d_d32);
  dfree(d_r);
  dfree(d_scal);
  dfree(d_red);
  hfree(h_pin);
  mat->decref();
  if (pc) pc->decref();
}
void TACSMixedGMRES::setTolerances(double r, double a) {
  rtol = r;
  atol = a;
}

int TACSMixedGMRES::setDeflation(int k) {
  if (k < 0 || k > 31 || (k > 0 && (k >= msub - 1 || msub + k > 250)))
    return tacs_set_error("setDeflation: need 0 <= k <= 31, k < m - 1 and m + k <= 250 "
                          "(k = %d, m = %d)", k, msub);
  HCHECK(hipStreamSynchronize(g_stream));
  if (rec) rec->decref();
  rec = NULL;
  ndefl = k;
  recycleDim = 0;
  recycleVersion = -1;
  if (k > 0) {
    rec = new TACSBatchedGMRES(mat, pc, msub, nrestart, 1, 32);
    rec->incref();
    if (rec->setDeflation(k) != 0) return 1;
  }
  return 0;
}

int TACSMixedGMRES::solve(TACSVec* tb, TACSVec* tx, int zero_guess) {
  if (ndefl > 0) return solveDeflated(tb, tx, zero_guess);
  TACSBVec* b = dynamic_cast<TACSBVec*>(tb);
  TACSBVec* x = dynamic_cast<TACSBVec*>(tx);
  const size_t n = (size_t)mat->nrows * mat->bsize;
  const int bs = mat->bsize, nr = mat->nrows;
  float* A32 = mat->getValues32(1);
  GmresScal L(msub);
  double* H = d_scal + L.H;
  double* res = d_scal + L.res;
  double* Qc = d_scal + L.Qc;
  double* Qs = d_scal + L.Qs;
  double* misc = d_scal + L.misc;
  double* tmp = d_scal + L.tmp;
  auto W = [&](int i) { return d_W + (size_t)i * n; };
  const double* bd = b->devRead();
  double* xd = x->devWrite();
  iterCount = 0;
  outerCount = 0;
  if (zero_guess) {
    KCHECK(tacs_hip_fill(64, n, xd, 0.0, g_stream));
    HCHECK(hipMemcpyAsync(d_r, bd, n * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
  } else {
    mat->mult64(xd, d_r);
    KCHECK(tacs_hip_axpby(64, 64, n, 1.0, bd, -1.0, d_r, g_stream));
  }
  double r0 = 0.0;
  int flag = 0;
  for (int count = 0; count <= nrestart; count++) {
    if (g_halt) return 0;  // a copy or launch failed: no further work
    // FP64 residual norm; W0 = r/|r| in FP32 (res[0] = 1)
    KCHECK(tacs_hip_dot(64, n, NULL, NULL, d_r, d_r, d_red, misc + MISC_NRM, 1, misc + MISC_INV,
                        g_stream));
    double rn = pin_read(h_pin, misc + MISC_NRM);
    if (monitor) monitor->printResidual(0, rn);
    if (count == 0) r0 = rn;
    resNorm = rn;
    if (rn <= rtol * r0 || rn < atol) {
      flag = 1;
      break;
    }
    if (!std::isfinite(rn)) break;
    KCHECK(tacs_hip_axpby(64, 32, n, 1.0 / rn, d_r, 0.0, W(0), g_stream));
    KCHECK(tacs_hip_fill(64, 1, res, 1.0, g_stream));
    const double target = std::max(rtol * r0, atol) / rn;  // in units of |r|
    int niters = 0;
    for (int i = 0; i < msub; i++) {
      if (g_halt) return 0;  // a copy or launch failed: no further work
      float* q = W(i + 1);
      KCHECK(tacs_hip_bsr_diag_apply_f32(bs, nr, pc->d_Dinv32, W(i), d_work32, g_stream));
      mat->mult32(d_work32, q);
      double* h = H + (size_t)i * (msub + 1);
      cgs2(32, n, i + 1, d_W, q, h, tmp, d_red);
      KCHECK(tacs_hip_dot(32, n, NULL, NULL, q, q, d_red, &h[i + 1], 1, misc + MISC_INV, g_stream));
      KCHECK(tacs_hip_scale(32, n, q, misc + MISC_INV, 0.0, 0, g_stream));
      KCHECK(tacs_hip_givens(i, msub, H, Qc, Qs, res, misc + MISC_ST, g_stream));
      niters++;
      double st = pin_read(h_pin, misc + MISC_ST);
      if (monitor) monitor->printResidual(i + 1, st * rn);
      if (st <= target || !std::isfinite(st)) break;
    }
    iterCount += niters;
    outerCount++;
    // x += |r| M^{-1} (W y);  r = b - A x in FP64
    KCHECK(tacs_hip_backsolve(niters, msub, H, res, g_stream));
    KCHECK(tacs_hip_lincomb(32, 32, n, niters, d_W, n, res, d_work32, 0, g_stream));
    KCHECK(tacs_hip_bsr_diag_apply_f32(bs, nr, pc->d_Dinv32, d_work32, d_d32, g_stream));
    KCHECK(tacs_hip_axpby(32, 64, n, rn, d_d32, 1.0, xd, g_stream));
    mat->mult64(xd, d_r);
    KCHECK(tacs_hip_axpby(64, 64, n, 1.0, bd, -1.0, d_r, g_stream));
  }
  HCHECK(hipStreamSynchronize(g_stream));
  return flag;
}

// ---------------------------------------------------------------------------
// Mixed precision GMRES with Krylov subspace recycling (GCRO-DR: M. L. Parks,
// E. de Sturler, G. Mackey, D. D. Johnson and S. Maiti, SIAM J. Sci. Comput.
// 28 (2006) 1651-1674).  The Arnoldi operator is the fixed FP32-rounded
// operator A32 D32^{-1} applied in FP64 arithmetic to FP64 vectors (so the
// recycled relation A32 D32^{-1} U = C stays exact across cycles, solves and
// right-hand sides); every cycle starts from the true FP64 residual
// r = b - A x.  A cycle: x += D^{-1} U C^T r, r -= C C^T r; m - k Arnoldi steps
// of (I - C C^T) A32 D32^{-1} from r (delayed classical Gram-Schmidt);
// least squares over [U D_u, V]; then the k harmonic Ritz vectors of smallest
// harmonic Ritz value of that projected problem become the next U, C.  The
// cycle is the one of the batched solver (TACSBatchedGMRES::
// solveBlockDeflated) with one right-hand side, whose least-squares problem
// and small eigenproblem run on the host.
// ---------------------------------------------------------------------------
namespace {
struct HostLS {
  int ld;                      // m + 1
  std::vector<double> Hb, Hr;  // Hessenberg (column-major, ld x m): plain / rotated
  std::vector<double> c, g;    // right-hand side: plain / rotated
  std::vector<int> rp;         // rotations: rows (rp, rp+1)
  std::vector<double> rc, rs;
  explicit HostLS(int m) : ld(m + 1), Hb((size_t)(m + 1) * m), Hr((size_t)(m + 1) * m), c(m + 1),
                           g(m + 1) {}
  void reset() {
    std::fill(Hb.begin(), Hb.end(), 0.0);
    std::fill(Hr.begin(), Hr.end(), 0.0);
    std::fill(c.begin(), c.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    rp.clear();
    rc.clear();
    rs.clear();
  }
  void rotate(double* v, int p, double cs, double sn) {
    double a = v[p], b = v[p + 1];
    v[p] = cs * a + sn * b;
    v[p + 1] = -sn * a + cs * b;
  }
  // zero Hr(row+1, col) against Hr(row, col): rotation of rows (row, row+1)
  void new_rotation(int row, int col, int ncols_from, int ncols_to) {
    double a = Hr[(size_t)col * ld + row], b = Hr[(size_t)col * ld + row + 1];
    double r = std::hypot(a, b);
    double cs = r == 0.0 ? 1.0 : a / r, sn = r == 0.0 ? 0.0 : b / r;
    for (int j = ncols_from; j < ncols_to; j++) rotate(&Hr[(size_t)j * ld], row, cs, sn);
    rotate(g.data(), row, cs, sn);
    rp.push_back(row);
    rc.push_back(cs);
    rs.push_back(sn);
  }
  // triangularise the leading dense (k+1) x k block
  void triangularize_leading(int k) {
    for (int i = 0; i < k; i++)
      for (int r = k; r > i; r--) new_rotation(r - 1, i, i, k);
  }
  // add column j (entries 0..j+1 in h); returns the residual estimate |g[j+1]|
  double add_column(int j, const double* h) {
    for (int i = 0; i < ld; i++) {
      Hb[(size_t)j * ld + i] = i <= j + 1 ? h[i] : 0.0;
      Hr[(size_t)j * ld + i] = Hb[(size_t)j * ld + i];
    }
    double* col = &Hr[(size_t)j * ld];
    for (size_t q = 0; q < rp.size(); q++) rotate(col, rp[q], rc[q], rs[q]);
    new_rotation(j, j, j, j + 1);
    return fabs(g[j + 1]);
  }
  // y = R^{-1} g for the first nc columns
  void solve(int nc, double* y) {
    for (int i = nc - 1; i >= 0; i--) {
      double s = g[i];
      for (int j = i + 1; j < nc; j++) s -= Hr[(size_t)j * ld + i] * y[j];
      y[i] = s / Hr[(size_t)i * ld + i];
    }
  }
};

// Solve A X = B in place (A row-major n x n, B row-major n x nrhs) by LU with
// partial pivoting; returns 1 when A is singular.
static int dense_solve_rowmajor(int n, const double* Ain, double* B, int nrhs) {
  std::vector<double> M(Ain, Ain + (size_t)n * n);
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (fabs(M[(size_t)i * n + k]) > fabs(M[(size_t)p * n + k])) p = i;
    if (M[(size_t)p * n + k] == 0.0) return 1;
    if (p != k) {
      for (int jj = 0; jj < n; jj++) std::swap(M[(size_t)k * n + jj], M[(size_t)p * n + jj]);
      for (int jj = 0; jj < nrhs; jj++) std::swap(B[(size_t)k * nrhs + jj], B[(size_t)p * nrhs + jj]);
    }
    for (int i = k + 1; i < n; i++) {
      double f = M[(size_t)i * n + k] / M[(size_t)k * n + k];
      if (f == 0.0) continue;
      for (int jj = k + 1; jj < n; jj++) M[(size_t)i * n + jj] -= f * M[(size_t)k * n + jj];
      for (int jj = 0; jj < nrhs; jj++) B[(size_t)i * nrhs + jj] -= f * B[(size_t)k * nrhs + jj];
    }
  }
  for (int i = n - 1; i >= 0; i--)
    for (int jj = 0; jj < nrhs; jj++) {
      double v = B[(size_t)i * nrhs + jj];
      for (int q = i + 1; q < n; q++) v -= M[(size_t)i * n + q] * B[(size_t)q * nrhs + jj];
      B[(size_t)i * nrhs + jj] = v / M[(size_t)i * n + i];
    }
  return 0;
}

// New recycle space of a GCRO-DR cycle from its harmonic Ritz vectors (Parks
// et al. 2006, sec. 2.4).  ls holds the cycle's plain Hessenberg Gbar
// (nrw = kd + jend + 1 rows [C; V_0..V_jend], nc = kd + jend columns
// [U D_u, V_0..V_jend-1]); CVtU ((kd + jend + 1) x kd, column b) holds
// [C, V_0..V_jend]^T U_b with the unscaled U.  The K harmonic Ritz vectors of
// smallest harmonic Ritz value, P_k, give Q R = Gbar P_k; C_new = [C V] Q and
// U_new = [U V] P_k R^{-1}.  Returns kk (the new dimension; 0: keep the old
// space) with Qm (nrw x kk), PRu (kd x kk, D_u folded in) and PRv (jend x kk),
// all column-major.
static int recycle_update(const HostLS& ls, int kd, int jend, int K, const double* dscale,
                          const double* CVtU, std::vector<double>& Qm, std::vector<double>& PRu,
                          std::vector<double>& PRv) {
  const int nc = kd + jend;
  const int nrw = nc + 1;
  const int ld = ls.ld;
  // Gbar^T Gbar and What^T W with What = [C, V_0..V_jend], W = [U D_u, V_0..V_jend-1]
  std::vector<double> GtG((size_t)nc * nc, 0.0), WtW((size_t)nrw * nc, 0.0);
  for (int a = 0; a < nc; a++)
    for (int bb = 0; bb < nc; bb++) {
      double acc = 0.0;
      for (int r = 0; r < nrw; r++) acc += ls.Hb[(size_t)a * ld + r] * ls.Hb[(size_t)bb * ld + r];
      GtG[(size_t)a * nc + bb] = acc;  // row-major
    }
  for (int i = 0; i < jend; i++) WtW[(size_t)(kd + i) * nrw + (kd + i)] = 1.0;  // V^T V
  for (int i = 0; i < kd; i++)
    for (int r = 0; r < nrw; r++) WtW[(size_t)i * nrw + r] = CVtU[(size_t)i * nrw + r] * dscale[i];
  std::vector<double> GtW((size_t)nc * nc, 0.0);  // row-major Gbar^T (What^T W)
  for (int a = 0; a < nc; a++)
    for (int bb = 0; bb < nc; bb++) {
      double acc = 0.0;
      for (int r = 0; r < nrw; r++) acc += ls.Hb[(size_t)a * ld + r] * WtW[(size_t)bb * nrw + r];
      GtW[(size_t)a * nc + bb] = acc;
    }
  // G = (Gbar^T What^T W)^{-1} Gbar^T Gbar: harmonic Ritz values theta, vectors z
  std::vector<double> Gm = GtG;
  if (dense_solve_rowmajor(nc, GtW.data(), Gm.data(), nc) != 0) return 0;
  std::vector<double> Z((size_t)nc * (K + 1), 0.0);
  int kout = tacs_dense_smallest_eigvecs(nc, Gm.data(), K, Z.data(), NULL, NULL);
  if (kout <= 0) return 0;
  if (kout > K) kout = K;
  // Q R = Gbar P_k (nrw x kout)
  std::vector<double> GP((size_t)nrw * kout, 0.0), R((size_t)kout * kout, 0.0);
  for (int jj = 0; jj < kout; jj++)
    for (int l = 0; l < nc; l++) {
      double pv = Z[(size_t)jj * nc + l];
      for (int r = 0; r < nrw; r++) GP[(size_t)jj * nrw + r] += ls.Hb[(size_t)l * ld + r] * pv;
    }
  int kk = 0;
  std::vector<int> keep;
  for (int jj = 0; jj < kout; jj++) {
    // modified Gram-Schmidt (twice) of GP column jj against the kept ones
    double* g = &GP[(size_t)jj * nrw];
    double nrm0 = 0.0;
    for (int r = 0; r < nrw; r++) nrm0 += g[r] * g[r];
    nrm0 = sqrt(nrm0);
    std::vector<double> rcol(kout, 0.0);
    for (int pass = 0; pass < 2; pass++)
      for (int q = 0; q < kk; q++) {
        const double* gq = &GP[(size_t)keep[q] * nrw];
        double d = 0.0;
        for (int r = 0; r < nrw; r++) d += gq[r] * g[r];
        for (int r = 0; r < nrw; r++) g[r] -= d * gq[r];
        rcol[q] += d;
      }
    double nrm = 0.0;
    for (int r = 0; r < nrw; r++) nrm += g[r] * g[r];
    nrm = sqrt(nrm);
    if (!(nrm > 1e-10 * nrm0)) continue;
    for (int r = 0; r < nrw; r++) g[r] /= nrm;
    rcol[kk] = nrm;
    for (int q = 0; q <= kk; q++) R[(size_t)kk * kout + q] = rcol[q];  // column kk of R
    keep.push_back(jj);
    kk++;
  }
  if (kk == 0) return 0;
  // P R^{-1} (nc x kk): the upper-triangular system column by column
  std::vector<double> PR((size_t)nc * kk, 0.0);
  for (int jj = 0; jj < kk; jj++) {
    for (int l = 0; l < nc; l++) {
      double v = Z[(size_t)keep[jj] * nc + l];
      for (int q = 0; q < jj; q++) v -= PR[(size_t)q * nc + l] * R[(size_t)jj * kout + q];
      PR[(size_t)jj * nc + l] = v / R[(size_t)jj * kout + jj];
    }
  }
  Qm.assign((size_t)nrw * kk, 0.0);
  PRu.assign((size_t)std::max(kd, 1) * kk, 0.0);
  PRv.assign((size_t)std::max(jend, 1) * kk, 0.0);
  for (int jj = 0; jj < kk; jj++) {
    for (int r = 0; r < nrw; r++) Qm[(size_t)jj * nrw + r] = GP[(size_t)keep[jj] * nrw + r];
    for (int i = 0; i < kd; i++) PRu[(size_t)jj * kd + i] = PR[(size_t)jj * nc + i] * dscale[i];
    for (int i = 0; i < jend; i++) PRv[(size_t)jj * jend + i] = PR[(size_t)jj * nc + kd + i];
  }
  return kk;
}
}  // namespace

int TACSMixedGMRES::solveDeflated(TACSVec* tb, TACSVec* tx, int zero_guess) {
  TACSBVec* b = dynamic_cast<TACSBVec*>(tb);
  TACSBVec* x = dynamic_cast<TACSBVec*>(tx);
  rec->rtol = rtol;
  rec->atol = atol;
  rec->nrestart = nrestart;
  rec->setMonitor(monitor);
  const double* bd = b->devRead();
  double* xd = zero_guess ? x->devWrite() : (x->devRead(), x->devWrite());
  int flag = rec->solveBlock(1, bd, xd, zero_guess);
  if (flag < 0) return 0;
  iterCount = rec->iters.empty() ? 0 : rec->iters[0];
  resNorm = rec->rn.empty() ? 0.0 : rec->rn[0];
  outerCount = rec->cycles;
  recycleDim = rec->kdj.empty() ? 0 : rec->kdj[0];
  recycleVersion = rec->defVersion;
  return flag;
}

// ---------------------------------------------------------------------------
// Batched GMRES: R restarted GMRES(m) recurrences (the recurrence of
// GMRES::solve src/bpmat/KSM.cpp:785-958, classical Gram-Schmidt twice)
// advanced together over interleaved block vectors.  Each cycle starts from
// the true FP64 residual R = B - A X with V_0(:, j) = R(:, j) / |R(:, j)| and
// res_j[0] = 1; a right-hand side whose Givens estimate reaches
// max(rtol |r_j(0)|, atol) / |r_j| keeps its step count cnt_j and its later
// columns are ignored; the cycle ends when every active right-hand side has
// converged or m steps are done; then X(:, j) += |r_j| M^{-1} V y_j.
// Right-hand sides that finished are masked out of later cycles (zero start
// vector, cnt_j = 0).
// ---------------------------------------------------------------------------
TACSBatchedGMRES::TACSBatchedGMRES(TACSParallelMat* m, TACSBlockJacobiPc* p, int _m, int _nrestart,
                                   int nrhs, int _prec)
    : rtol(1e-8), atol(1e-30), msub(_m), nrestart(_nrestart), R(nrhs), prec(_prec), cycles(0),
      kdefl(0), kVersion(-1), d_U(NULL), d_C(NULL), ndefl(0), defVersion(-1), defR(0),
      d_Ub(NULL), d_Cb(NULL), d_Unb(NULL), d_Cnb(NULL), d_prod(NULL), h_stage(NULL), h_up(NULL),
      ev_up(NULL), up_pending(0), d_lsst(NULL), d_ist(NULL), d_kdv(NULL), d_tgtv(NULL),
      h_flags(NULL), d_flags(NULL), ev_step{NULL, NULL, NULL}, d_dq(NULL),
      delayed(1), mat(m), pc(p) {
  allocate();
}
// Segmented systems: the nseg diagonal blocks (node ranges [node_off[j],
// node_off[j+1])) of the block-diagonal matrix m, each solved as its own
// system with its own Krylov space and stopping test.
TACSBatchedGMRES::TACSBatchedGMRES(TACSParallelMat* m, TACSBlockJacobiPc* p, int _m, int _nrestart,
                                   int _nseg, const int* node_off, int _prec)
    : rtol(1e-8), atol(1e-30), msub(_m), nrestart(_nrestart), R(_nseg), prec(_prec), cycles(0),
      kdefl(0), kVersion(-1), d_U(NULL), d_C(NULL), ndefl(0), defVersion(-1), defR(0),
      d_Ub(NULL), d_Cb(NULL), d_Unb(NULL), d_Cnb(NULL), d_prod(NULL), h_stage(NULL), h_up(NULL),
      ev_up(NULL), up_pending(0), d_lsst(NULL), d_ist(NULL), d_kdv(NULL), d_tgtv(NULL),
      h_flags(NULL), d_flags(NULL), ev_step{NULL, NULL, NULL}, d_dq(NULL),
      delayed(1), mat(m), pc(p) {
  // every column of a row of system j inside system j: the segmented
  // kernels index x relative to the system's first node
  if (R < 1 || R > 64 || !node_off || node_off[0] != 0 || node_off[R] != m->nrows ||
      m->bsize != 3) {
    tacs_set_error("TACSBatchedGMRES: segments need 1 <= nseg <= 64 ranges covering the %d rows "
                   "of a 3x3-block matrix",
                   m->nrows);
    R = 0;
  } else {
    for (int j = 0; j < R && R > 0; j++) {
      if (node_off[j + 1] < node_off[j]) {
        tacs_set_error("TACSBatchedGMRES: segment %d has a negative size", j);
        R = 0;
        break;
      }
      for (int r = node_off[j]; r < node_off[j + 1] && R > 0; r++)
        for (int s = m->h_rowp[r]; s < m->h_rowp[r + 1]; s++)
          if (m->h_cols[s] < node_off[j] || m->h_cols[s] >= node_off[j + 1]) {
            tacs_set_error("TACSBatchedGMRES: row %d of segment %d couples to node %d outside it",
                           r, j, m->h_cols[s]);
            R = 0;
            break;
          }
    }
  }
  if (R > 0) {
    nseg = R;
    segOff.assign(node_off, node_off + R + 1);
    segMaxNodes = 0;
    for (int j = 0; j < R; j++) segMaxNodes = std::max(segMaxNodes, node_off[j + 1] - node_off[j]);
    d_segoff = dalloc<int>(R + 1);
    h2d(d_segoff, node_off, R + 1);
    d_ux = dalloc<double>((size_t)m->nrows * 3);
    d_uy = dalloc<double>((size_t)m->nrows * 3);
  } else {
    R = 1;  // a valid (unused) size: the object is destroyed by the C ABI
  }
  allocate();
}
void TACSBatchedGMRES::allocate() {
  mat->incref();
  pc->incref();
  if (prec == 32) mat->getValues32(0);
  size_t N = (size_t)sysRows() * mat->bsize * R;
  d_V = dalloc<double>((size_t)(msub + 1) * N);
  d_work = dalloc<double>(N);
  d_Bb = dalloc<double>(N);
  d_Xb = dalloc<double>(N);
  d_scal = dalloc<double>(slice() * R);
  HCHECK(hipMemsetAsync(d_scal, 0, slice() * R * sizeof(double), g_stream));
  d_vec = dalloc<double>(4 * 64);
  d_cnt = dalloc<int>(64);
  size_t rb = tacs_hip_batch_scratch_bytes(R);
  HCHECK(hipMalloc(&d_red, rb));
  HCHECK(hipHostMalloc((void**)&h_pin, 512 * sizeof(double), 0));
}
TACSBatchedGMRES::~TACSBatchedGMRES() {
  HCHECK(hipStreamSynchronize(g_stream));
  void* bufs[] = {d_V, d_work, d_Bb,  d_Xb,  d_scal, d_vec, d_cnt, d_red,
                  d_U, d_C,    d_Ub,  d_Cb,  d_Unb,  d_Cnb, d_prod, d_dq, d_segoff, d_ux, d_uy,
                  d_cd, d_caz, d_cz32};
  for (void* q : bufs) dfree(q);
  hfree(h_pin);
  if (h_stage) hfree(h_stage);
  if (h_up) hfree(h_up);
  if (ev_up) HCHECK(hipEventDestroy(ev_up));
  void* st[] = {d_lsst, d_ist, d_kdv, d_tgtv};
  for (void* q : st) dfree(q);
  if (h_flags) hfree(h_flags);
  for (hipEvent_t e : ev_step)
    if (e) HCHECK(hipEventDestroy(e));
  if (monitor) monitor->decref();
  mat->decref();
  pc->decref();
}
void TACSBatchedGMRES::setTolerances(double r, double a) {
  rtol = r;
  atol = a;
}

int TACSBatchedGMRES::setRecycleSpace(TACSMixedGMRES* src) {
  if (nseg > 0) {
    tacs_set_error("setRecycleSpace: not for segmented systems");
    return -1;
  }
  HCHECK(hipStreamSynchronize(g_stream));
  dfree(d_U);
  dfree(d_C);
  d_U = d_C = NULL;
  kdefl = 0;
  kVersion = -1;
  if (src->mat != mat) {
    tacs_set_error("setRecycleSpace: the recycling solver belongs to another matrix");
    return -1;
  }
  int k = src->recycleDim;
  if (k <= 0 || !src->rec || src->recycleVersion != mat->version) return 0;
  if (k > KD_MAX) k = KD_MAX;
  const size_t n = (size_t)mat->nrows * mat->bsize;
  d_U = dalloc<double>((size_t)k * n);
  d_C = dalloc<double>((size_t)k * n);
  // the source's one right-hand side space: k vectors of n each
  HCHECK(hipMemcpyAsync(d_U, src->rec->d_Ub, (size_t)k * n * sizeof(double),
                        hipMemcpyDeviceToDevice, g_stream));
  HCHECK(hipMemcpyAsync(d_C, src->rec->d_Cb, (size_t)k * n * sizeof(double),
                        hipMemcpyDeviceToDevice, g_stream));
  HCHECK(hipStreamSynchronize(g_stream));
  if (g_halt) {
    dfree(d_U);
    dfree(d_C);
    d_U = d_C = NULL;
    return -1;
  }
  kdefl = k;
  kVersion = mat->version;
  return k;
}

int TACSBatchedGMRES::setDeflation(int k) {
  if (k < 0 || k > 31 || (k > 0 && (k >= msub - 1 || msub + k > 250))) {
    tacs_set_error("setDeflation: need 0 <= k <= 31, k < m - 1 and m + k <= 250 (k = %d, m = %d)",
                   k, msub);
    return -1;
  }
  HCHECK(hipStreamSynchronize(g_stream));
  auto release = [&]() {
    void* bufs[] = {d_Ub, d_Cb, d_Unb, d_Cnb, d_prod, d_dq};
    for (void* q : bufs) dfree(q);
    d_Ub = d_Cb = d_Unb = d_Cnb = d_prod = NULL;
    d_dq = NULL;
    hfree(h_stage);
    hfree(h_up);
    h_stage = h_up = NULL;
    up_pending = 0;
    void* st[] = {d_lsst, d_ist, d_kdv, d_tgtv};
    for (void* q : st) dfree(q);
    d_lsst = d_tgtv = NULL;
    d_ist = d_kdv = NULL;
    hfree(h_flags);
    h_flags = d_flags = NULL;
  };
  release();
  ndefl = k;
  defVersion = -1;
  defR = 0;
  kdj.assign(R, 0);
  dsc.assign((size_t)R * KD_MAX, 0.0);
  if (k > 0) {
    size_t N = (size_t)sysRows() * mat->bsize * R;
    d_Ub = dalloc<double>((size_t)k * N);
    d_Cb = dalloc<double>((size_t)k * N);
    d_Unb = dalloc<double>((size_t)k * N);
    d_Cnb = dalloc<double>((size_t)k * N);
    d_prod = dalloc<double>((size_t)R * prodStride());
    HCHECK(hipHostMalloc((void**)&h_stage, (size_t)R * prodStride() * sizeof(double), 0));
    HCHECK(hipHostMalloc((void**)&h_up, (size_t)R * prodStride() * sizeof(double), 0));
    if (!ev_up) HCHECK(hipEventCreateWithFlags(&ev_up, hipEventDisableTiming));
    const int ld = msub + 2;
    d_lsst = dalloc<double>((size_t)R * tacs_hip_bdcgs_state_stride(ld));
    d_ist = dalloc<int>((size_t)4 * R);
    d_kdv = dalloc<int>(R);
    d_tgtv = dalloc<double>(R);
    HCHECK(hipHostMalloc((void**)&h_flags, (size_t)(DCGS_LAG + 1) * 4 * R * sizeof(int), 0));
    if (h_flags) memset(h_flags, 0, (size_t)(DCGS_LAG + 1) * 4 * R * sizeof(int));
    HCHECK(hipHostGetDevicePointer((void**)&d_flags, h_flags, 0));
    for (hipEvent_t& e : ev_step)
      if (!e) HCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    d_dq = dalloc<char>(std::max(tacs_hip_bdotq_scratch_bytes(R), tacs_hip_btsgemm_scratch_bytes(R)));
    if (g_halt) {  // all of the recycling workspace, or none of it
      release();
      ndefl = 0;
      kdj.assign(R, 0);
      return -1;
    }
  }
  return 0;
}

int TACSBatchedGMRES::solveSegments(TACSBVec* b, TACSBVec* x, int zero_guess) {
  if (nseg <= 0) {
    tacs_set_error("solveSegments: the solver was not built for segmented systems");
    return -1;
  }
  const size_t n = (size_t)mat->nrows * 3;
  if (b->n != n || x->n != n) {
    tacs_set_error("solveSegments: vectors of %zu entries expected", n);
    return -1;
  }
  KCHECK(tacs_hip_seg_pack(segMaxNodes, R, d_segoff, b->devRead(), d_Bb, g_stream));
  if (!zero_guess) KCHECK(tacs_hip_seg_pack(segMaxNodes, R, d_segoff, x->devRead(), d_Xb, g_stream));
  int flag = solveBlock(R, d_Bb, d_Xb, zero_guess);
  KCHECK(tacs_hip_seg_unpack(mat->nrows, R, d_segoff, d_Xb, x->devWrite(), g_stream));
  HCHECK(hipStreamSynchronize(g_stream));
  return flag;
}
// Block-Jacobi and products of the batched cycles: the whole matrix for every
// right-hand side, or the segmented systems' own blocks
void TACSBatchedGMRES::bdiag(int pr, int Rn, const void* Dinv, const double* x, double* y) {
  if (nseg > 0)
    KCHECK(tacs_hip_seg_bdiag_apply(pr, segMaxNodes, Rn, d_segoff, Dinv, x, y, g_stream));
  else
    KCHECK(tacs_hip_bdiag_apply(mat->bsize, pr, mat->nrows, Rn, Dinv, x, y, g_stream));
}
void TACSBatchedGMRES::precond(int pr, int Rn, const void* Dinv, const float* A32,
                               const double* x, double* y) {
  const int deg = pc->chebDegree;
  if (deg <= 1) {
    bdiag(pr, Rn, Dinv, x, y);
    return;
  }
  // Chebyshev iteration for A z = x on [lmax / ratio, lmax] (Saad, Alg. 12.1),
  // the iterate z kept in the output (distinct from the input at every call);
  // segmented systems run it on whole-matrix vectors (their blocks are the
  // diagonal blocks of the matrix, so the polynomial acts on each alone)
  if (!d_cd) {
    const size_t Nmax = (size_t)mat->nrows * mat->bsize * (nseg > 0 ? 1 : R);
    d_cd = dalloc<double>(Nmax);
    d_caz = dalloc<double>(Nmax);
    if (g_halt) return;
  }
  const double b = pc->chebMax, a = b / pc->chebRatio;
  const double theta = 0.5 * (b + a), delta = 0.5 * (b - a), sigma = theta / delta;
  const int rc = nseg > 0 ? 1 : Rn;
  const double* xc = x;
  double* zc = y;
  if (nseg > 0) {
    KCHECK(tacs_hip_seg_to_union(segMaxNodes, Rn, d_segoff, x, d_ux, g_stream));
    xc = d_ux;
    zc = d_uy;
  }
  auto mult = [&](const double* in, double* out) {
    if (nseg == 0)
      bmult(pr, Rn, A32, in, out);
    else if (pr == 32)
      mat->multMixed(in, out);
    else
      mat->mult64(in, out);
  };
  double rho = 1.0 / sigma;
  if (rc == 1) {
    // one system: the step fused into the product, iterates alternating
    // between two buffers so that the last lands in zc; the products inside
    // the polynomial may read bfloat16 values (a fixed linear operator
    // still: FP64 vectors and sums)
    const int pin = (pr == 32 && pc->chebValues == 16) ? 16 : pr;
    double* buf[2];
    buf[(deg - 1) % 2] = zc;
    buf[deg % 2] = d_caz;
    KCHECK(tacs_hip_bcheb_step(mat->bsize, pr, mat->nrows, 1, Dinv, xc, NULL, d_cd, buf[0], 0.0,
                               1.0 / theta, 1, NULL, g_stream));
    // from the third iterate on, the three-term form (the iterate two steps
    // back is in the output buffer; no d vector moves)
    for (int k = 1; k < deg; k++) {
      const double rn = 1.0 / (2.0 * sigma - rho);
      if (k == 1)
        mat->chebStep(pin, Dinv, buf[0], buf[1], xc, d_cd, rn * rho, 2.0 * rn / delta);
      else
        mat->chebStep3(pin, Dinv, buf[(k - 1) % 2], buf[k % 2], xc, rn * rho, 2.0 * rn / delta);
      rho = rn;
    }
  } else if (pr == 32 && pc->chebValues == 16 && mat->bsize == 3 && rc % 4 == 0 && rc >= 16) {
    // R right-hand sides, reduced precision inside the polynomial: the
    // products read an FP32 copy of the iterate (half the gather traffic),
    // each step fused into its product; the iterate and d stay FP64
    if (!d_cz32) {
      d_cz32 = dalloc<float>(2 * (size_t)mat->nrows * 3 * R);
      if (g_halt) return;
    }
    float* z32[2] = {d_cz32, d_cz32 + (size_t)mat->nrows * 3 * R};
    const int* rowp = mat->d_rowp;
    const int* cols = mat->d_cols;
    KCHECK(tacs_hip_bcheb_step(3, 32, mat->nrows, rc, Dinv, xc, NULL, d_cd, zc, 0.0, 1.0 / theta,
                               1, z32[0], g_stream));
    // steps in place (z += d) until step t, which writes the next iterate
    // over d; from then on the three-term form alternates the iterate
    // between zc and d_cd (t of the degree's parity, so the last lands in zc)
    const int t = deg % 2 == 0 ? 2 : 3;
    double* cur = zc;
    double* prev = d_cd;
    for (int k = 1; k < deg; k++) {
      const double rn = 1.0 / (2.0 * sigma - rho);
      const int mode = (deg < t + 2 || k < t) ? 0 : k == t ? 2 : 1;
      KCHECK(tacs_hip_bsr3_spmm_cheb_mode(mode, mat->nrows, rc, rowp, cols, A32,
                                          z32[(k - 1) % 2], xc, (const float*)Dinv, prev, cur,
                                          z32[k % 2], rn * rho, 2.0 * rn / delta, g_stream));
      if (mode != 0) std::swap(cur, prev);
      rho = rn;
    }
  } else {
    KCHECK(tacs_hip_bcheb_step(mat->bsize, pr, mat->nrows, rc, Dinv, xc, NULL, d_cd, zc, 0.0,
                               1.0 / theta, 1, NULL, g_stream));
    for (int k = 1; k < deg; k++) {
      mult(zc, d_caz);
      const double rn = 1.0 / (2.0 * sigma - rho);
      KCHECK(tacs_hip_bcheb_step(mat->bsize, pr, mat->nrows, rc, Dinv, xc, d_caz, d_cd, zc,
                                 rn * rho, 2.0 * rn / delta, 0, NULL, g_stream));
      rho = rn;
    }
  }
  if (nseg > 0) KCHECK(tacs_hip_seg_from_union(segMaxNodes, Rn, d_segoff, zc, y, g_stream));
}
void TACSBatchedGMRES::bmult(int pr, int Rn, const float* A32, const double* x, double* y) {
  if (nseg > 0) {
    // through a whole-matrix vector: the matrix's own sliced ELL product
    KCHECK(tacs_hip_seg_to_union(segMaxNodes, Rn, d_segoff, x, d_ux, g_stream));
    if (pr == 32)
      mat->multMixed(d_ux, d_uy);
    else
      mat->mult64(d_ux, d_uy);
    KCHECK(tacs_hip_seg_from_union(segMaxNodes, Rn, d_segoff, d_uy, y, g_stream));
  } else if (Rn == 1 && pr == 32)
    mat->multMixed(x, y);
  else if (Rn == 1)
    mat->mult64(x, y);
  else if (pr == 32)
    KCHECK(tacs_hip_bsr3_spmm_mixed(mat->nrows, Rn, mat->d_rowp, mat->d_cols, A32, x, y, g_stream));
  else
    KCHECK(tacs_hip_bsr3_spmm(mat->nrows, Rn, mat->d_rowp, mat->d_cols, mat->d_A, x, y, g_stream));
}
int TACSBatchedGMRES::solve(int nrhs, TACSBVec** b, TACSBVec** x, int zero_guess) {
  if (nseg > 0) {
    tacs_set_error("TACSBatchedGMRES::solve: segmented systems are solved with solveSegments");
    return -1;
  }
  if (nrhs < 1 || nrhs > R) {
    tacs_set_error("TACSBatchedGMRES::solve: nrhs = %d outside [1, %d]", nrhs, R);
    return -1;
  }
  const size_t n = (size_t)mat->nrows * mat->bsize;
  for (int j = 0; j < nrhs; j++) {
    if (b[j]->n != n || x[j]->n != n) {
      tacs_set_error("TACSBatchedGMRES::solve: vector %d has the wrong size", j);
      return -1;
    }
    KCHECK(tacs_hip_bpack(n, nrhs, j, b[j]->devRead(), d_Bb, g_stream));
    if (!zero_guess) KCHECK(tacs_hip_bpack(n, nrhs, j, x[j]->devRead(), d_Xb, g_stream));
  }
  int flag = solveBlock(nrhs, d_Bb, d_Xb, zero_guess);
  for (int j = 0; j < nrhs; j++) KCHECK(tacs_hip_bunpack(n, nrhs, j, d_Xb, x[j]->devWrite(), g_stream));
  HCHECK(hipStreamSynchronize(g_stream));
  return flag;
}

int TACSBatchedGMRES::solveBlock(int Rn, const double* Bd, double* Xd, int zero_guess) {
  if (Rn < 1 || Rn > R) {
    tacs_set_error("TACSBatchedGMRES: nrhs = %d outside [1, %d]", Rn, R);
    return -1;
  }
  // several right-hand sides share the 3x3 multi-RHS SpMM; one right-hand
  // side runs the single-vector products of any block size
  if (Rn > 1 && mat->bsize != 3) {
    tacs_set_error("TACSBatchedGMRES: %d right-hand sides need 3x3 blocks (b = %d)", Rn,
                   mat->bsize);
    return -1;
  }
  if (nseg > 0 && Rn != nseg) {
    tacs_set_error("TACSBatchedGMRES: %d segmented systems, not %d", nseg, Rn);
    return -1;
  }
  if (ndefl > 0) return solveBlockDeflated(Rn, Bd, Xd, zero_guess);
  const int nr = sysRows(), m = msub;
  const size_t n = (size_t)nr * mat->bsize, N = n * Rn, S = slice();
  // per right-hand side slice (stride S)
  double* H = d_scal;
  double* res = d_scal + (size_t)(m + 1) * m;
  double* Qc = res + (m + 1);
  double* Qs = Qc + m;
  double* tmp = Qs + m;
  double* Bc = tmp + (m + 1);             // KD_MAX x m, column i at Bc + i*KD_MAX
  double* tmpC = Bc + (size_t)KD_MAX * m;
  double* zc = tmpC + KD_MAX;
  double* pco = zc + KD_MAX;
  double *nrm2 = d_vec, *nrm = d_vec + 64, *inv = d_vec + 128, *st = d_vec + 192;
  double *h_inv = h_pin, *h_res0 = h_pin + 64, *h_read = h_pin + 256;
  int* h_cnt = (int*)(h_pin + 128);
  auto V = [&](int i) { return d_V + (size_t)i * N; };
  const float* A32 = prec == 32 ? mat->getValues32(1) : NULL;
  const void* Dinv = prec == 32 ? (const void*)pc->d_Dinv32 : (const void*)pc->d_Dinv;
  const int kd = (kdefl > 0 && kVersion == mat->version) ? kdefl : 0;
  auto spmm64 = [&](const double* x, double* y) { bmult(64, Rn, NULL, x, y); };
  // y = A M^{-1} x with the Arnoldi operator of this precision
  auto op = [&](const double* x, double* y) {
    precond(prec, Rn, Dinv, A32, x, d_work);
    bmult(prec, Rn, A32, d_work, y);
  };
  // q(:, j) /= |q(:, j)|, with |q(:, j)| written to out[j*sOut]
  auto normalize = [&](double* q, double* out, size_t sOut) {
    KCHECK(tacs_hip_bmdot(n, Rn, 1, q, N, q, (double*)d_red, out, sOut, 0, g_stream));
    KCHECK(tacs_hip_bsqrt_inv(Rn, out, sOut, out, sOut, inv, g_stream));
    KCHECK(tacs_hip_bscale(n, Rn, inv, q, g_stream));
  };
  std::vector<char> active(Rn, 1), hit(Rn, 0), fail(Rn, 0), conv(Rn);
  std::vector<double> target(Rn);
  iters.assign(Rn, 0);
  flags.assign(Rn, 0);
  r0.assign(Rn, 0.0);
  rn.assign(Rn, 0.0);
  cycles = 0;
  if (zero_guess) {
    KCHECK(tacs_hip_fill(64, N, Xd, 0.0, g_stream));
    HCHECK(hipMemcpyAsync(V(0), Bd, N * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
  } else {
    spmm64(Xd, V(0));
    KCHECK(tacs_hip_axpby(64, 64, N, 1.0, Bd, -1.0, V(0), g_stream));
  }
  for (int count = 0;; count++) {
    if (g_halt) return 0;  // a copy or launch failed: no further work
    // FP64 residual norms
    KCHECK(tacs_hip_bmdot(n, Rn, 1, V(0), N, V(0), (double*)d_red, nrm2, 1, 0, g_stream));
    KCHECK(tacs_hip_bsqrt_inv(Rn, nrm2, 1, nrm, 1, inv, g_stream));
    d2h(h_read, nrm, Rn);
    int nact = 0;
    for (int j = 0; j < Rn; j++) {
      if (!active[j]) continue;
      double r = h_read[j];
      if (monitor) monitor->printResidual(0, r);
      if (count == 0) r0[j] = r;
      rn[j] = r;
      if (!std::isfinite(r) || fail[j]) {
        active[j] = 0;
      } else if (r <= rtol * r0[j] || r < atol || hit[j]) {
        active[j] = 0;
        flags[j] = 1;
      } else {
        nact++;
      }
    }
    if (nact == 0 || count > nrestart) break;
    cycles++;
    // V_0 = r / |r| for the active right-hand sides, zero for the others
    for (int j = 0; j < Rn; j++) {
      h_inv[j] = active[j] ? 1.0 / rn[j] : 0.0;
      target[j] = active[j] ? std::max(rtol * r0[j], atol) / rn[j] : 0.0;
      conv[j] = !active[j];
      h_cnt[j] = 0;
    }
    h2d(inv, h_inv, Rn);
    KCHECK(tacs_hip_bscale(n, Rn, inv, V(0), g_stream));
    if (kd > 0) {
      // x += |r| M^{-1} U C^T v_0;  v_0 -= C C^T v_0
      KCHECK(tacs_hip_bmdot_shared(n, Rn, kd, d_C, n, V(0), (double*)d_red, pco, S, 0, g_stream));
      KCHECK(tacs_hip_blincomb_shared(n, Rn, kd, d_U, n, pco, S, V(1), 0, g_stream));
      precond(prec, Rn, Dinv, A32, V(1), d_work);
      KCHECK(tacs_hip_bscale(n, Rn, nrm, d_work, g_stream));
      KCHECK(tacs_hip_axpby(64, 64, N, 1.0, d_work, 1.0, Xd, g_stream));
      KCHECK(tacs_hip_bcgs_update_shared(n, Rn, kd, d_C, n, pco, S, V(0), g_stream));
    }
    if (kd == 0) {
      // res_j[0] = 1 (0 for the inactive right-hand sides)
      for (int j = 0; j < Rn; j++) h_res0[j] = active[j] ? 1.0 : 0.0;
      HCHECK(hipMemcpy2DAsync(res, S * sizeof(double), h_res0, sizeof(double), sizeof(double), Rn,
                              hipMemcpyHostToDevice, g_stream));
    } else {
      // res_j[0] = |(I - C C^T) v_0(:, j)|, the deflated residual normalised
      normalize(V(0), res, S);
      HCHECK(hipMemcpy2DAsync(h_read, sizeof(double), res, S * sizeof(double), sizeof(double), Rn,
                              hipMemcpyDeviceToHost, g_stream));
      HCHECK(hipStreamSynchronize(g_stream));
      for (int j = 0; j < Rn; j++)
        if (!conv[j] && h_read[j] <= target[j]) conv[j] = 1;  // the projection alone converged
    }
    int niters = 0;
    bool all = true;
    for (int j = 0; j < Rn; j++) all = all && conv[j];
    for (int i = 0; i < m && !all; i++) {
      double* q = V(i + 1);
      double* h = H + (size_t)i * (m + 1);
      double* bc = Bc + (size_t)i * KD_MAX;
      op(V(i), q);
      // classical Gram-Schmidt, twice, against [C, V_0..V_i]: (bc, h) = [C V]^T q
      // (both passes), q -= C bc + V h
      for (int pass = 0; pass < 2; pass++) {
        double* cc = pass == 0 ? bc : tmpC;
        double* hh = pass == 0 ? h : tmp;
        if (kd > 0)
          KCHECK(tacs_hip_bmdot_shared(n, Rn, kd, d_C, n, q, (double*)d_red, cc, S, 0, g_stream));
        KCHECK(tacs_hip_bmdot(n, Rn, i + 1, V(0), N, q, (double*)d_red, hh, S, 0, g_stream));
        if (kd > 0) KCHECK(tacs_hip_bcgs_update_shared(n, Rn, kd, d_C, n, cc, S, q, g_stream));
        KCHECK(tacs_hip_bcgs_update(n, Rn, i + 1, V(0), N, hh, S, q, g_stream));
      }
      if (kd > 0) KCHECK(tacs_hip_badd(Rn, kd, tmpC, S, bc, S, g_stream));
      KCHECK(tacs_hip_badd(Rn, i + 1, tmp, S, h, S, g_stream));
      normalize(q, h + i + 1, S);
      KCHECK(tacs_hip_bgivens(Rn, i, m, H, S, Qc, Qs, S, res, S, st, g_stream));
      niters = i + 1;
      d2h(h_read, st, Rn);
      all = true;
      for (int j = 0; j < Rn; j++) {
        if (conv[j]) continue;
        double e = h_read[j];
        if (!std::isfinite(e)) {
          conv[j] = 1;
          fail[j] = 1;
          h_cnt[j] = i;
        } else if (e <= target[j]) {
          conv[j] = 1;
          h_cnt[j] = i + 1;
          // with the FP64 operator and no recycled space the estimate is the
          // true residual (GMRES::solve stops on it)
          if (prec == 64 && kd == 0) hit[j] = 1;
        } else {
          all = false;
        }
      }
    }
    for (int j = 0; j < Rn; j++) {
      if (!conv[j]) h_cnt[j] = niters;
      if (active[j]) iters[j] += h_cnt[j];
    }
    if (niters > 0) {
      h2d(d_cnt, h_cnt, Rn);
      // X(:, j) += |r_j| M^{-1} (V y_j - U B_j y_j);  R = B - A X
      KCHECK(tacs_hip_bbacksolve(Rn, m, d_cnt, H, S, res, S, g_stream));
      KCHECK(tacs_hip_blincomb(n, Rn, niters, V(0), N, res, S, d_cnt, d_work, 0, g_stream));
      if (kd > 0) {
        KCHECK(tacs_hip_bsmall_gemv(Rn, kd, d_cnt, Bc, S, KD_MAX, res, S, zc, S, g_stream));
        KCHECK(tacs_hip_bcgs_update_shared(n, Rn, kd, d_U, n, zc, S, d_work, g_stream));
      }
      precond(prec, Rn, Dinv, A32, d_work, V(1));
      KCHECK(tacs_hip_bscale(n, Rn, nrm, V(1), g_stream));
      KCHECK(tacs_hip_axpby(64, 64, N, 1.0, V(1), 1.0, Xd, g_stream));
    }
    spmm64(Xd, V(0));
    KCHECK(tacs_hip_axpby(64, 64, N, 1.0, Bd, -1.0, V(0), g_stream));
  }
  HCHECK(hipStreamSynchronize(g_stream));
  int ok = 1;
  for (int j = 0; j < Rn; j++) ok &= flags[j];
  return ok;
}

// ---------------------------------------------------------------------------
// Batched GCRO-DR: the cycle of TACSMixedGMRES::solveDeflated for R
// right-hand sides at once.  Right-hand side j keeps its own recycled space
// (U_j, C_j) as interleaved block vectors (columns beyond its dimension k_j
// are zero), its own least-squares problem and its own harmonic Ritz update
// on the host; the SpMM, the Gram-Schmidt products against [C_j, V_j], the
// corrections and the basis changes run batched over all right-hand sides.
// The Arnoldi operator is A M^{-1} with the FP64 (prec 64) or FP32-rounded
// (prec 32) matrix and block-Jacobi values, in FP64 arithmetic, so
// A M^{-1} U_j = C_j holds across cycles and solves.
// ---------------------------------------------------------------------------
int TACSBatchedGMRES::solveBlockDeflated(int Rn, const double* Bd, double* Xd, int zero_guess) {
  const int nr = sysRows(), m = msub, K = ndefl;
  const size_t n = (size_t)nr * mat->bsize, N = n * Rn, S = slice(), SP = prodStride();
  double* H = d_scal;
  double* res = d_scal + (size_t)(m + 1) * m;
  double* tmp = res + (m + 1) + 2 * (size_t)m;
  double* Bc = tmp + (m + 1);  // KD_MAX x m, column i at Bc + i*KD_MAX
  double* tmpC = Bc + (size_t)KD_MAX * m;
  double* zc = tmpC + KD_MAX;
  double* pco = zc + KD_MAX;
  double* dA = pco + KD_MAX;             // dot products with u (KD_MAX + m + 1)
  double* dB = dA + (KD_MAX + m + 1);    // dot products with z (KD_MAX + m + 1)
  double* dcf = dB + (KD_MAX + m + 1);   // update coefficients (2 (KD_MAX + m) + 2)
  double *nrm2 = d_vec, *nrm = d_vec + 64, *inv = d_vec + 128;
  double* h_inv = h_pin;
  auto V = [&](int i) { return d_V + (size_t)i * N; };
  auto Ub = [&](int v) { return d_Ub + (size_t)v * N; };
  auto Cb = [&](int v) { return d_Cb + (size_t)v * N; };
  // pr: precision of the operator's values; a solve that stalls with the
  // FP32-valued operator continues with the FP64 one (see the cycle loop)
  int pr = prec;
  bool fell_back = false;
  const float* A32 = prec == 32 ? mat->getValues32(1) : NULL;
  const void* Dinv = prec == 32 ? (const void*)pc->d_Dinv32 : (const void*)pc->d_Dinv;
  auto spmm64 = [&](const double* x, double* y) { bmult(64, Rn, NULL, x, y); };
  auto op = [&](const double* x, double* y) {
    precond(pr, Rn, Dinv, A32, x, d_work);
    bmult(pr, Rn, A32, d_work, y);
  };
  auto normalize = [&](double* q, double* out, size_t sOut) {
    KCHECK(tacs_hip_bmdot(n, Rn, 1, q, N, q, (double*)d_red, out, sOut, 0, g_stream));
    KCHECK(tacs_hip_bsqrt_inv(Rn, out, sOut, out, sOut, inv, g_stream));
    KCHECK(tacs_hip_bscale(n, Rn, inv, q, g_stream));
  };
  // Rn x cols blocks between host (dense) and device (row stride ds), through
  // the pinned staging buffer; both wait for the copy
  auto d2h_rows = [&](double* h, int cols, const double* d, size_t ds) {
    if (cols <= 0) return;
    HCHECK(hipMemcpy2DAsync(h_stage, cols * sizeof(double), d, ds * sizeof(double),
                            cols * sizeof(double), Rn, hipMemcpyDeviceToHost, g_stream));
    HCHECK(hipStreamSynchronize(g_stream));
    memcpy(h, h_stage, (size_t)Rn * cols * sizeof(double));
  };
  // uploads are not waited for (the kernels that read them follow in stream
  // order); the staging buffer is reused once the previous upload is done
  auto h2d_rows = [&](double* d, size_t ds, const double* h, int cols) {
    if (cols <= 0) return;
    if (up_pending) HCHECK(hipEventSynchronize(ev_up));
    memcpy(h_up, h, (size_t)Rn * cols * sizeof(double));
    HCHECK(hipMemcpy2DAsync(d, ds * sizeof(double), h_up, cols * sizeof(double),
                            cols * sizeof(double), Rn, hipMemcpyHostToDevice, g_stream));
    HCHECK(hipEventRecord(ev_up, g_stream));
    up_pending = 1;
  };
  std::vector<double> stage;
  // dsc[j][v] = 1 / |U_j(:, v)|
  auto update_dscale = [&](int k) {
    for (int v = 0; v < k; v++)
      KCHECK(tacs_hip_bmdot(n, Rn, 1, Ub(v), N, Ub(v), (double*)d_red, d_prod + v, SP, 0, g_stream));
    stage.assign((size_t)Rn * std::max(k, 1), 0.0);
    d2h_rows(stage.data(), k, d_prod, SP);
    for (int j = 0; j < Rn; j++)
      for (int v = 0; v < KD_MAX; v++) {
        double d = v < k ? stage[(size_t)j * k + v] : 0.0;
        dsc[(size_t)j * KD_MAX + v] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
      }
  };

  // Refresh of the recycled spaces: C_j = A M^-1 U_j recomputed with the
  // current operator, then C_j = Q R by Cholesky QR, twice, and U_j <- U_j R^-1,
  // so A M^-1 U_j = C_j with C_j orthonormal to rounding again.  The cycle-end
  // updates keep the relation only in exact arithmetic; over thousands of
  // cycles the drift makes the projection's residual estimate optimistic and
  // the true residual stalls.
  auto refresh_space = [&](int Kd) {
    if (Kd <= 0) return;
    for (int v = 0; v < Kd; v++) op(Ub(v), Cb(v));
    std::vector<double> G((size_t)Rn * Kd * Kd), P((size_t)Rn * Kd * Kd);
    for (int pass = 0; pass < 2; pass++) {
      KCHECK(tacs_hip_btsgemm(n, Rn, Kd, Cb(0), Kd, Cb(0), N, (double*)d_dq, d_prod, SP, Kd,
                              g_stream));
      d2h_rows(G.data(), Kd * Kd, d_prod, SP);
      std::fill(P.begin(), P.end(), 0.0);
      for (int j = 0; j < Rn; j++) {
        const int kd = kdj[j];
        const double* g = &G[(size_t)j * Kd * Kd];  // g[a + b Kd] = C_a . C_b
        double* pj = &P[(size_t)j * Kd * Kd];        // pj[v + c Kd] = (R^-1)_{vc}
        std::vector<double> L((size_t)kd * kd, 0.0);  // G = L L^T, row-major
        bool ok = true;
        for (int r = 0; r < kd && ok; r++)
          for (int c = 0; c <= r; c++) {
            double acc = g[r + (size_t)c * Kd];
            for (int q = 0; q < c; q++) acc -= L[(size_t)r * kd + q] * L[(size_t)c * kd + q];
            if (r == c) {
              if (!(acc > 0.0)) {
                ok = false;
                break;
              }
              L[(size_t)r * kd + r] = sqrt(acc);
            } else {
              L[(size_t)r * kd + c] = acc / L[(size_t)c * kd + c];
            }
          }
        if (!ok) {
          kdj[j] = 0;  // a space that lost its rank is dropped
          continue;
        }
        // R = L^T; R^-1 column by column (upper triangular)
        for (int c = 0; c < kd; c++)
          for (int v = c; v >= 0; v--) {
            double acc = v == c ? 1.0 : 0.0;
            for (int q = v + 1; q <= c; q++) acc -= L[(size_t)q * kd + v] * pj[q + (size_t)c * Kd];
            pj[v + (size_t)c * Kd] = acc / L[(size_t)v * kd + v];
          }
        for (int c = kd; c < Kd; c++) pj[c + (size_t)c * Kd] = kd == 0 ? 0.0 : 1.0;
      }
      h2d_rows(d_prod, SP, P.data(), Kd * Kd);
      KCHECK(tacs_hip_bbasis_combine(n, Rn, Kd, Cb(0), N, d_prod, SP, Kd, Kd, d_Cnb, N, 0,
                                     g_stream));
      KCHECK(tacs_hip_bbasis_combine(n, Rn, Kd, Ub(0), N, d_prod, SP, Kd, Kd, d_Unb, N, 0,
                                     g_stream));
      HCHECK(hipStreamSynchronize(g_stream));
      std::swap(d_Cb, d_Cnb);
      std::swap(d_Ub, d_Unb);
    }
    update_dscale(Kd);
  };
  int refreshes = 0;
  bool nodefl = false;  // deflation switched off after repeated stalls

  // the spaces belong to one matrix and one number of right-hand sides
  if (defVersion != mat->version || defR != Rn) std::fill(kdj.begin(), kdj.end(), 0);
  defVersion = mat->version;
  defR = Rn;
  int Kd = 0;
  for (int j = 0; j < Rn; j++) Kd = std::max(Kd, kdj[j]);
  if (Kd == 0 && kdefl > 0 && kVersion == mat->version) {
    // start every right-hand side from the shared space
    Kd = std::min(kdefl, K);
    for (int v = 0; v < Kd; v++) {
      KCHECK(tacs_hip_bbroadcast(n, Rn, d_U + (size_t)v * n, Ub(v), g_stream));
      KCHECK(tacs_hip_bbroadcast(n, Rn, d_C + (size_t)v * n, Cb(v), g_stream));
    }
    for (int j = 0; j < Rn; j++) kdj[j] = Kd;
    update_dscale(Kd);
  }

  std::vector<char> active(Rn, 1), fail(Rn, 0), conv(Rn);
  std::vector<double> target(Rn), beta(Rn), tgt(Rn), hb, hh, hc((size_t)KD_MAX + m + 2);
  std::vector<int> stall(Rn, 0);
  std::vector<char> est_hit(Rn, 0);
  std::vector<int> jend(Rn);
  std::vector<HostLS> ls(Rn, HostLS(m + 1));
  iters.assign(Rn, 0);
  flags.assign(Rn, 0);
  r0.assign(Rn, 0.0);
  rn.assign(Rn, 0.0);
  cycles = 0;

  // Arnoldi steps with delayed reorthogonalisation (DCGS2).  Step i applies
  // the operator to u_i (in V(i), orthogonalised once at step i - 1); one
  // sweep of dot products of Q_i = [C, V_0..V_{i-1}] and u_i with u_i and
  // z = A M^{-1} u_i gives the reorthogonalisation coefficients s of u_i,
  // |u_i|^2 and the first-pass coefficients t of z.  On the host,
  // omega = sqrt(|u_i|^2 - |s|^2) finishes column i - 1 (its first-pass
  // coefficients plus s, subdiagonal omega), and since
  // A M^{-1} v_i = (z - A M^{-1} Q_i s) / omega, with A M^{-1} V s_V expanded
  // by the finished columns (the C part of s, of rounding size, is dropped),
  // the first-pass column i is ([t; tau] - [B; H] s_V) / omega.  One update
  // sweep then forms v_i = (u_i - Q_i s) / omega and
  // u_{i+1} = (z - Q_i t - v_i tau) / omega.  The stopping test of column
  // i - 1 is made at step i.  Returns the number of finished columns.
  static_assert(DCGS_LAG + 1 == sizeof(ev_step) / sizeof(ev_step[0]), "one event per queued step");
  auto arnoldi_delayed = [&](int s, int Kd) -> int {
    std::vector<std::vector<double>> pend(Rn);
    std::vector<double> Ah, cf;
    bool all = true;
    for (int j = 0; j < Rn; j++) all = all && conv[j];
    if (all) return 0;
    // the least-squares state of every right-hand side goes to the device,
    // which runs the per-step small computation (tacs_hip_bdcgs_coeffs), so
    // the host only queues kernels; it reads each step's stopping flags
    // DCGS_LAG steps later and stops queueing once every system stopped
    const int ld = m + 2;
    const size_t sLS = tacs_hip_bdcgs_state_stride(ld);
    const size_t HH = (size_t)ld * (ld - 1);
    std::vector<double> hls((size_t)Rn * sLS, 0.0);
    std::vector<int> hist((size_t)4 * Rn, 0), hkd(Rn);
    for (int j = 0; j < Rn; j++) {
      const HostLS& L = ls[j];
      double* q = &hls[(size_t)j * sLS];
      std::copy(L.Hb.begin(), L.Hb.end(), q);
      std::copy(L.Hr.begin(), L.Hr.end(), q + HH);
      double* g = q + 2 * HH;
      std::copy(L.g.begin(), L.g.end(), g);
      std::copy(L.c.begin(), L.c.end(), g + ld);
      std::copy(L.rc.begin(), L.rc.end(), g + 2 * ld);
      std::copy(L.rs.begin(), L.rs.end(), g + 3 * ld);
      hist[4 * j] = conv[j];
      hist[4 * j + 1] = fail[j];
      hist[4 * j + 2] = jend[j];
      hist[4 * j + 3] = (int)L.rc.size();
      hkd[j] = kdj[j];
    }
    HCHECK(hipMemcpyAsync(d_lsst, hls.data(), hls.size() * sizeof(double), hipMemcpyHostToDevice,
                          g_stream));
    HCHECK(hipMemcpyAsync(d_ist, hist.data(), hist.size() * sizeof(int), hipMemcpyHostToDevice,
                          g_stream));
    HCHECK(hipMemcpyAsync(d_kdv, hkd.data(), Rn * sizeof(int), hipMemcpyHostToDevice, g_stream));
    HCHECK(hipMemcpyAsync(d_tgtv, tgt.data(), Rn * sizeof(double), hipMemcpyHostToDevice,
                          g_stream));
    int i = 0;
    for (; i < s; i++) {
      if (i >= DCGS_LAG) {
        const int q = (i - DCGS_LAG) % (DCGS_LAG + 1);
        HCHECK(hipEventSynchronize(ev_step[q]));
        const volatile int* f = h_flags + (size_t)q * 4 * R;
        bool alld = true;
        for (int j = 0; j < Rn; j++) alld = alld && f[4 * j] != 0;
        if (alld) break;
      }
      op(V(i), V(i + 1));
      KCHECK(tacs_hip_bdotq(n, Rn, 2, Kd, Cb(0), i + 1, V(0), N, V(i), V(i + 1), (double*)d_dq,
                            dA, dB, S, g_stream));
      KCHECK(tacs_hip_bdcgs_coeffs(Rn, i, Kd, ld, d_kdv, d_tgtv, dA, dB, S, d_lsst, sLS, d_ist,
                                   dcf, S, d_flags + (size_t)(i % (DCGS_LAG + 1)) * 4 * R,
                                   g_stream));
      HCHECK(hipEventRecord(ev_step[i % (DCGS_LAG + 1)], g_stream));
      KCHECK(tacs_hip_bdcgs_update(n, Rn, Kd, Cb(0), i, V(0), N, dcf, S, V(i), V(i + 1), g_stream));
    }
    // the state back to the host
    HCHECK(hipMemcpyAsync(hls.data(), d_lsst, hls.size() * sizeof(double), hipMemcpyDeviceToHost,
                          g_stream));
    HCHECK(hipMemcpyAsync(hist.data(), d_ist, hist.size() * sizeof(int), hipMemcpyDeviceToHost,
                          g_stream));
    HCHECK(hipStreamSynchronize(g_stream));
    all = true;
    int last = -1;  // the step at which the last system stopped
    for (int j = 0; j < Rn; j++) {
      HostLS& L = ls[j];
      const double* q = &hls[(size_t)j * sLS];
      std::copy(q, q + HH, L.Hb.begin());
      std::copy(q + HH, q + 2 * HH, L.Hr.begin());
      const double* g = q + 2 * HH;
      std::copy(g, g + ld, L.g.begin());
      std::copy(g + ld, g + 2 * ld, L.c.begin());
      const int ncol = hist[4 * j + 3];
      L.rp.resize(ncol);
      for (int r = 0; r < ncol; r++) L.rp[r] = r;
      L.rc.assign(g + 2 * ld, g + 2 * ld + ncol);
      L.rs.assign(g + 3 * ld, g + 3 * ld + ncol);
      pend[j].assign(g + 4 * ld, g + 5 * ld);
      conv[j] = (char)hist[4 * j];
      fail[j] = (char)hist[4 * j + 1];
      if (conv[j]) {
        jend[j] = hist[4 * j + 2];
        last = std::max(last, fail[j] ? jend[j] + 1 : jend[j]);
      } else {
        all = false;
      }
    }
    if (all) return last + 1;
    // m - Kd steps done: finish column s - 1 and v_s
    const int p = Kd + s;
    KCHECK(tacs_hip_bdotq(n, Rn, 1, Kd, Cb(0), s + 1, V(0), N, V(s), NULL, (double*)d_dq, dA, NULL,
                          S, g_stream));
    Ah.assign((size_t)Rn * (p + 1), 0.0);
    d2h_rows(Ah.data(), p + 1, dA, S);
    cf.assign((size_t)Rn * (2 * p + 2), 0.0);
    for (int j = 0; j < Rn; j++) {
      if (conv[j]) continue;
      const int kd = kdj[j];
      const double* a = &Ah[(size_t)j * (p + 1)];
      double ss = 0.0;
      for (int r = 0; r < p; r++) ss += a[r] * a[r];
      const double w2 = a[p] - ss;
      const double om = (w2 > 0.0 && std::isfinite(w2)) ? sqrt(w2) : 0.0;
      std::vector<double>& col = pend[j];
      col.resize((size_t)kd + s + 1);
      for (int r = 0; r < kd + s; r++) col[r] += r < kd ? a[r] : a[Kd + (r - kd)];
      col[kd + s] = om;
      const double st = ls[j].add_column(kd + s - 1, col.data());
      if (!std::isfinite(st)) {
        conv[j] = 1;
        fail[j] = 1;
        jend[j] = s - 1;
        continue;
      }
      double* cj = &cf[(size_t)j * (2 * p + 2)];
      for (int r = 0; r < p; r++) cj[r] = a[r];
      cj[2 * p] = om > 0.0 ? 1.0 / om : 0.0;
    }
    h2d_rows(dcf, S, cf.data(), 2 * p + 2);
    KCHECK(tacs_hip_bdcgs_update(n, Rn, Kd, Cb(0), s, V(0), N, dcf, S, V(s), NULL, g_stream));
    return s;
  };

  if (zero_guess) {
    KCHECK(tacs_hip_fill(64, N, Xd, 0.0, g_stream));
    HCHECK(hipMemcpyAsync(V(0), Bd, N * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
  } else {
    spmm64(Xd, V(0));
    KCHECK(tacs_hip_axpby(64, 64, N, 1.0, Bd, -1.0, V(0), g_stream));
  }
  for (int count = 0;; count++) {
    if (g_halt) return 0;  // a copy or launch failed: no further work
    KCHECK(tacs_hip_bmdot(n, Rn, 1, V(0), N, V(0), (double*)d_red, nrm2, 1, 0, g_stream));
    KCHECK(tacs_hip_bsqrt_inv(Rn, nrm2, 1, nrm, 1, inv, g_stream));
    d2h(h_pin + 256, nrm, Rn);
    int nact = 0;
    for (int j = 0; j < Rn; j++) {
      if (!active[j]) continue;
      double r = h_pin[256 + j];
      if (monitor) monitor->printResidual(0, r);
      if (count == 0) r0[j] = r;
      rn[j] = r;
      if (!std::isfinite(r) || fail[j]) {
        active[j] = 0;
      } else if (r <= rtol * r0[j] || r < atol) {
        active[j] = 0;
        flags[j] = 1;
      } else {
        nact++;
      }
    }
    if (nact == 0 || count > nrestart) break;
    // The cycle's least-squares estimate and the true FP64 residual part when
    // rounding in the operator (about 6e-8 |A| |A^-1| for FP32 values) or in
    // the recycled relation A M^-1 U = C matters at the requested tolerance.
    // A stall: 5 cycles in a row whose estimate reached the target while the
    // true residual did not; it is answered in stages below.
    bool stalled = false;
    for (int j = 0; j < Rn; j++) {
      if (!active[j]) continue;
      stall[j] = est_hit[j] ? stall[j] + 1 : 0;
      if (stall[j] >= 5) stalled = true;
    }
    if (stalled) {
      // first refresh the spaces; a second stall with the FP32-valued
      // operator continues with the FP64 one (spaces refreshed for it); a
      // further stall stops deflating for the rest of the solve: a recycled
      // vector of a small eigenvalue of A M^-1 carries its rounding times
      // |U| ~ 1/lambda into the true residual, which bounds what deflated
      // cycles can reach on very ill-conditioned systems
      if (Kd > 0 && refreshes > 0 && pr == 64) {
        Kd = 0;
        std::fill(kdj.begin(), kdj.end(), 0);
        nodefl = true;
        fell_back = true;
        if (monitor)
          monitor->print("recycling GMRES: stalled, continuing without deflation, classical "
                         "Gram-Schmidt\n");
      } else {
        if (refreshes > 0 && pr == 32) {
          pr = 64;
          Dinv = pc->d_Dinv;
          fell_back = true;
          if (monitor) monitor->print("recycling GMRES: stalled, continuing with the FP64 operator\n");
        } else if (monitor) {
          monitor->print("recycling GMRES: stalled, recycled spaces refreshed\n");
        }
        refresh_space(Kd);
        refreshes++;
      }
      std::fill(stall.begin(), stall.end(), 0);
    }
    cycles++;
    for (int j = 0; j < Rn; j++) {
      h_inv[j] = active[j] ? 1.0 / rn[j] : 0.0;
      target[j] = active[j] ? std::max(rtol * r0[j], atol) / rn[j] : 0.0;
      conv[j] = !active[j];
      jend[j] = 0;
    }
    h2d(inv, h_inv, Rn);
    KCHECK(tacs_hip_bscale(n, Rn, inv, V(0), g_stream));
    double* tmp2 = d_Cnb;  // block scratch (holds the new C only at the end of a cycle)
    if (Kd > 0) {
      // x += |r| M^{-1} U C^T v_0;  v_0 -= C C^T v_0
      KCHECK(tacs_hip_bmdot(n, Rn, Kd, Cb(0), N, V(0), (double*)d_red, pco, S, 0, g_stream));
      KCHECK(tacs_hip_blincomb(n, Rn, Kd, Ub(0), N, pco, S, NULL, d_work, 0, g_stream));
      precond(pr, Rn, Dinv, A32, d_work, tmp2);
      KCHECK(tacs_hip_bscale(n, Rn, nrm, tmp2, g_stream));
      KCHECK(tacs_hip_axpby(64, 64, N, 1.0, tmp2, 1.0, Xd, g_stream));
      KCHECK(tacs_hip_bcgs_update(n, Rn, Kd, Cb(0), N, pco, S, V(0), g_stream));
    }
    // beta_j = |(I - C C^T) v_0(:, j)| into res_j[0]; v_0 normalised
    normalize(V(0), res, S);
    stage.assign(Rn, 0.0);
    d2h_rows(stage.data(), 1, res, S);
    for (int j = 0; j < Rn; j++) {
      beta[j] = stage[j];
      if (conv[j]) continue;
      if (beta[j] <= target[j]) {
        conv[j] = 1;  // the projection alone converged
        continue;
      }
      // least squares over [U D_u, V]: columns 0..k_j-1 are D_u, then Arnoldi
      const int kd = kdj[j];
      HostLS& L = ls[j];
      L.reset();
      L.c[kd] = 1.0;
      L.g[kd] = 1.0;
      for (int i = 0; i < kd; i++) {
        std::fill(hc.begin(), hc.end(), 0.0);
        hc[i] = dsc[(size_t)j * KD_MAX + i];
        L.add_column(i, hc.data());
      }
      tgt[j] = target[j] / beta[j];
    }
    const int s = m - Kd;
    int niters = 0;
    bool all = true;
    for (int j = 0; j < Rn; j++) all = all && conv[j];
    // after repeated stalls the cycles use classical Gram-Schmidt twice with
    // each column finished in its own step (no lagged normalisation)
    if (delayed && !nodefl) {
      niters = arnoldi_delayed(s, Kd);
      all = true;  // skips the classical loop below
    }
    for (int i = 0; i < s && !all; i++) {
      double* q = V(i + 1);
      double* h = H + (size_t)i * (m + 1);
      double* bc = Bc + (size_t)i * KD_MAX;
      op(V(i), q);
      // classical Gram-Schmidt, twice, against [C_j, V_0..V_i]
      for (int pass = 0; pass < 2; pass++) {
        double* cc = pass == 0 ? bc : tmpC;
        double* hq = pass == 0 ? h : tmp;
        if (Kd > 0)
          KCHECK(tacs_hip_bmdot(n, Rn, Kd, Cb(0), N, q, (double*)d_red, cc, S, 0, g_stream));
        KCHECK(tacs_hip_bmdot(n, Rn, i + 1, V(0), N, q, (double*)d_red, hq, S, 0, g_stream));
        if (Kd > 0) KCHECK(tacs_hip_bcgs_update(n, Rn, Kd, Cb(0), N, cc, S, q, g_stream));
        KCHECK(tacs_hip_bcgs_update(n, Rn, i + 1, V(0), N, hq, S, q, g_stream));
      }
      if (Kd > 0) KCHECK(tacs_hip_badd(Rn, Kd, tmpC, S, bc, S, g_stream));
      KCHECK(tacs_hip_badd(Rn, i + 1, tmp, S, h, S, g_stream));
      normalize(q, h + i + 1, S);
      niters = i + 1;
      hb.assign((size_t)Rn * std::max(Kd, 1), 0.0);
      hh.assign((size_t)Rn * (i + 2), 0.0);
      d2h_rows(hb.data(), Kd, bc, S);
      d2h_rows(hh.data(), i + 2, h, S);
      all = true;
      for (int j = 0; j < Rn; j++) {
        if (conv[j]) continue;
        const int kd = kdj[j];
        for (int r = 0; r < kd; r++) hc[r] = hb[(size_t)j * Kd + r];
        for (int r = 0; r < i + 2; r++) hc[kd + r] = hh[(size_t)j * (i + 2) + r];
        double st = ls[j].add_column(kd + i, hc.data());
        if (!std::isfinite(st)) {
          conv[j] = 1;
          fail[j] = 1;
          jend[j] = i;
        } else if (st <= tgt[j]) {
          conv[j] = 1;
          jend[j] = i + 1;
        } else {
          all = false;
        }
      }
    }
    for (int j = 0; j < Rn; j++) {
      // the cycle ended on the least-squares estimate (or the projection)
      // reaching the target; whether the true residual agrees is checked at
      // the next cycle start
      est_hit[j] = active[j] && conv[j] && !fail[j];
      if (!conv[j]) jend[j] = niters;
      if (active[j]) iters[j] += jend[j];
    }
    // y_j from the least squares;  X(:, j) += |r_j| beta_j M^{-1} (U D_u y_u + V y_v)
    if (niters > 0) {
      std::vector<double> cu((size_t)Rn * std::max(Kd, 1), 0.0), cv((size_t)Rn * niters, 0.0);
      std::vector<double> y(m + 1);
      for (int j = 0; j < Rn; j++) {
        h_inv[j] = 0.0;
        if (!active[j] || jend[j] == 0) continue;
        const int kd = kdj[j];
        ls[j].solve(kd + jend[j], y.data());
        for (int i = 0; i < kd; i++) cu[(size_t)j * Kd + i] = dsc[(size_t)j * KD_MAX + i] * y[i];
        for (int i = 0; i < jend[j]; i++) cv[(size_t)j * niters + i] = y[kd + i];
        h_inv[j] = rn[j] * beta[j];
      }
      h2d_rows(zc, S, cu.data(), Kd);
      h2d_rows(res, S, cv.data(), niters);
      h2d(inv, h_inv, Rn);
      if (Kd > 0) KCHECK(tacs_hip_blincomb(n, Rn, Kd, Ub(0), N, zc, S, NULL, d_work, 0, g_stream));
      KCHECK(tacs_hip_blincomb(n, Rn, niters, V(0), N, res, S, NULL, d_work, Kd > 0, g_stream));
      precond(pr, Rn, Dinv, A32, d_work, tmp2);
      KCHECK(tacs_hip_bscale(n, Rn, inv, tmp2, g_stream));
      KCHECK(tacs_hip_axpby(64, 64, N, 1.0, tmp2, 1.0, Xd, g_stream));
    }

    // ---- new recycled spaces from each right-hand side's harmonic Ritz vectors
    std::vector<char> cand(Rn, 0);
    bool any = false;
    for (int j = 0; j < Rn; j++) {
      cand[j] = !nodefl && active[j] && !fail[j] && jend[j] > 0 && kdj[j] + jend[j] > K + 1;
      any = any || cand[j];
    }
    if (any) {
      const int jm = niters;      // V_0..V_jm hold the Arnoldi vectors
      const int p = Kd + jm + 1;  // rows of [C, V_0..V_jm]^T U_b
      std::vector<double> prodh((size_t)Rn * std::max(p * Kd, 1), 0.0);
      if (Kd > 0) {
        // column b of [C, V_0..V_jm]^T U_b at d_prod + b p (per right-hand side)
        KCHECK(tacs_hip_btsgemm(n, Rn, Kd, Cb(0), Kd, Ub(0), N, (double*)d_dq, d_prod, SP, p,
                                g_stream));
        KCHECK(tacs_hip_btsgemm(n, Rn, jm + 1, V(0), Kd, Ub(0), N, (double*)d_dq, d_prod + Kd, SP,
                                p, g_stream));
        d2h_rows(prodh.data(), p * Kd, d_prod, SP);
      }
      std::vector<int> kk(Rn, 0);
      std::vector<std::vector<double>> Qm(Rn), PRu(Rn), PRv(Rn);
      // the systems' harmonic Ritz updates are independent: host threads
      // take them in turn
      auto update_one = [&](int j, std::vector<double>& cvtu) {
        const int kd = kdj[j];
        const int nrw = kd + jend[j] + 1;
        cvtu.assign((size_t)nrw * std::max(kd, 1), 0.0);
        for (int b = 0; b < kd; b++) {
          const double* src = &prodh[(size_t)j * p * Kd + (size_t)b * p];
          for (int r = 0; r < kd; r++) cvtu[(size_t)b * nrw + r] = src[r];
          for (int v = 0; v <= jend[j]; v++) cvtu[(size_t)b * nrw + kd + v] = src[Kd + v];
        }
        kk[j] = recycle_update(ls[j], kd, jend[j], K, &dsc[(size_t)j * KD_MAX], cvtu.data(),
                               Qm[j], PRu[j], PRv[j]);
      };
      std::vector<int> todo;
      for (int j = 0; j < Rn; j++)
        if (cand[j]) todo.push_back(j);
      const int nth = std::min<int>((int)todo.size(), host_threads());
      if (nth <= 1) {
        std::vector<double> cvtu;
        for (int j : todo) update_one(j, cvtu);
      } else {
        std::atomic<int> next(0);
        std::vector<std::thread> pool;
        for (int t = 0; t < nth; t++)
          pool.emplace_back([&]() {
            std::vector<double> cvtu;
            for (int q = next++; q < (int)todo.size(); q = next++) update_one(todo[q], cvtu);
          });
        for (std::thread& th : pool) th.join();
      }
      int newKd = 0;
      for (int j = 0; j < Rn; j++) newKd = std::max(newKd, kk[j] > 0 ? kk[j] : kdj[j]);
      bool changed = false;
      for (int j = 0; j < Rn; j++) changed = changed || kk[j] > 0;
      if (changed && newKd > 0) {
        // per right-hand side P blocks: A (Kd x newKd) and B ((jm+1) x newKd) give
        // C_new = C A + V B;  Cm (jm x newKd) and D (Kd x newKd) give U_new = V Cm + U D
        const size_t oB = (size_t)Kd * newKd, oC = oB + (size_t)(jm + 1) * newKd,
                     oD = oC + (size_t)jm * newKd, tot = oD + (size_t)Kd * newKd;
        std::vector<double> P((size_t)Rn * tot, 0.0);
        for (int j = 0; j < Rn; j++) {
          double* Pj = &P[(size_t)j * tot];
          const int kd = kdj[j];
          if (kk[j] > 0) {
            const int nrw = kd + jend[j] + 1, jn = jend[j];
            for (int c = 0; c < kk[j]; c++) {
              for (int r = 0; r < kd; r++) Pj[r + (size_t)c * Kd] = Qm[j][(size_t)c * nrw + r];
              for (int v = 0; v <= jn; v++)
                Pj[oB + v + (size_t)c * (jm + 1)] = Qm[j][(size_t)c * nrw + kd + v];
              for (int v = 0; v < jn; v++) Pj[oC + v + (size_t)c * jm] = PRv[j][(size_t)c * jn + v];
              for (int r = 0; r < kd; r++) Pj[oD + r + (size_t)c * Kd] = PRu[j][(size_t)c * kd + r];
            }
          } else {
            for (int r = 0; r < kd; r++) {
              Pj[r + (size_t)r * Kd] = 1.0;
              Pj[oD + r + (size_t)r * Kd] = 1.0;
            }
          }
        }
        h2d_rows(d_prod, SP, P.data(), (int)tot);
        if (Kd > 0)
          KCHECK(tacs_hip_bbasis_combine(n, Rn, Kd, Cb(0), N, d_prod, SP, Kd, newKd, d_Cnb, N, 0,
                                         g_stream));
        KCHECK(tacs_hip_bbasis_combine(n, Rn, jm + 1, V(0), N, d_prod + oB, SP, jm + 1, newKd,
                                       d_Cnb, N, Kd > 0, g_stream));
        KCHECK(tacs_hip_bbasis_combine(n, Rn, jm, V(0), N, d_prod + oC, SP, jm, newKd, d_Unb, N,
                                       0, g_stream));
        if (Kd > 0)
          KCHECK(tacs_hip_bbasis_combine(n, Rn, Kd, Ub(0), N, d_prod + oD, SP, Kd, newKd, d_Unb,
                                         N, 1, g_stream));
        HCHECK(hipStreamSynchronize(g_stream));
        std::swap(d_Cb, d_Cnb);
        std::swap(d_Ub, d_Unb);
        for (int j = 0; j < Rn; j++)
          if (kk[j] > 0) kdj[j] = kk[j];
        Kd = newKd;
        update_dscale(Kd);
      }
    }
    spmm64(Xd, V(0));
    KCHECK(tacs_hip_axpby(64, 64, N, 1.0, Bd, -1.0, V(0), g_stream));
  }
  HCHECK(hipStreamSynchronize(g_stream));
  // spaces built with the FP64 operator do not satisfy A32 D32^-1 U = C
  if (fell_back && prec == 32) defVersion = -1;
  int ok = 1;
  for (int j = 0; j < Rn; j++) ok &= flags[j];
  return ok;
}

// ---------------------------------------------------------------------------
// Functions
// ---------------------------------------------------------------------------
TACSFunction::TACSFunction(TACSAssembler* a) : assembler(a) { assembler->incref(); }
TACSFunction::~TACSFunction() { assembler->decref(); }
TACSCompliance::TACSCompliance(TACSAssembler* a) : TACSFunction(a) {}
TACSKSFailure::TACSKSFailure(TACSAssembler* a, double w, double al, double sf)
    : TACSFunction(a), ksWeight(w), alpha(al), safetyFactor(sf) {}
// TACSKSFailure::getFunctionValue src/functions/TACSKSFailure.cpp:74
TacsScalar TACSKSFailure::getFunctionValue() {
  return maxFail + std::log(ksFailSum / alpha) / ksWeight;
}

// ---------------------------------------------------------------------------
// TACSAssembler
// ---------------------------------------------------------------------------
// Number of processes of comm, asked of the MPI the process has loaded and
// initialised (the driver's); 1 when there is none.  The library links no
// MPI and calls nothing else of it.
static int comm_size(TACSComm c) {
  typedef int (*initialized_fn)(int*);
  typedef int (*size_fn)(MPI_Comm, int*);
  initialized_fn mi = (initialized_fn)dlsym(RTLD_DEFAULT, "MPI_Initialized");
  size_fn ms = (size_fn)dlsym(RTLD_DEFAULT, "MPI_Comm_size");
  int flag = 0, n = 1;
  if (!mi || !ms || mi(&flag) != 0 || !flag) return 1;
  if (ms(c, &n) != 0) return 1;
  return n;
}

TACSAssembler::TACSAssembler(TACSComm _comm, int vpn, int nowned, int nelems, int ndep)
    : varsPerNode(vpn), numNodes(nowned), numElements(nelems), npe(0), order(0),
      initialized(0), numDesignVars(0), h_ptr(NULL), h_conn(NULL), elements(NULL),
      auxElements(NULL), aux_version(-1), naux(0), naux_pressure(0), nnzb(0), jac_ms(0.0) {
  if (ndep != 0) tacs_set_error("TACSAssembler: dependent nodes are not supported");
  bcMap = new TACSBcMap(varsPerNode, 64);
  bcMap->incref();
  memset(&mat, 0, sizeof(mat));
  memset(&smat, 0, sizeof(smat));
  family = FAMILY_SOLID;
  d_ke = NULL;
  d_qgeo = NULL;
  d_qstress = NULL;
  d_aux_pres = NULL;
  d_conn = d_n2e_ptr = d_n2e = d_bcvars = d_bcnodes = d_bcmask = NULL;
  d_bcvals = NULL;
  nbc_nodes = 0;
  d_dvnum = d_dv2e_ptr = d_dv2e = NULL;
  d_tfixed = d_telem = d_X = d_u = d_ev = d_es = NULL;
  d_mtab = NULL;
  d_emat = NULL;
  d_aux_elem = d_aux_face = d_n2a_ptr = d_n2a = NULL;
  d_aux_trac = d_aux_ev = NULL;
  d_status = NULL;
  d_red = NULL;
  d_scalar = NULL;
  h_rowp = h_cols = NULL;
  comm = _comm;
  elemPrec = 64;
  jacKernel = 0;
  jacVar = -1;
  d_keb = NULL;
}

TACSAssembler::~TACSAssembler() {
  if (elements) {
    for (int i = 0; i < numElements; i++)
      if (elements[i]) elements[i]->decref();
    delete[] elements;
  }
  delete[] h_ptr;
  delete[] h_conn;
  delete[] h_rowp;
  delete[] h_cols;
  bcMap->decref();
  if (auxElements) auxElements->decref();
  void* ds[] = {d_conn, d_n2e_ptr, d_n2e, d_bcvars, d_bcnodes, d_bcmask, d_bcvals, d_dvnum,
                d_dv2e_ptr, d_dv2e, d_tfixed, d_telem, d_X, d_u, d_ev, d_es, d_aux_elem,
                d_aux_face, d_n2a_ptr, d_n2a, d_aux_trac, d_aux_ev, d_status, d_red, d_scalar,
                d_ke, d_aux_pres, d_qgeo, d_qstress, d_keb, d_mtab, d_emat};
  for (void* p : ds) dfree(p);
}

int TACSAssembler::setElementPrecision(int prec) {
  if (prec != 64 && prec != 32) return tacs_set_error("setElementPrecision: 64 or 32, not %d", prec);
  elemPrec = prec;
  return 0;
}
int TACSAssembler::setJacobianKernel(int kernel) {
  if (kernel < 0 || kernel > 2) return tacs_set_error("setJacobianKernel: 0, 1 or 2, not %d", kernel);
  jacKernel = kernel;
  return 0;
}
// The element-block path needs the isotropic tangent and block indices that
// fit int32; elsewhere (and when asked) the row kernel runs.  jacKernel = 0
// follows the assembler's autotune choice or the tuning table, whose
// built-in defaults are the MI355X timings: 27-node hexahedra (40^3) 5.0 ms
// with element blocks against 27.8 ms with rows; 8-node (128^3) 11.5 ms
// against 8.2 ms, so FP64 8-node elements keep the rows.  FP32 element
// arithmetic exists only as element blocks.
int TACSAssembler::jacobianVariant() {
  const int nv = tacs_hip_solid_jacobian_nvariants(order);
  const int v = jacVar >= 0 ? jacVar : tacs_tune_variant(TACS_TUNE_JACOBIAN, order, numNodes);
  return v > nv ? nv : v;
}
int TACSAssembler::jacobianKernelInUse() {
  if (family != FAMILY_SOLID) return 0;
  const bool fits = mat.isotropic && (double)numElements * npe * npe < 2147483647.0;
  if (jacKernel == 1 || !fits) return 1;
  if (jacKernel == 2 || elemPrec == 32) return 2;
  return jacobianVariant() == tacs_hip_solid_jacobian_nvariants(order) ? 2 : 1;
}
int TACSAssembler::autotune(TACSParallelMat* A, int reps, double* ms, int maxv) {
  if (!initialized || family != FAMILY_SOLID || !mat.isotropic || !A)
    return tacs_set_error("autotune: an initialised assembler of isotropic solid elements and a "
                          "matrix");
  if (jacKernel == 2 || elemPrec == 32) return 0;  // the element-block form only: nothing to time
  const int nv = tacs_hip_solid_jacobian_nvariants(order);
  const bool fits = (double)numElements * npe * npe < 2147483647.0;
  const int nc = nv + (jacKernel == 0 && fits ? 1 : 0);
  if (reps < 1 || !ms || maxv < nc)
    return tacs_set_error("autotune: reps >= 1 and room for %d choices", nc);
  const int saved = jacVar;
  int best = 0;
  for (int c = 0; c < nc && !g_halt; c++) {
    jacVar = c;
    assembleJacobian(1.0, 0.0, 0.0, NULL, A, TACS_MAT_NORMAL, 1.0, 1);
    double t = 0.0;
    for (int r = 0; r < reps && !g_halt; r++) {
      assembleJacobian(1.0, 0.0, 0.0, NULL, A, TACS_MAT_NORMAL, 1.0, 1);
      t += jac_ms;
    }
    ms[c] = t / reps;
    if (ms[c] < ms[best]) best = c;
  }
  jacVar = g_halt ? saved : best;
  return g_halt ? 1 : nc;
}

int TACSAssembler::setElementConnectivity(const int* ptr, const int* conn) {
  if (initialized) return tacs_set_error("setElementConnectivity after initialize()");
  if (numElements < 0 || numElements >= (1 << 26))
    return tacs_set_error("setElementConnectivity: %d elements; the node-to-element packing holds "
                          "fewer than 2^26 = 67108864",
                          numElements);
  delete[] h_ptr;
  delete[] h_conn;
  h_ptr = new int[numElements + 1];
  memcpy(h_ptr, ptr, (numElements + 1) * sizeof(int));
  size_t len = ptr[numElements];
  h_conn = new int[len > 0 ? len : 1];
  memcpy(h_conn, conn, len * sizeof(int));
  for (size_t i = 0; i < len; i++)
    if (conn[i] < 0 || conn[i] >= numNodes)
      return tacs_set_error("setElementConnectivity: node %d out of range", conn[i]);
  return 0;
}

int TACSAssembler::setElements(TACSElement** elems) {
  if (initialized) return tacs_set_error("setElements after initialize()");
  if (!elements) {
    elements = new TACSElement*[numElements];
    memset(elements, 0, numElements * sizeof(TACSElement*));
  }
  for (int i = 0; i < numElements; i++) {
    elems[i]->incref();
    if (elements[i]) elements[i]->decref();
    elements[i] = elems[i];
  }
  return 0;
}

int TACSAssembler::addBCs(int nnodes, const int* nodes, int nbcs, const int* vars,
                          const TacsScalar* vals) {
  // TACSAssembler::addBCs src/TACSAssembler.cpp:801: without vars (or with
  // nbcs < 0) every variable of the node is constrained
  if (initialized) return tacs_set_error("Cannot call addBC() after initialize()");
  if (!vars || nbcs < 0) nbcs = varsPerNode;
  for (int i = 0; i < nnodes; i++) {
    if (nodes[i] < 0 || nodes[i] >= numNodes) return tacs_set_error("addBCs: node out of range");
    bcMap->addBC(nodes[i], nbcs, vars, vals);
  }
  return 0;
}

void TACSAssembler::setAuxElements(TACSAuxElements* aux) {
  if (aux) aux->incref();
  if (auxElements) auxElements->decref();
  auxElements = aux;
  aux_version = -1;
}

int TACSAssembler::initialize() {
  if (initialized) return 0;
  const int nranks = comm_size(comm);
  if (nranks > 1)
    return tacs_set_error("initialize: the communicator has %d processes; this build runs one "
                          "rank on one GPU",
                          nranks);
  if (!h_ptr || !h_conn) return tacs_set_error("initialize: connectivity not set");
  if (!elements) return tacs_set_error("initialize: elements not set");
  std::vector<int> dvnum(numElements);
  std::vector<double> tfixed(numElements);
  std::vector<int> matIndex(numElements, 0);  // per-element materials
  std::vector<ShellMaterial> shellTab;
  std::vector<double> solidProps;
  bool solidMixed = false;
  TACSMaterialProperties* props0 = NULL;
  if (numElements > 0 && dynamic_cast<TACSQuad4Shell*>(elements[0])) {
    // Shell family: TACSQuad4Shell + TACSShellNaturalTransform or
    // TACSShellRefAxisTransform + TACSIsoShellConstitutive, one material,
    // kcorr and tOffset.
    family = FAMILY_SHELL;
    if (varsPerNode != 6) return tacs_set_error("initialize: shells need 6 variables per node");
    TACSShellTransform* tr0 = NULL;
    for (int e = 0; e < numElements; e++) {
      TACSQuad4Shell* el = dynamic_cast<TACSQuad4Shell*>(elements[e]);
      if (!el) return tacs_set_error("initialize: element %d is not a TACSQuad4Shell", e);
      TACSShellRefAxisTransform* ra = dynamic_cast<TACSShellRefAxisTransform*>(el->transform);
      if (!dynamic_cast<TACSShellNaturalTransform*>(el->transform) && !ra)
        return tacs_set_error("initialize: the shell transform must be TACSShellNaturalTransform "
                              "or TACSShellRefAxisTransform");
      // one frame per assembler: one transform object, or transforms of one kind
      // and axis
      if (!tr0) tr0 = el->transform;
      if (el->transform != tr0) {
        TACSShellRefAxisTransform* ra0 = dynamic_cast<TACSShellRefAxisTransform*>(tr0);
        const bool same = (ra && ra0) ? (ra->axis[0] == ra0->axis[0] && ra->axis[1] == ra0->axis[1] &&
                                         ra->axis[2] == ra0->axis[2])
                                      : (!ra && !ra0);
        if (!same) return tacs_set_error("initialize: elements must share one shell transform");
      }
      TACSIsoShellConstitutive* con = dynamic_cast<TACSIsoShellConstitutive*>(el->con);
      if (!con) return tacs_set_error("initialize: element %d constitutive is not isotropic", e);
      TACSMaterialProperties* p = con->props;
      if (!props0) props0 = p;
      // the element's material: plane-stress Q (TACSMaterialProperties::
      // evalTangentStiffness2D :323), kcorr, drilling regularization
      // (TACSShellConstitutive::DRILLING_REGULARIZATION), ys, rho, tOffset
      ShellMaterial me = {};
      const double D = p->E / (1.0 - p->nu * p->nu);
      me.Q[0] = D;
      me.Q[1] = p->nu * D;
      me.Q[3] = D;
      me.Q[5] = p->G;
      me.kcorr = con->kcorr;
      me.drill_reg = 0.1;
      me.ys = p->ys;
      me.rho = p->rho;
      me.toff = con->tOffset;
      size_t k = 0;
      for (; k < shellTab.size(); k++) {
        const ShellMaterial& o = shellTab[k];
        if (o.Q[0] == me.Q[0] && o.Q[1] == me.Q[1] && o.Q[5] == me.Q[5] && o.kcorr == me.kcorr &&
            o.ys == me.ys && o.rho == me.rho && o.toff == me.toff)
          break;
      }
      if (k == shellTab.size()) shellTab.push_back(me);
      matIndex[e] = (int)k;
      dvnum[e] = con->tNum;
      tfixed[e] = con->t;
    }
    order = 2;
    npe = 4;
    smat = shellTab[0];
    TACSShellRefAxisTransform* ra0 = dynamic_cast<TACSShellRefAxisTransform*>(tr0);
    smat.ref_axis = ra0 ? 1 : 0;
    for (int k = 0; k < 3; k++) smat.axis[k] = ra0 ? ra0->axis[k] : 0.0;
  } else {
    // Solid family: TACSElement3D + TACSLinearElasticity3D +
    // TACSSolidConstitutive on one hexahedral basis and one material.
    family = FAMILY_SOLID;
    if (varsPerNode != 3) return tacs_set_error("initialize: solids need 3 variables per node");
    solidProps.resize(4 * (size_t)numElements);
    for (int e = 0; e < numElements; e++) {
      TACSElement3D* el = dynamic_cast<TACSElement3D*>(elements[e]);
      if (!el) return tacs_set_error("initialize: element %d is not a TACSElement3D", e);
      TACSLinearElasticity3D* model = dynamic_cast<TACSLinearElasticity3D*>(el->model);
      if (!model)
        return tacs_set_error("initialize: element %d model is not TACSLinearElasticity3D", e);
      if (model->strain_type != TACS_LINEAR_STRAIN)
        return tacs_set_error("initialize: only TACS_LINEAR_STRAIN is supported");
      int o = el->basis->getOrder();
      if (order == 0) order = o;
      if (o != order) return tacs_set_error("initialize: mixed element bases are not supported");
      TACSMaterialProperties* p = model->stiff->props;
      if (!props0) props0 = p;
      // the element's isotropic constants (lam, mu, rho, ys)
      double C[21];
      p->evalTangentStiffness3D(C);
      solidProps[4 * (size_t)e] = C[1];
      solidProps[4 * (size_t)e + 1] = p->G;
      solidProps[4 * (size_t)e + 2] = p->rho;
      solidProps[4 * (size_t)e + 3] = p->ys;
      if (p != props0 && (p->E != props0->E || p->nu != props0->nu || p->rho != props0->rho ||
                          p->ys != props0->ys))
        solidMixed = true;
      dvnum[e] = model->stiff->tNum;
      tfixed[e] = model->stiff->t;
    }
    npe = order * order * order;
    props0->evalTangentStiffness3D(mat.C);
    mat.rho = props0->rho;
    mat.ys = props0->ys;
    mat.isotropic = 1;
    mat.mu = props0->G;
    mat.lam = mat.C[1];  // nu*E/((1+nu)(1-2nu))
  }
  if (!props0) return tacs_set_error("initialize: no elements");
  for (int e = 0; e < numElements; e++)
    if (h_ptr[e + 1] - h_ptr[e] != npe)
      return tacs_set_error("initialize: element %d has %d nodes, expected %d", e,
                            h_ptr[e + 1] - h_ptr[e], npe);

  // node -> element CSR (TACSAssembler::computeNodeToElementCSR :1791),
  // entries (element << 5 | local index), ascending element order: the
  // packing holds element indices below 2^26 (67,108,864 elements) and
  // int32 offsets hold at most 2^31 - 1 (element, node) incidences
  if (numElements >= (1 << 26))
    return tacs_set_error("initialize: %d elements; the node-to-element packing holds fewer than "
                          "2^26 = 67108864",
                          numElements);
  if ((long)numElements * npe > 0x7fffffffL)
    return tacs_set_error("initialize: %ld element-node incidences exceed int32 offsets",
                          (long)numElements * npe);
  std::vector<int> n2e_ptr(numNodes + 1, 0);
  for (int e = 0; e < numElements; e++)
    for (int a = 0; a < npe; a++) n2e_ptr[h_conn[(size_t)e * npe + a] + 1]++;
  for (int i = 0; i < numNodes; i++) n2e_ptr[i + 1] += n2e_ptr[i];
  std::vector<int> n2e(n2e_ptr[numNodes]);
  {
    std::vector<int> pos(n2e_ptr.begin(), n2e_ptr.end() - 1);
    for (int e = 0; e < numElements; e++)
      for (int a = 0; a < npe; a++) n2e[pos[h_conn[(size_t)e * npe + a]]++] = (e << 5) | a;
  }
  // node -> node CSR (computeLocalNodeToNodeCSR :1899): sorted, unique,
  // diagonal kept
  h_rowp = new int[numNodes + 1];
  std::vector<int> cols;
  cols.reserve((size_t)numNodes * (family == FAMILY_SHELL ? 9 : (order == 2 ? 27 : 64)));
  h_rowp[0] = 0;
  std::vector<int> tmp;
  for (int i = 0; i < numNodes; i++) {
    tmp.clear();
    for (int p = n2e_ptr[i]; p < n2e_ptr[i + 1]; p++) {
      int e = n2e[p] >> 5;
      for (int a = 0; a < npe; a++) tmp.push_back(h_conn[(size_t)e * npe + a]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    cols.insert(cols.end(), tmp.begin(), tmp.end());
    if (cols.size() > 0x7fffffffL) return tacs_set_error("initialize: too many nonzero blocks");
    h_rowp[i + 1] = (int)cols.size();
  }
  nnzb = (long)cols.size();
  h_cols = new int[nnzb > 0 ? nnzb : 1];
  memcpy(h_cols, cols.data(), nnzb * sizeof(int));
  int maxrow = 0;
  for (int i = 0; i < numNodes; i++) maxrow = std::max(maxrow, h_rowp[i + 1] - h_rowp[i]);
  int maxc = (family == FAMILY_SHELL) ? (1 << 30) : ((order == 2) ? 64 : 125);
  if (maxrow > maxc)
    return tacs_set_error("initialize: a node couples to %d nodes (limit %d)", maxrow, maxc);

  // Boundary conditions: merged per node (mask OR; later values win)
  {
    const int *bn, *bv;
    TacsScalar* bval;
    int nbc = bcMap->getBCs(&bn, &bv, &bval);
    std::vector<int> mask(numNodes, 0), order_nodes;
    const int b = varsPerNode;
    std::vector<double> vals((size_t)numNodes * b, 0.0);
    for (int q = 0; q < nbc; q++) {
      int nd = bn[q];
      if (!mask[nd] && bv[q]) order_nodes.push_back(nd);
      for (int k = 0; k < b; k++)
        if (bv[q] & (1 << k)) vals[(size_t)b * nd + k] = bval[(size_t)b * q + k];
      mask[nd] |= bv[q];
    }
    std::sort(order_nodes.begin(), order_nodes.end());
    nbc_nodes = (int)order_nodes.size();
    std::vector<int> masks(nbc_nodes);
    std::vector<double> bvals((size_t)nbc_nodes * b);
    for (int q = 0; q < nbc_nodes; q++) {
      masks[q] = mask[order_nodes[q]];
      for (int k = 0; k < b; k++) bvals[(size_t)b * q + k] = vals[(size_t)b * order_nodes[q] + k];
    }
    d_bcvars = dalloc<int>(numNodes);
    h2d(d_bcvars, mask.data(), numNodes);
    d_bcnodes = dalloc<int>(nbc_nodes);
    d_bcmask = dalloc<int>(nbc_nodes);
    d_bcvals = dalloc<double>((size_t)b * nbc_nodes);
    h2d(d_bcnodes, order_nodes.data(), nbc_nodes);
    h2d(d_bcmask, masks.data(), nbc_nodes);
    h2d(d_bcvals, bvals.data(), (size_t)b * nbc_nodes);
    HCHECK(hipStreamSynchronize(g_stream));
  }

  // Design variables: one per element constitutive (tNum), dv -> elements CSR
  int maxdv = -1;
  for (int e = 0; e < numElements; e++) maxdv = std::max(maxdv, dvnum[e]);
  numDesignVars = maxdv + 1;
  std::vector<int> dv2e_ptr(numDesignVars + 1, 0), dv2e;
  for (int e = 0; e < numElements; e++)
    if (dvnum[e] >= 0) dv2e_ptr[dvnum[e] + 1]++;
  for (int d = 0; d < numDesignVars; d++) dv2e_ptr[d + 1] += dv2e_ptr[d];
  dv2e.resize(dv2e_ptr[numDesignVars]);
  {
    std::vector<int> pos(dv2e_ptr.begin(), dv2e_ptr.end() - 1);
    for (int e = 0; e < numElements; e++)
      if (dvnum[e] >= 0) dv2e[pos[dvnum[e]]++] = e;
  }

  // Upload
  size_t clen = (size_t)numElements * npe;
  d_conn = dalloc<int>(clen);
  h2d(d_conn, h_conn, clen);
  d_n2e_ptr = dalloc<int>(numNodes + 1);
  d_n2e = dalloc<int>(n2e.size());
  h2d(d_n2e_ptr, n2e_ptr.data(), numNodes + 1);
  h2d(d_n2e, n2e.data(), n2e.size());
  d_dvnum = dalloc<int>(numElements);
  d_tfixed = dalloc<double>(numElements);
  d_telem = dalloc<double>(numElements);
  h2d(d_dvnum, dvnum.data(), numElements);
  h2d(d_tfixed, tfixed.data(), numElements);
  h2d(d_telem, tfixed.data(), numElements);
  if (family == FAMILY_SOLID && solidMixed) {
    double* tab = dalloc<double>(solidProps.size());
    h2d(tab, solidProps.data(), solidProps.size());
    d_mtab = tab;
    mat.elem_props = tab;
  }
  if (family == FAMILY_SHELL && shellTab.size() > 1) {
    ShellMaterial* tab = dalloc<ShellMaterial>(shellTab.size());
    d_emat = dalloc<int>(numElements);
    h2d(tab, shellTab.data(), shellTab.size());
    h2d(d_emat, matIndex.data(), numElements);
    d_mtab = tab;
    smat.table = tab;
    smat.elem_mat = d_emat;
  }
  d_dv2e_ptr = dalloc<int>(numDesignVars + 1);
  d_dv2e = dalloc<int>(dv2e.size());
  h2d(d_dv2e_ptr, dv2e_ptr.data(), numDesignVars + 1);
  h2d(d_dv2e, dv2e.data(), dv2e.size());
  d_X = dalloc<double>((size_t)3 * numNodes);
  d_u = dalloc<double>((size_t)varsPerNode * numNodes);
  HCHECK(hipMemsetAsync(d_X, 0, (size_t)3 * numNodes * sizeof(double), g_stream));
  HCHECK(hipMemsetAsync(d_u, 0, (size_t)varsPerNode * numNodes * sizeof(double), g_stream));
  d_status = dalloc<int>(1);
  HCHECK(hipMemsetAsync(d_status, 0, sizeof(int), g_stream));
  size_t rb = tacs_hip_red_scratch_bytes();
  d_red = dalloc<char>(rb);
  HCHECK(hipMemsetAsync(d_red, 0, rb, g_stream));
  d_scalar = dalloc<double>(16);
  HCHECK(hipStreamSynchronize(g_stream));
  if (g_halt) return 1;  // an upload failed: stays uninitialised (error set)
  initialized = 1;
  return 0;
}

int TACSAssembler::status() {
  int s = 0;
  d2h(&s, d_status, 1);
  if (s) HCHECK(hipMemsetAsync(d_status, 0, sizeof(int), g_stream));
  return s;
}

TACSBVec* TACSAssembler::createVec() { return new TACSBVec(numNodes, varsPerNode); }
TACSBVec* TACSAssembler::createNodeVec() { return new TACSBVec(numNodes, 3); }
TACSBVec* TACSAssembler::createDesignVec() { return new TACSBVec(numDesignVars, 1); }
TACSParallelMat* TACSAssembler::createMat() {
  if (!initialized) {
    tacs_set_error("Cannot call createMat() before initialize()");
    return NULL;
  }
  return new TACSParallelMat(varsPerNode, numNodes, h_rowp, h_cols);
}

void TACSAssembler::setNodes(TACSBVec* X) {
  HCHECK(hipMemcpyAsync(d_X, X->devRead(), (size_t)3 * numNodes * sizeof(double),
                        hipMemcpyDeviceToDevice, g_stream));
}
void TACSAssembler::getNodes(TACSBVec* X) {
  HCHECK(hipMemcpyAsync(X->devWrite(), d_X, (size_t)3 * numNodes * sizeof(double),
                        hipMemcpyDeviceToDevice, g_stream));
}
void TACSAssembler::setVariables(TACSBVec* q, TACSBVec*, TACSBVec*) {
  HCHECK(hipMemcpyAsync(d_u, q->devRead(), (size_t)varsPerNode * numNodes * sizeof(double),
                        hipMemcpyDeviceToDevice, g_stream));
}
void TACSAssembler::getVariables(TACSBVec* q, TACSBVec*, TACSBVec*) {
  HCHECK(hipMemcpyAsync(q->devWrite(), d_u, (size_t)varsPerNode * numNodes * sizeof(double),
                        hipMemcpyDeviceToDevice, g_stream));
}
void TACSAssembler::zeroVariables() {
  KCHECK(tacs_hip_fill(64, (size_t)varsPerNode * numNodes, d_u, 0.0, g_stream));
}
void TACSAssembler::setDesignVars(TACSBVec* dvs) {
  KCHECK(tacs_hip_gather_design_vars(numElements, d_dvnum, d_tfixed,
                                     numDesignVars > 0 ? dvs->devRead() : NULL, d_telem, g_stream));
}
void TACSAssembler::getDesignVars(TACSBVec* dvs) {
  std::vector<double> t(numElements);
  d2h(t.data(), d_telem, numElements);
  std::vector<int> dvn(numElements);
  d2h(dvn.data(), d_dvnum, numElements);
  double* h;
  dvs->getArray(&h);
  for (int e = numElements - 1; e >= 0; e--)
    if (dvn[e] >= 0) h[dvn[e]] = t[e];
}
void TACSAssembler::applyBCs(TACSVec* vec) {
  TACSBVec* v = dynamic_cast<TACSBVec*>(vec);
  KCHECK(tacs_hip_apply_bc_vec(varsPerNode, nbc_nodes, d_bcnodes, d_bcmask, NULL, NULL, 1.0,
                               v->devWrite(), g_stream));
}
void TACSAssembler::applyBCs(TACSMat* m) { m->applyBCs(bcMap); }

int TACSAssembler::uploadAux() {
  if (!auxElements) {
    naux = naux_pressure = 0;
    return 0;
  }
  if (aux_version == auxElements->version) return 0;
  // sort aux elements by element index (TACSAuxElements::sort), stable
  int n = auxElements->getNumAux();
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int a, int b) { return auxElements->nums[a] < auxElements->nums[b]; });
  std::vector<int> ae(n), af(n);
  std::vector<double> at((size_t)3 * n, 0.0), ap(n, 0.0);
  int npres = 0;
  for (int q = 0; q < n; q++) {
    if (family == FAMILY_SHELL) {
      TACSShellPressure* pr = dynamic_cast<TACSShellPressure*>(auxElements->elems[idx[q]]);
      if (!pr) return tacs_set_error("aux element %d is not a TACSShellPressure", idx[q]);
      ae[q] = auxElements->nums[idx[q]];
      af[q] = 0;
      ap[q] = pr->p;
      continue;
    }
    if (TACSPressure3D* pr = dynamic_cast<TACSPressure3D*>(auxElements->elems[idx[q]])) {
      if (pr->basis->getOrder() != order) return tacs_set_error("aux element basis mismatch");
      if (pr->faceIndex < 0 || pr->faceIndex > 5)
        return tacs_set_error("aux element %d: face %d out of range", idx[q], pr->faceIndex);
      ae[q] = auxElements->nums[idx[q]];
      af[q] = pr->faceIndex;
      ap[q] = pr->p;
      npres++;
      continue;
    }
    TACSTraction3D* tr = dynamic_cast<TACSTraction3D*>(auxElements->elems[idx[q]]);
    if (!tr) return tacs_set_error("aux element %d is not a TACSTraction3D or TACSPressure3D", idx[q]);
    if (!tr->coordComponent)
      return tacs_set_error("aux element %d: only coordinate-component tractions supported", idx[q]);
    if (tr->basis->getOrder() != order) return tacs_set_error("aux element basis mismatch");
    ae[q] = auxElements->nums[idx[q]];
    af[q] = tr->faceIndex;
    for (int k = 0; k < 3; k++) at[3 * q + k] = tr->trac[k];
  }
  // node -> (aux, local node) CSR, ascending aux order
  std::vector<int> ptr(numNodes + 1, 0);
  for (int q = 0; q < n; q++)
    for (int a = 0; a < npe; a++) ptr[h_conn[(size_t)ae[q] * npe + a] + 1]++;
  for (int i = 0; i < numNodes; i++) ptr[i + 1] += ptr[i];
  std::vector<int> list(ptr[numNodes]);
  std::vector<int> pos(ptr.begin(), ptr.end() - 1);
  for (int q = 0; q < n; q++)
    for (int a = 0; a < npe; a++) list[pos[h_conn[(size_t)ae[q] * npe + a]]++] = (q << 5) | a;
  void* old[] = {d_aux_elem, d_aux_face, d_aux_trac, d_n2a_ptr, d_n2a, d_aux_ev, d_aux_pres};
  HCHECK(hipStreamSynchronize(g_stream));
  for (void* p : old) dfree(p);
  d_aux_elem = dalloc<int>(n);
  d_aux_face = dalloc<int>(n);
  d_aux_trac = dalloc<double>((size_t)3 * n);
  d_n2a_ptr = dalloc<int>(numNodes + 1);
  d_n2a = dalloc<int>(list.size());
  d_aux_ev = dalloc<double>((size_t)n * varsPerNode * npe);
  d_aux_pres = dalloc<double>(n);
  h2d(d_aux_pres, ap.data(), n);
  h2d(d_aux_elem, ae.data(), n);
  h2d(d_aux_face, af.data(), n);
  h2d(d_aux_trac, at.data(), (size_t)3 * n);
  h2d(d_n2a_ptr, ptr.data(), numNodes + 1);
  h2d(d_n2a, list.data(), list.size());
  HCHECK(hipStreamSynchronize(g_stream));
  if (g_halt) {  // no partial aux state survives a failed upload
    void* fresh[] = {d_aux_elem, d_aux_face, d_aux_trac, d_n2a_ptr, d_n2a, d_aux_ev, d_aux_pres};
    for (void* p : fresh) dfree(p);
    d_aux_elem = d_aux_face = d_n2a_ptr = d_n2a = NULL;
    d_aux_trac = d_aux_ev = d_aux_pres = NULL;
    naux = naux_pressure = 0;
    aux_version = -1;
    return 1;
  }
  naux = n;
  naux_pressure = family == FAMILY_SHELL ? 0 : npres;
  aux_version = auxElements->version;
  return 0;
}

int TACSAssembler::elementVectorsToNodes(const double* ev, double* out, int accumulate) {
  if (family == FAMILY_SHELL)
    return tacs_hip_gather_elem_vectors_b(6, 24, numNodes, d_n2e_ptr, d_n2e, ev, accumulate, out,
                                          g_stream);
  return tacs_hip_gather_elem_vectors(3 * npe, numNodes, d_n2e_ptr, d_n2e, ev, accumulate, out,
                                      g_stream);
}

// Shell element kernel on the assembler's state (mode numbers of
// tacs_hip_shell_elem_kernel)
int TACSAssembler::shellKernel(int prec, int mode, const double* psi, double scale,
                               double ks_weight, double ks_max, double ks_sum, double safety,
                               void* ke, double* ev, double* es) {
  return tacs_hip_shell_elem_kernel(prec, mode, numElements, d_conn, d_X, d_u, psi, d_telem,
                                    &smat, scale, ks_weight, ks_max, ks_sum, safety, ke, ev, es,
                                    g_stream);
}

int TACSAssembler::scatterElementScalarsToDVs(const double* es, double* dfdx) {
  return tacs_hip_scatter_dv_sens(numDesignVars, d_dv2e_ptr, d_dv2e, es, dfdx, g_stream);
}

static double* ensure(double*& p, size_t n) {
  if (!p) p = dalloc<double>(n);
  return p;
}

// TACSAssembler::assembleRes src/TACSAssembler.cpp:4133
void TACSAssembler::assembleRes(TACSBVec* residual, const TacsScalar lambda, const bool bcs) {
  if (!initialized) {
    tacs_set_error("assembleRes before initialize()");
    return;
  }
  double* r = residual->devWrite();
  ensure(d_ev, (size_t)numElements * varsPerNode * npe);
  if (family == FAMILY_SHELL)
    KCHECK(shellKernel(64, 1, NULL, 1.0, 0, 0, 1, 1, NULL, d_ev, NULL));
  else
    KCHECK(tacs_hip_solid_elem_vectors(order, 0, numElements, d_conn, d_X, d_u, d_telem, &mat,
                                       1.0, 0, 0, 1, 1, d_ev, g_stream));
  KCHECK(elementVectorsToNodes(d_ev, r, 0));
  addAuxResidual(r, lambda);
  if (bcs)
    KCHECK(tacs_hip_apply_bc_vec(varsPerNode, nbc_nodes, d_bcnodes, d_bcmask, d_bcvals, d_u,
                                 lambda, r, g_stream));
}

// Aux-element loads added to a residual (TACSAssembler.cpp:4210-4225)
void TACSAssembler::addAuxResidual(double* r, TacsScalar lambda) {
  if (uploadAux() != 0 || naux == 0) return;
  if (family == FAMILY_SHELL) {
    KCHECK(tacs_hip_shell_pressure_vectors(naux, d_aux_elem, d_aux_pres, d_conn, d_X, lambda,
                                           d_aux_ev, g_stream));
    KCHECK(tacs_hip_gather_elem_vectors_b(6, 24, numNodes, d_n2a_ptr, d_n2a, d_aux_ev, 1, r,
                                          g_stream));
    return;
  }
  {
    KCHECK(tacs_hip_solid_traction_vectors(order, naux, d_aux_elem, d_aux_face, d_aux_trac,
                                           naux_pressure ? d_aux_pres : NULL, d_conn, d_X, lambda,
                                           d_aux_ev, g_stream));
    KCHECK(tacs_hip_gather_elem_vectors(3 * npe, numNodes, d_n2a_ptr, d_n2a, d_aux_ev, 1, r,
                                        g_stream));
  }
}

// TACSAssembler::assembleJacobian src/TACSAssembler.cpp:4291.  Element
// matrices are symmetric, so TACS_MAT_TRANSPOSE assembles the same values.
void TACSAssembler::assembleJacobian(TacsScalar alpha, TacsScalar beta, TacsScalar gamma,
                                     TACSBVec* residual, TACSMat* tA, MatrixOrientation matOr,
                                     const TacsScalar lambda, const bool bcs) {
  if (!initialized) {
    tacs_set_error("assembleJacobian before initialize()");
    return;
  }
  TACSParallelMat* A = dynamic_cast<TACSParallelMat*>(tA);
  if (!A || A->nrows != numNodes || A->nnzb != nnzb || A->bsize != varsPerNode) {
    tacs_set_error("assembleJacobian: matrix does not match the assembler");
    return;
  }
  hipEvent_t e0, e1;
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  if (family == FAMILY_SHELL) {
    // element matrices (and residuals K u) in one pass, then the row gather
    // (FP32 element matrices: the FP64 residual is a pass of its own)
    ensure(d_ke, (size_t)numElements * 576);
    double* ev = NULL;
    if (residual) ev = ensure(d_ev, (size_t)numElements * 24);
    if (residual && elemPrec == 32)
      KCHECK(shellKernel(64, 1, NULL, 1.0, 0, 0, 1, 1, NULL, ev, NULL));
    HCHECK(hipEventRecord(e0, g_stream));
    KCHECK(tacs_hip_shell_mat_type(elemPrec, numElements, d_conn, d_X, d_u, d_telem, &smat, alpha,
                                   gamma, 0.0, d_ke, elemPrec == 64 ? ev : NULL, g_stream));
    KCHECK(tacs_hip_gather_elem_matrices(6, 4, elemPrec, numNodes, d_n2e_ptr, d_n2e, d_conn, d_ke,
                                         A->d_rowp, A->d_cols, bcs ? d_bcvars : NULL, A->d_A,
                                         A->d_A32, g_stream));
    HCHECK(hipEventRecord(e1, g_stream));
    if (residual) {
      double* r = residual->devWrite();
      KCHECK(elementVectorsToNodes(d_ev, r, 0));
      addAuxResidual(r, lambda);
      if (bcs)
        KCHECK(tacs_hip_apply_bc_vec(varsPerNode, nbc_nodes, d_bcnodes, d_bcmask, d_bcvals, d_u,
                                     lambda, r, g_stream));
    }
    HCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HCHECK(hipEventElapsedTime(&ms, e0, e1));
    jac_ms = ms;
    HCHECK(hipEventDestroy(e0));
    HCHECK(hipEventDestroy(e1));
    A->A32_current = A->d_A32 ? 1 : 0;
    A->version++;
    return;
  }
  if (residual) assembleRes(residual, lambda, bcs);
  const int kern = jacobianKernelInUse();
  if (elemPrec == 32 && kern != 2) {
    HCHECK(hipEventDestroy(e0));
    HCHECK(hipEventDestroy(e1));
    tacs_set_error("assembleJacobian: FP32 element kernels need the element-block kernel "
                   "(isotropic material)");
    return;
  }
  if (kern == 2) {
    // element blocks in the element precision, then the row gather
    if (!d_keb) d_keb = dalloc<char>((size_t)numElements * npe * npe * 9 * sizeof(double));
    HCHECK(hipEventRecord(e0, g_stream));
    KCHECK(tacs_hip_solid_elem_blocks(order, elemPrec, numElements, d_conn, d_X, d_telem, &mat,
                                      alpha, gamma, d_keb, g_stream));
    KCHECK(tacs_hip_gather_solid_blocks(order, elemPrec, numNodes, A->d_rowp, A->d_cols,
                                        d_n2e_ptr, d_n2e, d_conn, d_keb, bcs ? d_bcvars : NULL,
                                        A->d_A, A->d_A32, d_status, g_stream));
  } else {
    HCHECK(hipEventRecord(e0, g_stream));
    ensure(d_qgeo, (size_t)numElements * npe * 10);
    int v = jacobianVariant();
    if (v >= tacs_hip_solid_jacobian_nvariants(order)) v = 0;
    KCHECK(tacs_hip_solid_jacobian_variant(v, order, numNodes, numElements, d_n2e_ptr, d_n2e,
                                           d_conn, d_X, d_telem, &mat, alpha, gamma, A->d_rowp,
                                           A->d_cols, bcs ? d_bcvars : NULL, d_qgeo, A->d_A,
                                           A->d_A32, d_status, g_stream));
  }
  HCHECK(hipEventRecord(e1, g_stream));
  HCHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HCHECK(hipEventElapsedTime(&ms, e0, e1));
  jac_ms = ms;
  HCHECK(hipEventDestroy(e0));
  HCHECK(hipEventDestroy(e1));
  A->A32_current = A->d_A32 ? 1 : 0;
  A->version++;
  int s = status();
  if (s) tacs_set_error("assembleJacobian: kernel status %d", s);
}

// TACSAssembler::assembleMatType src/TACSAssembler.cpp:4418 and
// assembleMatCombo :4519.  Solids: one pass of the row kernel with the summed
// coefficients of the stiffness, the consistent mass and the geometric
// stiffness at the current state; aux elements (tractions, pressures) add no
// matrix terms (TACSElement::getMatType src/elements/TACSElement.cpp:334 over
// their zero Jacobians).  Shells: the stiffness, the consistent mass and the
// geometric stiffness along the current state.
void TACSAssembler::assembleMatCombo(ElementMatrixType matTypes[], TacsScalar scale[], int nmats,
                                     TACSMat* tA, MatrixOrientation matOr, const TacsScalar lambda,
                                     const bool bcs) {
  if (!initialized) {
    tacs_set_error("assembleMatCombo before initialize()");
    return;
  }
  TACSParallelMat* A = dynamic_cast<TACSParallelMat*>(tA);
  if (!A || A->nrows != numNodes || A->nnzb != nnzb || A->bsize != varsPerNode) {
    tacs_set_error("assembleMatType: matrix does not match the assembler");
    return;
  }
  double kc = 0.0, mc = 0.0, gc = 0.0;
  for (int i = 0; i < nmats; i++) {
    switch (matTypes[i]) {
      case TACS_STIFFNESS_MATRIX:
        kc += scale[i];
        break;
      case TACS_MASS_MATRIX:
        mc += scale[i];
        break;
      case TACS_GEOMETRIC_STIFFNESS_MATRIX:
        gc += scale[i];
        break;
      default:
        tacs_set_error("assembleMatType: matrix type %d is not supported", (int)matTypes[i]);
        return;
    }
  }
  if (family == FAMILY_SHELL) {
    ensure(d_ke, (size_t)numElements * 576);
    KCHECK(tacs_hip_shell_mat_type(64, numElements, d_conn, d_X, gc != 0.0 ? d_u : NULL, d_telem,
                                   &smat, kc, mc, gc, d_ke, NULL, g_stream));
    KCHECK(tacs_hip_gather_elem_matrices(6, 4, 64, numNodes, d_n2e_ptr, d_n2e, d_conn, d_ke,
                                         A->d_rowp, A->d_cols, bcs ? d_bcvars : NULL, A->d_A,
                                         A->d_A32, g_stream));
  } else {
    ensure(d_qgeo, (size_t)numElements * npe * 10);
    if (gc != 0.0) ensure(d_qstress, (size_t)numElements * npe * 6);
    KCHECK(tacs_hip_solid_mat_type(order, numNodes, numElements, d_n2e_ptr, d_n2e, d_conn, d_X,
                                   d_u, d_telem, &mat, kc, mc, gc, A->d_rowp, A->d_cols,
                                   bcs ? d_bcvars : NULL, d_qgeo, gc != 0.0 ? d_qstress : NULL,
                                   A->d_A, A->d_A32, d_status, g_stream));
  }
  A->A32_current = A->d_A32 ? 1 : 0;
  A->version++;
  int s = status();
  if (s) tacs_set_error("assembleMatType: kernel status %d", s);
}
void TACSAssembler::assembleMatType(ElementMatrixType matType, TACSMat* A,
                                    MatrixOrientation matOr, const TacsScalar lambda,
                                    const bool bcs) {
  ElementMatrixType t[1] = {matType};
  TacsScalar sc[1] = {1.0};
  assembleMatCombo(t, sc, 1, A, matOr, lambda, bcs);
}

// ---------------------------------------------------------------------------
// Matrix-free products, TACSAssembler.cpp:5502-5717 and
// src/TACSMatrixFreeMat.cpp.  The reference stores per quadrature point the
// inverse Jacobian and the 84 weak-form Jacobian entries of
// TACSLinearElasticity3D (TACSElement3D.cpp:362-425); here the material
// enters the kernel as (lambda, mu) or C[21], so a point keeps the inverse
// Jacobian and w det(J) t (10 values), and the coefficients of the stiffness
// and mass terms head the array.
// ---------------------------------------------------------------------------
static int matfree_coefs(ElementMatrixType matType, double alpha, double gamma, double* kc,
                         double* mc) {
  // TACSLinearElasticity3D::getWeakMatrixNonzeros (TACSLinearElasticity.cpp:1137)
  // with TACSElementBasis::scaleWeakMatrix (basis/TACSElementBasis.cpp:840):
  // stiffness pairs scale by alpha, acceleration pairs by gamma
  switch (matType) {
    case TACS_JACOBIAN_MATRIX:
      *kc = alpha;
      *mc = gamma;
      return 0;
    case TACS_STIFFNESS_MATRIX:
      *kc = alpha;
      *mc = 0.0;
      return 0;
    case TACS_MASS_MATRIX:
      *kc = 0.0;
      *mc = gamma;
      return 0;
    default:
      return 1;
  }
}

void TACSAssembler::getMatrixFreeDataSize(ElementMatrixType matType, int* ds, int* ts) {
  long d = 0, t = 0;
  if (family == FAMILY_SOLID && initialized) {
    d = 4 + 10L * npe * numElements;
    t = 3L * npe * numElements;
  }
  if (d > 2147483647L) {
    tacs_set_error("getMatrixFreeDataSize: %ld values exceed the int size of the interface", d);
    d = t = 0;
  }
  if (ds) *ds = (int)d;
  if (ts) *ts = (int)t;
}

void TACSAssembler::assembleMatrixFreeData(ElementMatrixType matType, TacsScalar alpha,
                                           TacsScalar beta, TacsScalar gamma, TacsScalar data[]) {
  if (!initialized) {
    tacs_set_error("assembleMatrixFreeData before initialize()");
    return;
  }
  double kc, mc;
  if (matfree_coefs(matType, alpha, gamma, &kc, &mc)) {
    tacs_set_error("assembleMatrixFreeData: matrix type %d is not supported", (int)matType);
    return;
  }
  if (family != FAMILY_SOLID) return;  // shells: no matrix-free data (TACSElement defaults)
  if (!data || ((uintptr_t)data & 15) != 0) {
    tacs_set_error("assembleMatrixFreeData: expected a 16-byte aligned device data array");
    return;
  }
  KCHECK(tacs_hip_solid_matfree_data(order, numElements, d_conn, d_X, d_telem, kc, mc, data,
                                     g_stream));
}

int TACSAssembler::matFreeProduct(const double* data, double* temp, const double* x, double* y,
                                  int accumulate, int bcs) {
  const size_t n = (size_t)numNodes * varsPerNode;
  if (family == FAMILY_SOLID) {
    if (!data || !temp) return tacs_set_error("matrix-free product: no data (assemble it first)");
    if (((uintptr_t)data & 15) != 0)
      return tacs_set_error("matrix-free product: the data array must be 16-byte aligned");
    KCHECK(tacs_hip_solid_matfree_ev(order, numElements, d_conn, data, &mat, x, temp, g_stream));
    KCHECK(tacs_hip_gather_elem_vectors(3 * npe, numNodes, d_n2e_ptr, d_n2e, temp, accumulate, y,
                                        g_stream));
  } else if (!accumulate) {
    KCHECK(tacs_hip_fill(64, n, y, 0.0, g_stream));
  }
  if (bcs)
    KCHECK(tacs_hip_apply_bc_vec(varsPerNode, nbc_nodes, d_bcnodes, d_bcmask, NULL, NULL, 1.0, y,
                                 g_stream));
  return 0;
}

void TACSAssembler::addMatrixFreeVecProduct(ElementMatrixType matType, const TacsScalar data[],
                                            TacsScalar temp[], TACSBVec* x, TACSBVec* y,
                                            MatrixOrientation matOr, const TacsScalar lambda,
                                            const bool bcs) {
  // element matrices of linear elasticity are symmetric: both orientations
  // give the same product; aux elements (tractions, pressures) carry no
  // matrix-free data, so lambda scales nothing
  if (!initialized) {
    tacs_set_error("addMatrixFreeVecProduct before initialize()");
    return;
  }
  const double* xd = x->devRead();
  double* yd = (x == y) ? NULL : y->devWrite();
  if (!yd) {
    tacs_set_error("addMatrixFreeVecProduct: x and y must differ");
    return;
  }
  matFreeProduct(data, temp, xd, yd, 1, bcs ? 1 : 0);
}

TACSMatrixFreeMat::TACSMatrixFreeMat(TACSAssembler* a)
    : assembler(a), matType(TACS_JACOBIAN_MATRIX), data_size(0), temp_size(0), data(NULL),
      temp(NULL) {
  assembler->incref();
}
TACSMatrixFreeMat::~TACSMatrixFreeMat() {
  dfree(data);
  dfree(temp);
  assembler->decref();
}
// TACSMatrixFreeMat::assembleMatrixFreeData src/TACSMatrixFreeMat.cpp:37
void TACSMatrixFreeMat::assembleMatrixFreeData(ElementMatrixType mt, double alpha, double beta,
                                               double gamma) {
  int ds = 0, ts = 0;
  assembler->getMatrixFreeDataSize(mt, &ds, &ts);
  if (ds != data_size) {
    dfree(data);
    data = ds > 0 ? dalloc<double>((size_t)ds) : NULL;
    data_size = ds;
  }
  if (ts != temp_size) {
    dfree(temp);
    temp = ts > 0 ? dalloc<double>((size_t)ts) : NULL;
    temp_size = ts;
  }
  matType = mt;
  assembler->assembleMatrixFreeData(mt, alpha, beta, gamma, data);
}
TACSVec* TACSMatrixFreeMat::createVec() { return assembler->createVec(); }
size_t TACSMatrixFreeMat::length() {
  return (size_t)assembler->getNumNodes() * assembler->getVarsPerNode();
}
void TACSMatrixFreeMat::mult64(const double* x, double* y) {
  assembler->matFreeProduct(data, temp, x, y, 0, 1);
}
// TACSMatrixFreeMat::mult src/TACSMatrixFreeMat.cpp:55: y = 0, then the
// assembler's product with the Dirichlet rows zeroed
void TACSMatrixFreeMat::mult(TACSVec* tx, TACSVec* ty) {
  TACSBVec* x = dynamic_cast<TACSBVec*>(tx);
  TACSBVec* y = dynamic_cast<TACSBVec*>(ty);
  if (!x || !y || x == y) {
    tacs_set_error("TACSMatrixFreeMat::mult: expected two distinct TACSBVec");
    return;
  }
  mult64(x->devRead(), y->devWrite());
}
void TACSMatrixFreeMat::multTranspose(TACSVec* x, TACSVec* y) { mult(x, y); }

// TACSAssembler::evalFunctions src/TACSAssembler.cpp:4617
void TACSAssembler::evalFunctions(int numFuncs, TACSFunction** funcs, TacsScalar* vals) {
  ensure(d_es, numElements);
  for (int k = 0; k < numFuncs; k++) {
    vals[k] = 0.0;
    if (!funcs[k]) continue;
    const bool sh = (family == FAMILY_SHELL);
    if (TACSCompliance* c = dynamic_cast<TACSCompliance*>(funcs[k])) {
      if (sh)
        KCHECK(shellKernel(64, 2, NULL, 1.0, 0, 0, 1, 1, NULL, NULL, d_es));
      else
        KCHECK(tacs_hip_solid_elem_scalars(order, 0, numElements, d_conn, d_X, d_u, NULL, d_telem,
                                           &mat, 1.0, 0, 0, 1, d_es, g_stream));
      KCHECK(tacs_hip_reduce(numElements, d_es, 0, d_red, d_scalar, g_stream));
      c->compliance = read_scalar(d_scalar);
      vals[k] = c->compliance;
    } else if (TACSKSFailure* f = dynamic_cast<TACSKSFailure*>(funcs[k])) {
      if (sh)
        KCHECK(shellKernel(64, 3, NULL, 1.0, f->ksWeight, 0, 1, f->safetyFactor, NULL, NULL, d_es));
      else
        KCHECK(tacs_hip_solid_elem_scalars(order, 1, numElements, d_conn, d_X, d_u, NULL, d_telem,
                                           &mat, 1.0, f->ksWeight, 0, f->safetyFactor, d_es,
                                           g_stream));
      KCHECK(tacs_hip_reduce(numElements, d_es, 1, d_red, d_scalar, g_stream));
      f->maxFail = read_scalar(d_scalar);
      if (sh)
        KCHECK(shellKernel(64, 4, NULL, 1.0, f->ksWeight, f->maxFail, 1, f->safetyFactor, NULL,
                           NULL, d_es));
      else
        KCHECK(tacs_hip_solid_elem_scalars(order, 2, numElements, d_conn, d_X, d_u, NULL, d_telem,
                                           &mat, 1.0, f->ksWeight, f->maxFail, f->safetyFactor,
                                           d_es, g_stream));
      KCHECK(tacs_hip_reduce(numElements, d_es, 0, d_red, d_scalar, g_stream));
      f->ksFailSum = read_scalar(d_scalar);
      vals[k] = f->getFunctionValue();
    } else {
      tacs_set_error("evalFunctions: unsupported function type");
    }
  }
}

// TACSAssembler::addSVSens src/TACSAssembler.cpp:4941 (no BCs applied)
void TACSAssembler::addSVSens(TacsScalar alpha, TacsScalar beta, TacsScalar gamma, int numFuncs,
                              TACSFunction** funcs, TACSBVec** dfdu) {
  ensure(d_ev, (size_t)numElements * varsPerNode * npe);
  const bool sh = (family == FAMILY_SHELL);
  for (int k = 0; k < numFuncs; k++) {
    if (!funcs[k]) continue;
    if (sh && dynamic_cast<TACSCompliance*>(funcs[k])) {
      KCHECK(shellKernel(64, 8, NULL, alpha, 0, 0, 1, 1, NULL, d_ev, NULL));
    } else if (TACSKSFailure* f = sh ? dynamic_cast<TACSKSFailure*>(funcs[k]) : NULL) {
      KCHECK(shellKernel(64, 5, NULL, alpha, f->ksWeight, f->maxFail, f->ksFailSum,
                         f->safetyFactor, NULL, d_ev, NULL));
    } else if (dynamic_cast<TACSCompliance*>(funcs[k])) {
      KCHECK(tacs_hip_solid_elem_vectors(order, 1, numElements, d_conn, d_X, d_u, d_telem, &mat,
                                         alpha, 0, 0, 1, 1, d_ev, g_stream));
    } else if (TACSKSFailure* f = dynamic_cast<TACSKSFailure*>(funcs[k])) {
      KCHECK(tacs_hip_solid_elem_vectors(order, 2, numElements, d_conn, d_X, d_u, d_telem, &mat,
                                         alpha, f->ksWeight, f->maxFail, f->ksFailSum,
                                         f->safetyFactor, d_ev, g_stream));
    } else {
      tacs_set_error("addSVSens: unsupported function type");
      continue;
    }
    KCHECK(elementVectorsToNodes(d_ev, dfdu[k]->devWrite(), 1));
  }
}

// TACSAssembler::addDVSens src/TACSAssembler.cpp:4767
void TACSAssembler::addDVSens(TacsScalar coef, int numFuncs, TACSFunction** funcs,
                              TACSBVec** dfdx) {
  ensure(d_es, numElements);
  const bool sh = (family == FAMILY_SHELL);
  for (int k = 0; k < numFuncs; k++) {
    if (!funcs[k]) continue;
    if (sh && dynamic_cast<TACSCompliance*>(funcs[k])) {
      KCHECK(shellKernel(64, 9, NULL, coef, 0, 0, 1, 1, NULL, NULL, d_es));
      KCHECK(scatterElementScalarsToDVs(d_es, dfdx[k]->devWrite()));
    } else if (TACSKSFailure* f = sh ? dynamic_cast<TACSKSFailure*>(funcs[k]) : NULL) {
      KCHECK(shellKernel(64, 6, NULL, coef, f->ksWeight, f->maxFail, f->ksFailSum,
                         f->safetyFactor, NULL, NULL, d_es));
      KCHECK(scatterElementScalarsToDVs(d_es, dfdx[k]->devWrite()));
    } else if (dynamic_cast<TACSCompliance*>(funcs[k])) {
      KCHECK(tacs_hip_solid_elem_scalars(order, 3, numElements, d_conn, d_X, d_u, NULL, d_telem,
                                         &mat, coef, 0, 0, 1, d_es, g_stream));
      KCHECK(scatterElementScalarsToDVs(d_es, dfdx[k]->devWrite()));
    } else if (dynamic_cast<TACSKSFailure*>(funcs[k])) {
      // The solid failure index does not depend on t
      // (TACSSolidConstitutive::evalFailure :274, no addFailureDVSens).
    } else {
      tacs_set_error("addDVSens: unsupported function type");
    }
  }
}

// TACSAssembler::addAdjointResProducts src/TACSAssembler.cpp:5033
void TACSAssembler::addAdjointResProducts(TacsScalar scale, int numAdjoints, TACSBVec** adjoint,
                                          TACSBVec** dfdx, const TacsScalar lambda) {
  ensure(d_es, numElements);
  for (int k = 0; k < numAdjoints; k++) {
    if (family == FAMILY_SHELL)
      KCHECK(shellKernel(64, 7, adjoint[k]->devRead(), scale, 0, 0, 1, 1, NULL, NULL, d_es));
    else
      KCHECK(tacs_hip_solid_elem_scalars(order, 4, numElements, d_conn, d_X, d_u,
                                         adjoint[k]->devRead(), d_telem, &mat, scale, 0, 0, 1,
                                         d_es, g_stream));
    KCHECK(scatterElementScalarsToDVs(d_es, dfdx[k]->devWrite()));
  }
}

// TACSAssembler::addXptSens src/TACSAssembler.cpp:4858: partial derivatives
// of the functions w.r.t. the node coordinates at fixed state (solid family).
void TACSAssembler::addXptSens(TacsScalar coef, int numFuncs, TACSFunction** funcs,
                               TACSBVec** dfdXpt) {
  if (family == FAMILY_SHELL) {
    // exact node derivatives of the element terms, 12 per element
    ensure(d_ev, (size_t)numElements * 24);
    for (int k = 0; k < numFuncs; k++) {
      if (!funcs[k]) continue;
      if (dynamic_cast<TACSCompliance*>(funcs[k])) {
        KCHECK(tacs_hip_shell_xpt_kernel(1, numElements, d_conn, d_X, d_u, NULL, d_telem, &smat,
                                         coef, 0, 0, 1, 1, d_ev, g_stream));
      } else if (TACSKSFailure* f = dynamic_cast<TACSKSFailure*>(funcs[k])) {
        KCHECK(tacs_hip_shell_xpt_kernel(2, numElements, d_conn, d_X, d_u, NULL, d_telem, &smat,
                                         coef, f->ksWeight, f->maxFail, f->ksFailSum,
                                         f->safetyFactor, d_ev, g_stream));
      } else {
        tacs_set_error("addXptSens: unsupported function type");
        continue;
      }
      KCHECK(tacs_hip_gather_elem_vectors_b(3, 12, numNodes, d_n2e_ptr, d_n2e, d_ev, 1,
                                            dfdXpt[k]->devWrite(), g_stream));
    }
    return;
  }
  ensure(d_ev, (size_t)numElements * 3 * npe);
  for (int k = 0; k < numFuncs; k++) {
    if (!funcs[k]) continue;
    if (dynamic_cast<TACSCompliance*>(funcs[k])) {
      KCHECK(tacs_hip_solid_xpt_vectors(order, 0, numElements, d_conn, d_X, d_u, NULL, d_telem,
                                        &mat, coef, 0, 0, 1, 1, d_ev, g_stream));
    } else if (TACSKSFailure* f = dynamic_cast<TACSKSFailure*>(funcs[k])) {
      KCHECK(tacs_hip_solid_xpt_vectors(order, 1, numElements, d_conn, d_X, d_u, NULL, d_telem,
                                        &mat, coef, f->ksWeight, f->maxFail, f->ksFailSum,
                                        f->safetyFactor, d_ev, g_stream));
    } else {
      tacs_set_error("addXptSens: unsupported function type");
      continue;
    }
    KCHECK(tacs_hip_gather_elem_vectors(3 * npe, numNodes, d_n2e_ptr, d_n2e, d_ev, 1,
                                        dfdXpt[k]->devWrite(), g_stream));
  }
}

// TACSAssembler::addAdjointResXptSensProducts src/TACSAssembler.cpp:5136:
// dfdXpt += scale psi^T dR/dXpt, element residuals plus aux elements (whose
// contribution is scaled by lambda, as there).
void TACSAssembler::addAdjointResXptSensProducts(TacsScalar scale, int numAdjoints,
                                                 TACSBVec** adjoint, TACSBVec** dfdXpt,
                                                 const TacsScalar lambda) {
  if (family == FAMILY_SHELL) {
    ensure(d_ev, (size_t)numElements * 24);
    for (int k = 0; k < numAdjoints; k++) {
      const double* psi = adjoint[k]->devRead();
      double* out = dfdXpt[k]->devWrite();
      KCHECK(tacs_hip_shell_xpt_kernel(0, numElements, d_conn, d_X, d_u, psi, d_telem, &smat,
                                       scale, 0, 0, 1, 1, d_ev, g_stream));
      KCHECK(tacs_hip_gather_elem_vectors_b(3, 12, numNodes, d_n2e_ptr, d_n2e, d_ev, 1, out,
                                            g_stream));
      if (uploadAux() == 0 && naux > 0) {
        // pressure loads (aux residual scaled by lambda, TACSAssembler.cpp:4210-4225)
        KCHECK(tacs_hip_shell_pressure_xpt(naux, d_aux_elem, d_aux_pres, d_conn, d_X, psi,
                                           lambda * scale, d_aux_ev, g_stream));
        KCHECK(tacs_hip_gather_elem_vectors_b(3, 12, numNodes, d_n2a_ptr, d_n2a, d_aux_ev, 1, out,
                                              g_stream));
      }
    }
    return;
  }
  ensure(d_ev, (size_t)numElements * 3 * npe);
  for (int k = 0; k < numAdjoints; k++) {
    const double* psi = adjoint[k]->devRead();
    double* out = dfdXpt[k]->devWrite();
    KCHECK(tacs_hip_solid_xpt_vectors(order, 2, numElements, d_conn, d_X, d_u, psi, d_telem, &mat,
                                      scale, 0, 0, 1, 1, d_ev, g_stream));
    KCHECK(tacs_hip_gather_elem_vectors(3 * npe, numNodes, d_n2e_ptr, d_n2e, d_ev, 1, out,
                                        g_stream));
    if (uploadAux() == 0 && naux > 0) {
      if (naux_pressure) {
        tacs_set_error("addAdjointResXptSensProducts: pressure aux elements are not supported");
        return;
      }
      KCHECK(tacs_hip_solid_traction_xpt_vectors(order, naux, d_aux_elem, d_aux_face, d_aux_trac,
                                                 d_conn, d_X, psi, lambda * scale, d_aux_ev,
                                                 g_stream));
      KCHECK(tacs_hip_gather_elem_vectors(3 * npe, numNodes, d_n2a_ptr, d_n2a, d_aux_ev, 1, out,
                                          g_stream));
    }
  }
}
