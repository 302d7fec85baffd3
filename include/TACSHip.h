/*
  C++ API of the MI355X implementation of TACS's assembly / block-CSR /
  GMRES / adjoint path.  The class names, constructor argument orders, method
  signatures, reference-counting rules (objects start at refcount 0; callers
  incref/decref; arguments are borrowed) and int/stderr error convention
  follow the reference headers cited on each class, so a C++ driver written
  against the reference's subset below compiles against this header.

  Everything lives on one GPU: vectors and matrices are device resident;
  TACSBVec::getArray returns a host mirror that is synchronised on demand.
  The constructors take the reference's MPI_Comm (src/TACSAssembler.h:83);
  this build runs one rank on one GPU, so a communicator of more than one
  process is refused by TACSAssembler::initialize (the library calls no MPI
  routine itself and does not link MPI; it asks the process's MPI, when one
  is loaded and initialised, for the communicator's size).
*/
#ifndef TACS_HIP_CXX_H
#define TACS_HIP_CXX_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdio.h>

#include <string>
#include <vector>

#define MPICH_SKIP_MPICXX
#define OMPI_SKIP_MPICXX
#include <mpi.h>

#include "tacs_hip.h"

typedef double TacsScalar;
typedef MPI_Comm TACSComm;
#define TACS_COMM_SELF MPI_COMM_SELF

enum MatrixOrientation { TACS_MAT_NORMAL, TACS_MAT_TRANSPOSE };
enum ElementStrainType { TACS_LINEAR_STRAIN, TACS_NONLINEAR_STRAIN };
enum OrthoType { CLASSICAL_GRAM_SCHMIDT, MODIFIED_GRAM_SCHMIDT };
// src/elements/TACSElementTypes.h:113
enum ElementMatrixType {
  TACS_JACOBIAN_MATRIX,
  TACS_STIFFNESS_MATRIX,
  TACS_MASS_MATRIX,
  TACS_GEOMETRIC_STIFFNESS_MATRIX,
  TACS_STIFFNESS_PRODUCT_DERIVATIVE
};

// src/TACSObject.h:112-140
class TACSObject {
 public:
  TACSObject();
  virtual ~TACSObject();
  void incref();
  void decref();
  int refcount() const { return ref_count; }
  virtual const char* getObjectName();

 private:
  int ref_count;
};

// src/constitutive/TACSMaterialProperties.h (isotropic constructor)
class TACSMaterialProperties : public TACSObject {
 public:
  TACSMaterialProperties(TacsScalar rho, TacsScalar specific_heat, TacsScalar E,
                         TacsScalar nu, TacsScalar ys, TacsScalar alpha = 0.0,
                         TacsScalar kappa = 0.0);
  TacsScalar getDensity() const { return rho; }
  void evalTangentStiffness3D(TacsScalar C[]);
  TacsScalar rho, specific_heat, E, nu, G, ys, alpha, kappa;
};

class TACSConstitutive : public TACSObject {};

// src/constitutive/TACSSolidConstitutive.h
class TACSSolidConstitutive : public TACSConstitutive {
 public:
  TACSSolidConstitutive(TACSMaterialProperties* props, TacsScalar t = 1.0,
                        int tNum = -1, TacsScalar tlb = 0.0, TacsScalar tub = 1.0);
  ~TACSSolidConstitutive();
  TACSMaterialProperties* getMaterialProperties() { return props; }
  TACSMaterialProperties* props;
  TacsScalar t, tlb, tub;
  int tNum;
};

// src/constitutive/TACSShellConstitutive.h, TACSIsoShellConstitutive.h:33
class TACSShellConstitutive : public TACSConstitutive {};
class TACSIsoShellConstitutive : public TACSShellConstitutive {
 public:
  TACSIsoShellConstitutive(TACSMaterialProperties* props, TacsScalar t = 1.0, int tNum = -1,
                           TacsScalar tlb = 0.0, TacsScalar tub = 10.0,
                           TacsScalar tOffset = 0.0, TacsScalar kcorr = 5.0 / 6.0);
  ~TACSIsoShellConstitutive();
  TACSMaterialProperties* getMaterialProperties() { return props; }
  TACSMaterialProperties* props;
  TacsScalar t, tlb, tub, tOffset, kcorr;
  int tNum;
};

// src/elements/basis/TACSHexaBasis.h
class TACSElementBasis : public TACSObject {
 public:
  virtual int getOrder() = 0;  // nodes per direction (2: 8-node, 3: 27-node)
  int getNumNodes() {
    int o = getOrder();
    return o * o * o;
  }
};
class TACSLinearHexaBasis : public TACSElementBasis {
 public:
  int getOrder() { return 2; }
};
class TACSQuadraticHexaBasis : public TACSElementBasis {
 public:
  int getOrder() { return 3; }
};

class TACSElementModel : public TACSObject {
 public:
  virtual int getVarsPerNode() = 0;
};

// src/elements/TACSLinearElasticity.h:805 (TACSLinearElasticity3D)
class TACSLinearElasticity3D : public TACSElementModel {
 public:
  TACSLinearElasticity3D(TACSSolidConstitutive* stiff,
                         ElementStrainType strain_type = TACS_LINEAR_STRAIN);
  ~TACSLinearElasticity3D();
  int getVarsPerNode() { return 3; }
  TACSSolidConstitutive* stiff;
  ElementStrainType strain_type;
};

// src/elements/TACSElement.h:37
class TACSElement : public TACSObject {
 public:
  virtual int getVarsPerNode() = 0;
  virtual int getNumNodes() = 0;
  int getNumVariables() { return getVarsPerNode() * getNumNodes(); }
};

// src/elements/TACSElement3D.h
class TACSElement3D : public TACSElement {
 public:
  TACSElement3D(TACSElementModel* model, TACSElementBasis* basis);
  ~TACSElement3D();
  int getVarsPerNode() { return model->getVarsPerNode(); }
  int getNumNodes() { return basis->getNumNodes(); }
  // TACSElement3D::createElementTraction / createElementPressure
  // src/elements/TACSElement3D.cpp:48, :54 (new objects, reference count 0)
  TACSElement* createElementTraction(int faceIndex, const TacsScalar t[]);
  TACSElement* createElementPressure(int faceIndex, TacsScalar p);
  TACSElementModel* model;
  TACSElementBasis* basis;
};

// src/elements/TACSPressure3D.h: constant pressure p on face faceIndex of a
// solid element; the residual adds -p n (n the outward unit normal of the
// face) integrated against the basis (TACSPressure3D::addResidual
// src/elements/TACSPressure3D.cpp:60).
class TACSPressure3D : public TACSElement {
 public:
  TACSPressure3D(int varsPerNode, int faceIndex, TACSElementBasis* basis, TacsScalar p);
  ~TACSPressure3D();
  int getVarsPerNode() { return varsPerNode; }
  int getNumNodes() { return basis->getNumNodes(); }
  int varsPerNode, faceIndex;
  TACSElementBasis* basis;
  TacsScalar p;
};

// src/elements/TACSTraction3D.h:22 (constant traction, coordinate components)
class TACSTraction3D : public TACSElement {
 public:
  TACSTraction3D(int varsPerNode, int faceIndex, TACSElementBasis* basis,
                 const TacsScalar trac[], int tractionCoordinateComponent = 1);
  ~TACSTraction3D();
  int getVarsPerNode() { return varsPerNode; }
  int getNumNodes() { return basis->getNumNodes(); }
  int varsPerNode, faceIndex, coordComponent;
  TACSElementBasis* basis;
  TacsScalar trac[24];
};

// src/elements/shell/TACSShellElementTransform.h:8-92
class TACSShellTransform : public TACSObject {};
class TACSShellNaturalTransform : public TACSShellTransform {};
// src/elements/shell/TACSShellElementTransform.h:95: frame from the reference
// axis projected on the tangent plane
class TACSShellRefAxisTransform : public TACSShellTransform {
 public:
  TACSShellRefAxisTransform(const TacsScalar _axis[]);
  void getRefAxis(TacsScalar _axis[]);
  TacsScalar axis[3];
};

// TACSQuad4Shell: src/elements/shell/TACSShellElementDefs.h:18-20 over
// TACSShellElement (src/elements/shell/TACSShellElement.h:39); linear model,
// linearized rotations, 6 variables per node.
class TACSQuad4Shell : public TACSElement {
 public:
  TACSQuad4Shell(TACSShellTransform* transform, TACSShellConstitutive* con);
  ~TACSQuad4Shell();
  int getVarsPerNode() { return 6; }
  int getNumNodes() { return 4; }
  // TACSShellElement::createElementPressure (TACSShellElement.h:109)
  TACSElement* createElementPressure(int faceIndex, TacsScalar p);
  TACSShellTransform* transform;
  TACSShellConstitutive* con;
};

// src/elements/shell/TACSShellPressure.h:9 (uniform pressure)
class TACSShellPressure : public TACSElement {
 public:
  TACSShellPressure(TacsScalar p) : p(p) {}
  int getVarsPerNode() { return 6; }
  int getNumNodes() { return 4; }
  TacsScalar p;
};

// src/TACSAuxElements.h:52-101
class TACSAuxElements : public TACSObject {
 public:
  TACSAuxElements(int size = 100);
  ~TACSAuxElements();
  void addElement(int num, TACSElement* elem);
  int getNumAux() const { return n; }
  void clear();
  int n, cap, version;
  int* nums;
  TACSElement** elems;
};

// src/bpmat/KSM.h:392 (TACSBcMap)
class TACSBcMap : public TACSObject {
 public:
  TACSBcMap(int bsize, int num_bcs);
  ~TACSBcMap();
  void addBC(int node, int nvals, const int* bc_vars = NULL,
             const TacsScalar* bc_vals = NULL);
  int getBCs(const int** nodes, const int** vars, TacsScalar** values);
  int bsize, nbcs, max_size;
  int *nodes, *vars;
  TacsScalar* values;
};

// src/bpmat/KSM.h:43 (TACSVec) and src/bpmat/TACSBVec.h:67 (TACSBVec)
class TACSVec : public TACSObject {
 public:
  virtual TacsScalar norm() = 0;
  virtual void scale(TacsScalar alpha) = 0;
  virtual TacsScalar dot(TACSVec* x) = 0;
  virtual void axpy(TacsScalar alpha, TACSVec* x) = 0;
  virtual void copyValues(TACSVec* x) = 0;
  virtual void axpby(TacsScalar alpha, TacsScalar beta, TACSVec* x) = 0;
  virtual void zeroEntries() = 0;
  virtual void setRand(double lower = -1.0, double upper = 1.0) = 0;
};

class TACSBVec : public TACSVec {
 public:
  TACSBVec(int nnodes, int bsize);
  ~TACSBVec();
  void getSize(int* size);
  int getBlockSize() { return bsize; }
  int getArray(TacsScalar** array);  // host mirror; host becomes current
  TacsScalar norm();
  void scale(TacsScalar alpha);
  TacsScalar dot(TACSVec* x);
  void axpy(TacsScalar alpha, TACSVec* x);
  void copyValues(TACSVec* x);
  void axpby(TacsScalar alpha, TacsScalar beta, TACSVec* x);
  void zeroEntries();
  void set(TacsScalar val);
  void setRand(double lower = -1.0, double upper = 1.0);
  void applyBCs(TACSBcMap* map, TACSVec* vec = NULL, TacsScalar lambda = 1.0);
  void setBCs(TACSBcMap* map, TacsScalar lambda = 1.0);
  // Device access: devRead uploads a newer host copy; devWrite also marks
  // the device copy as the current one.
  const double* devRead();
  double* devWrite();
  size_t length() const { return n; }
  int nnodes, bsize;
  size_t n;
  double* d_x;
  double* h_x;
  int state;  // 0 both current, 1 device newer, 2 host newer
};

// src/bpmat/KSM.h:91 (TACSMat), src/bpmat/TACSParallelMat.h:52 (one rank:
// the [Aloc] block-CSR matrix of BCSRMat, src/bpmat/BCSRMat.h:34)
class TACSMat : public TACSObject {
 public:
  virtual TACSVec* createVec() = 0;
  virtual void mult(TACSVec* x, TACSVec* y) = 0;
  virtual void zeroEntries() = 0;
  virtual void applyBCs(TACSBcMap* bcmap) = 0;
};

class TACSParallelMat : public TACSMat {
 public:
  TACSParallelMat(int bsize, int nrows, const int* rowp, const int* cols);
  ~TACSParallelMat();
  TACSVec* createVec();
  void mult(TACSVec* x, TACSVec* y);
  void zeroEntries();
  void applyBCs(TACSBcMap* bcmap);
  int getSize(int* nr, int* nc);
  // Host copies of the block-CSR arrays (BCSRMat::getArrays BCSRMat.h:82)
  void getArrays(int* bsize, int* nrows, const int** rowp, const int** cols,
                 TacsScalar** A);
  // FP32 shadow of the values for the mixed-precision solver
  float* getValues32(int refresh);
  // y = A x on device vectors: FP64 values and vectors; FP32-rounded values
  // with FP64 vectors (the recycling solver's operator); FP32 values and
  // vectors.  The products read a sliced ELL copy of the values (3x3 or 6x6
  // blocks; slices of 64 rows, one lane per row), refreshed when the values
  // change.
  void mult64(const double* x, double* y);
  void multMixed(const double* x, double* y);
  void mult32(const float* x, float* y);
  // one Chebyshev step for A z = rhs fused into the product (values of
  // precision pr, Dinv in that precision): d = c1 d + c2 D^{-1} (rhs - A z),
  // zout = z + d
  void chebStep(int pr, const void* Dinv, const double* z, double* zout, const double* rhs,
                double* d, double c1, double c2);
  // the same step in three-term form: zprev <- (1 + c1) z - c1 zprev
  // + c2 D^{-1} (rhs - A z)
  void chebStep3(int pr, const void* Dinv, const double* z, double* zprev, const double* rhs,
                 double c1, double c2);
  int bsize, nrows;
  long nnzb;
  long version;  // incremented whenever the values change
  // sliced ELL copy
  void sellUpdate(int prec);
  int* d_sw;
  long* d_soff;
  int* d_scols;
  double* d_s64;
  float* d_s32;
  unsigned short* d_s16;  // bfloat16 copy (Chebyshev polynomial products)
  long sellVer64, sellVer32, sellVer16;
  size_t nslots;
  int *d_rowp, *d_cols;
  double* d_A;
  float* d_A32;
  int A32_current;
  int* h_rowp;
  int* h_cols;
  double* h_A;
  // launch variant of the sliced ELL product per kind (0: FP64, 1: FP32
  // values with FP64 vectors, 2: FP32), -1 for the tuning table's; every
  // variant gives the same sums.  autotune() times each variant of each kind
  // on this matrix (reps products each) and keeps the fastest; ms receives
  // the average times [kind][variant] (room for 3 x maxv); returns the
  // number of variants per kind.  No counterpart in the reference.
  int sellVar[3];
  int autotune(int reps, double* ms, int maxv);
  // times every launch variant of the nrhs right-hand-side block product
  // (3x3 blocks; prec 64: FP64 values, 32: FP32 values) on this matrix and
  // enters the fastest in the tuning table for this size class (the batched
  // solvers launch the table's variant); returns the number of variants.
  int autotuneSpMM(int nrhs, int prec, int reps, double* ms, int maxv);
};

// src/bpmat/KSM.h:147 (TACSPc)
class TACSPc : public TACSObject {
 public:
  virtual void factor() = 0;
  virtual void applyFactor(TACSVec* x, TACSVec* y) = 0;
};

// Block-Jacobi: inverse diagonal blocks of the assembled matrix (the quantity
// BCSRMat::factorDiag src/bpmat/BCSRMat.cpp:1527 computes), in FP64 and FP32.
class TACSBlockJacobiPc : public TACSPc {
 public:
  TACSBlockJacobiPc(TACSParallelMat* mat);
  ~TACSBlockJacobiPc();
  void factor();
  void applyFactor(TACSVec* x, TACSVec* y);
  void applyFactor32(const float* x, float* y);
  TACSParallelMat* mat;
  int* d_diag;
  double* d_Dinv;
  float* d_Dinv32;
  // Chebyshev polynomial of the block-Jacobi iteration (no counterpart in
  // the reference): with degree > 1 the recycling and batched solvers apply
  // M^{-1} x = z_degree of the Chebyshev iteration for A z = x (degree - 1
  // products with the solver's operator values), on the interval
  // [lmax / ratio, lmax] of the eigenvalues of D^{-1} A; factor() estimates
  // lmax by 30 Lanczos steps (FP32-rounded values, x 1.05).  The
  // smallest eigenvalues are left to the recycled space.  degree 1 (the
  // default) is plain block-Jacobi.
  // values 16: the products inside the polynomial read a bfloat16 copy of
  // the values (single systems; the outer operator keeps its values) or,
  // in the batched solver with 16 or more right-hand sides, an FP32 copy of
  // the iterate
  int setChebyshev(int degree, double ratio, int values = 32);
  double lanczosMax(int steps);
  int chebDegree = 1, chebValues = 32;
  double chebRatio = 20.0, chebMax = 0.0;
};

// src/bpmat/KSM.h:200 (TACSKsm), :266 (GMRES)
// Residual monitors (KSMPrint src/bpmat/KSM.h:229, KSMPrintStdout :293,
// KSMPrintFile :310): printResidual(iter, res) prints "descript[iter]: res"
// when iter == 0 or iter is a multiple of freq, on rank 0; the file variant
// flushes each line.
class KSMPrint : public TACSObject {
 public:
  virtual ~KSMPrint() {}
  virtual void printResidual(int iter, TacsScalar res) = 0;
  virtual void print(const char* cstr) = 0;
};
class KSMPrintStdout : public KSMPrint {
 public:
  KSMPrintStdout(const char* descript, int rank, int freq);
  void printResidual(int iter, TacsScalar res);
  void print(const char* cstr);
  std::string descript;
  int rank, freq;
};
class KSMPrintFile : public KSMPrint {
 public:
  KSMPrintFile(const char* filename, const char* descript, int rank, int freq);
  ~KSMPrintFile();
  void printResidual(int iter, TacsScalar res);
  void print(const char* cstr);
  FILE* fp;
  std::string descript;
  int rank, freq;
};

class TACSKsm : public TACSObject {
 public:
  virtual ~TACSKsm() {
    if (monitor) monitor->decref();
  }
  virtual int solve(TACSVec* b, TACSVec* x, int zero_guess = 1) = 0;
  virtual void setTolerances(double rtol, double atol) = 0;
  // GMRES::setMonitor src/bpmat/KSM.cpp:739 (the monitor is reference counted)
  virtual void setMonitor(KSMPrint* m) {
    if (m) m->incref();
    if (monitor) monitor->decref();
    monitor = m;
  }
  int getIterCount() { return iterCount; }
  TacsScalar getResidualNorm() { return resNorm; }
  int iterCount = 0;
  TacsScalar resNorm = 0.0;
  KSMPrint* monitor = NULL;
};

class TACSMatrixFreeMat;

// GMRES::solve src/bpmat/KSM.cpp:785: right-preconditioned; with isFlexible
// the preconditioned vectors Z_i = M^{-1} v_i are kept and the update is
// x += Z y (FGMRES), otherwise x += M^{-1} (V y).
class GMRES : public TACSKsm {
 public:
  GMRES(TACSMat* mat, TACSPc* pc, int m, int nrestart, int isFlexible = 0);
  ~GMRES();
  int solve(TACSVec* b, TACSVec* x, int zero_guess = 1);
  void setTolerances(double rtol, double atol);
  void setOrthoType(enum OrthoType otype);
  // Run each restart cycle's m Arnoldi steps as one captured HIP graph
  // (assembled operator only): no host round trip per step; each step's
  // residual estimate goes to a history array and the host then takes the
  // first step that meets the stopping test, which gives the same iterate as
  // stopping there (later steps touch no earlier Hessenberg column, residual
  // entry or basis vector).  The steps after convergence in the last cycle
  // are extra work.  On by default (MI355X, FP64 GMRES(30) + MGS per step:
  // 8^3 131 -> 129 us, 12^3 179 -> 141 us, 20^3 175 -> 154 us, 40^3 206 ->
  // 192 us); setCycleGraph(0) runs the steps one by one.  No counterpart in
  // the reference.
  void setCycleGraph(int on);
  double rtol, atol;
  int msub, nrestart, ortho;
  int useGraph;
  hipStream_t gstream;     // capture stream (non-blocking, default priority)
  hipGraphExec_t gexec;    // the captured cycle (NULL: none)
  int gexec_ortho;         // Gram-Schmidt the captured cycle was built with
  double* d_hist;          // per step |residual estimate| of the cycle
  double* h_hist;          // pinned host copy
  void arnoldiStep(int i, double* st);
  int captureCycle();
  // the operator: an assembled matrix or a matrix-free one (the other NULL)
  TACSParallelMat* mat;
  TACSMatrixFreeMat* mfree;
  size_t n;
  void opMult(const double* x, double* y);
  TACSBlockJacobiPc* pc;
  int flexible;
  double* d_W;      // (msub+1) basis vectors, contiguous
  double* d_Z;      // msub preconditioned vectors (flexible only)
  double* d_work;
  double* d_scal;   // H, res, Qc, Qs, status scalars
  void* d_red;
  double* h_pin;    // pinned host scalars
};

class TACSBatchedGMRES;

// Mixed-precision restarted GMRES(m): each cycle starts from the FP64
// residual b - A x and runs its Arnoldi steps on the FP32 matrix, FP32
// block-Jacobi and an FP32 basis (Gram-Schmidt dot products, Hessenberg and
// Givens rotations in FP64); the correction is accumulated in FP64.  No
// counterpart in the reference.
class TACSMixedGMRES : public TACSKsm {
 public:
  TACSMixedGMRES(TACSParallelMat* mat, TACSBlockJacobiPc* pc, int m, int nrestart);
  ~TACSMixedGMRES();
  int solve(TACSVec* b, TACSVec* x, int zero_guess = 1);
  int solveDeflated(TACSVec* b, TACSVec* x, int zero_guess);
  void setTolerances(double rtol, double atol);
  // Krylov subspace recycling (GCRO-DR, Parks et al., SIAM J. Sci. Comput.
  // 28 (2006)): each cycle keeps the k harmonic Ritz vectors of smallest
  // harmonic Ritz value of the fixed FP32-valued operator, and the space is
  // reused by later cycles and later solves with the same matrix
  // (k <= 31, k < m - 1, m + k <= 250); k = 0 restarts plainly.
  int setDeflation(int k);
  double rtol, atol;
  int msub, nrestart, outerCount, ndefl;
  int recycleDim;       // k of the current recycle space (0: none)
  long recycleVersion;  // matrix version the recycle space belongs to
  TACSParallelMat* mat;
  TACSBlockJacobiPc* pc;
  float* d_W;
  float* d_work32;
  float* d_d32;
  TACSBatchedGMRES* rec;  // the recycling cycle (one right-hand side), k > 0
  double* d_r;
  double* d_scal;
  void* d_red;
  double* h_pin;
};

// Batched restarted GMRES(m) for up to R right-hand sides of one 3x3-block
// matrix (R <= 64): the R Arnoldi recurrences advance together, so every step
// reads the matrix once for all right-hand sides (multi-RHS SpMM over
// interleaved block vectors V[i*R + j]) and the Gram-Schmidt products of all
// right-hand sides run as one batched multi-dot.  Every cycle restarts from
// the true FP64 residual B - A X.  prec 64 runs the Arnoldi steps on the FP64
// matrix and block-Jacobi inverse, prec 32 on their FP32-rounded values (FP64
// vectors and arithmetic).  Each right-hand side stops on its own criterion
// and keeps its own iteration count.  No counterpart in the reference (one
// GMRES::solve per load case there, src/bpmat/KSM.cpp:785).
class TACSBatchedGMRES : public TACSObject {
 public:
  TACSBatchedGMRES(TACSParallelMat* mat, TACSBlockJacobiPc* pc, int m, int nrestart, int nrhs,
                   int prec);
  // Segmented systems: mat is block diagonal with nseg <= 64 diagonal
  // blocks, block j the node range [node_off[j], node_off[j+1]) (the meshes
  // of a multi-mesh assembler); each block is solved as its own system, with
  // its own Krylov space (and recycled space, setDeflation), Hessenberg
  // matrix and stopping test, all advancing together (solveSegments).  A
  // block vector holds local node i of system j at (3 i + c) nseg + j, sized
  // by the largest block.  Rows that couple two blocks are refused.
  TACSBatchedGMRES(TACSParallelMat* mat, TACSBlockJacobiPc* pc, int m, int nrestart, int nseg,
                   const int* node_off, int prec);
  ~TACSBatchedGMRES();
  // x = A^{-1} b for segmented systems (b, x: whole-matrix vectors); each
  // block stops when its own residual meets rtol |b_j| (or atol); returns 1
  // when every block converged; iters, flags, r0 and rn are per block
  int solveSegments(TACSBVec* b, TACSBVec* x, int zero_guess = 1);
  int nseg = 0;           // segmented systems (0: right-hand sides of one matrix)
  int segMaxNodes = 0;    // nodes of the largest block
  int* d_segoff = NULL;   // node_off on the device
  double *d_ux = NULL, *d_uy = NULL;  // whole-matrix work vectors of the products
  std::vector<int> segOff;
  int sysRows() const { return nseg > 0 ? segMaxNodes : mat->nrows; }
  void allocate();
  void bdiag(int pr, int Rn, const void* Dinv, const double* x, double* y);
  // y = M^{-1} x: block-Jacobi, or the preconditioner's Chebyshev polynomial
  void precond(int pr, int Rn, const void* Dinv, const float* A32, const double* x, double* y);
  double *d_cd = NULL, *d_caz = NULL;  // Chebyshev work vectors
  float* d_cz32 = NULL;                // FP32 copies of the iterate (two buffers)
  void bmult(int pr, int Rn, const float* A32, const double* x, double* y);
  // x[j] = A^{-1} b[j], j < nrhs <= R; returns 1 when every system converged
  int solve(int nrhs, TACSBVec** b, TACSBVec** x, int zero_guess = 1);
  // residual monitor: at the start of each cycle, the true residual norm of
  // every system still iterating, printResidual(0, |r_j|)
  void setMonitor(KSMPrint* m) {
    if (m) m->incref();
    if (monitor) monitor->decref();
    monitor = m;
  }
  KSMPrint* monitor = NULL;
  // the same on interleaved device block vectors B[i*nrhs + j], X[i*nrhs + j]
  int solveBlock(int nrhs, const double* B, double* X, int zero_guess);
  void setTolerances(double rtol, double atol);
  // Deflate every right-hand side by the recycled space (U, C) of a
  // recycling solver over the same matrix (A32 D32^{-1} U = C, C orthonormal):
  // each cycle starts with x += D^{-1} U C^T r, r -= C C^T r and runs its
  // Arnoldi steps on (I - C C^T) A M^{-1}; the space stays fixed.  Copies the
  // space; it is dropped when the matrix values change.  Returns its
  // dimension (0: src has none).
  int setRecycleSpace(TACSMixedGMRES* src);
  // Krylov subspace recycling per right-hand side (GCRO-DR, Parks et al.,
  // SIAM J. Sci. Comput. 28 (2006), as TACSMixedGMRES::setDeflation): every
  // right-hand side keeps its own k-dimensional space (U_j, C_j), updated each
  // cycle from its harmonic Ritz vectors; the spaces persist across solves
  // with the same matrix values and number of right-hand sides and start from
  // the setRecycleSpace space when there is one (0 <= k <= 31, k < m - 1,
  // m + k <= 250; k = 0 turns recycling off).
  int setDeflation(int k);
  int solveBlockDeflated(int nrhs, const double* B, double* X, int zero_guess);
  double rtol, atol;
  int msub, nrestart, R, prec, cycles;
  int kdefl;         // dimension of the shared recycled space (0: none)
  long kVersion;     // matrix version of the shared recycled space
  double *d_U, *d_C;
  int ndefl;         // per right-hand side recycling dimension (0: off)
  long defVersion;   // matrix version of the per right-hand side spaces
  int defR;          // number of right-hand sides they were built for
  double *d_Ub, *d_Cb, *d_Unb, *d_Cnb;  // K block vectors each
  double* d_prod;    // per right-hand side small matrices (products, basis changes)
  double* h_stage;   // pinned staging of the small matrices, R x prodStride()
  double* h_up;      // pinned staging of uploads (same size); an upload is not
                     // waited for: the next one waits on ev_up before reusing it
  hipEvent_t ev_up;
  int up_pending;
  // device-side DCGS2 steps (tacs_hip_bdcgs_coeffs): least-squares state,
  // 4 ints per right-hand side, recycled dimensions and stopping targets;
  // host-visible flags of the last DCGS_LAG + 1 steps and their events
  double* d_lsst;
  int* d_ist;
  int* d_kdv;
  double* d_tgtv;
  int* h_flags;
  int* d_flags;
  hipEvent_t ev_step[3];
  void* d_dq;        // scratch of the delayed Gram-Schmidt dots and the products
  // Gram-Schmidt of the recycling (GCRO-DR) Arnoldi steps: 1 (default)
  // delayed classical Gram-Schmidt (DCGS2, two sweeps over the basis per
  // step), 0 classical Gram-Schmidt twice (four sweeps)
  int delayed;
  std::vector<int> kdj;      // dimension of each right-hand side's space
  std::vector<double> dsc;   // 1 / |U_j(:, v)|, R x KD_MAX
  TACSParallelMat* mat;
  TACSBlockJacobiPc* pc;
  double* d_V;     // (m+1) block vectors of n*R
  double* d_work;  // block work vector
  double* d_Bb;    // packed right-hand sides / solutions of solve()
  double* d_Xb;
  double* d_scal;  // per right-hand side: H, res, Qc, Qs, tmp (stride slice())
  double* d_vec;   // |r|^2, |r|, 1/|r|, Givens estimates (64 each)
  int* d_cnt;
  void* d_red;
  double* h_pin;
  std::vector<int> iters, flags;
  std::vector<double> r0, rn;
  // per right-hand side: H (m+1) x m, res m+1, Qc m, Qs m, tmp m+1, then the
  // recycled-space coefficients Bc (KD_MAX x m), tmpC, z, projection (KD_MAX
  // each), then the delayed Gram-Schmidt dot products (2 x (KD_MAX + m + 1))
  // and update coefficients (2 (KD_MAX + m) + 2)
  static const int KD_MAX = 32;
  // steps queued ahead of the host's look at a step's stopping flags
  static const int DCGS_LAG = 2;
  // per right-hand side: four (KD_MAX + m + 1) x KD_MAX matrices
  size_t prodStride() const { return (size_t)4 * (KD_MAX + msub + 1) * KD_MAX; }
  size_t slice() const {
    return (size_t)(msub + 1) * msub + 4 * (size_t)msub + 2 + (size_t)KD_MAX * (msub + 3) +
           4 * (size_t)(KD_MAX + msub) + 4;
  }
};

class TACSAssembler;

// src/functions/TACSFunction.h:107
class TACSFunction : public TACSObject {
 public:
  TACSFunction(TACSAssembler* assembler);
  ~TACSFunction();
  virtual TacsScalar getFunctionValue() = 0;
  TACSAssembler* getAssembler() { return assembler; }
  TACSAssembler* assembler;
};

// src/functions/TACSCompliance.h
class TACSCompliance : public TACSFunction {
 public:
  TACSCompliance(TACSAssembler* assembler);
  TacsScalar getFunctionValue() { return compliance; }
  TacsScalar compliance = 0.0;
};

// src/functions/TACSKSFailure.h (KS_CONTINUOUS aggregation)
class TACSKSFailure : public TACSFunction {
 public:
  TACSKSFailure(TACSAssembler* assembler, double ksWeight = 80.0,
                double alpha = 1.0, double safetyFactor = 1.0);
  TacsScalar getFunctionValue();
  TacsScalar getMaximumFailure() { return maxFail; }
  double ksWeight, alpha, safetyFactor;
  TacsScalar maxFail = -1e20, ksFailSum = 0.0;
};

// src/TACSAssembler.h:61-325 (single rank, one GPU)
class TACSAssembler : public TACSObject {
 public:
  TACSAssembler(TACSComm comm, int varsPerNode, int numOwnedNodes,
                int numElements, int numDependentNodes = 0);
  ~TACSAssembler();

  int setElementConnectivity(const int* ptr, const int* conn);
  int setElements(TACSElement** elements);
  int addBCs(int nnodes, const int* nodes, int nbcs = -1, const int* vars = NULL,
             const TacsScalar* vals = NULL);
  int initialize();
  void setAuxElements(TACSAuxElements* aux);
  TACSAuxElements* getAuxElements() { return auxElements; }

  int getVarsPerNode() { return varsPerNode; }
  int getNumNodes() { return numNodes; }
  int getNumOwnedNodes() { return numNodes; }
  int getNumElements() { return numElements; }
  int getNumDesignVars() { return numDesignVars; }
  TACSBcMap* getBcMap() { return bcMap; }
  void setNumThreads(int t) {}
  TACSComm getMPIComm() { return comm; }

  TACSBVec* createVec();
  TACSBVec* createNodeVec();
  TACSBVec* createDesignVec();
  TACSParallelMat* createMat();

  void setNodes(TACSBVec* X);
  void getNodes(TACSBVec* X);
  void setVariables(TACSBVec* q, TACSBVec* qdot = NULL, TACSBVec* qddot = NULL);
  void getVariables(TACSBVec* q, TACSBVec* qdot = NULL, TACSBVec* qddot = NULL);
  void zeroVariables();
  void setDesignVars(TACSBVec* dvs);
  void getDesignVars(TACSBVec* dvs);
  void applyBCs(TACSVec* vec);
  void applyBCs(TACSMat* mat);

  void assembleRes(TACSBVec* residual, const TacsScalar lambda = 1.0,
                   const bool applyBCs = true);
  void assembleJacobian(TacsScalar alpha, TacsScalar beta, TacsScalar gamma,
                        TACSBVec* residual, TACSMat* A,
                        MatrixOrientation matOr = TACS_MAT_NORMAL,
                        const TacsScalar lambda = 1.0, const bool applyBCs = true);
  // Stiffness, mass and geometric stiffness (at the current state) matrices
  // and their combinations (src/TACSAssembler.cpp:4418, :4519)
  void assembleMatType(ElementMatrixType matType, TACSMat* A,
                       MatrixOrientation matOr = TACS_MAT_NORMAL, const TacsScalar lambda = 1.0,
                       const bool applyBCs = true);
  void assembleMatCombo(ElementMatrixType matTypes[], TacsScalar scale[], int nmats, TACSMat* A,
                        MatrixOrientation matOr = TACS_MAT_NORMAL, const TacsScalar lambda = 1.0,
                        const bool applyBCs = true);
  void evalFunctions(int numFuncs, TACSFunction** funcs, TacsScalar* funcVals);
  void addDVSens(TacsScalar coef, int numFuncs, TACSFunction** funcs, TACSBVec** dfdx);
  void addSVSens(TacsScalar alpha, TacsScalar beta, TacsScalar gamma, int numFuncs,
                 TACSFunction** funcs, TACSBVec** dfdu);
  void addAdjointResProducts(TacsScalar scale, int numAdjoints, TACSBVec** adjoint,
                             TACSBVec** dfdx, const TacsScalar lambda = 1.0);
  void addXptSens(TacsScalar coef, int numFuncs, TACSFunction** funcs, TACSBVec** dfdXpt);
  void addAdjointResXptSensProducts(TacsScalar scale, int numAdjoints, TACSBVec** adjoint,
                                    TACSBVec** dfdXpt, const TacsScalar lambda = 1.0);

  // Matrix-free products (src/TACSAssembler.cpp:5502-5717).  data and temp
  // are device arrays of getMatrixFreeDataSize's sizes (doubles): data holds
  // per quadrature point the inverse Jacobian and w det(J) t behind the
  // coefficients of the stiffness and mass terms; temp receives the element
  // products before their node sums.
  void getMatrixFreeDataSize(ElementMatrixType matType, int* data_size, int* temp_size);
  void assembleMatrixFreeData(ElementMatrixType matType, TacsScalar alpha, TacsScalar beta,
                              TacsScalar gamma, TacsScalar data[]);
  void addMatrixFreeVecProduct(ElementMatrixType matType, const TacsScalar data[],
                               TacsScalar temp[], TACSBVec* x, TACSBVec* y,
                               MatrixOrientation matOr = TACS_MAT_NORMAL,
                               const TacsScalar lambda = 1.0, const bool applyBCs = true);
  // y (+)= the product on device vectors; zeroes the Dirichlet rows of y when bcs
  int matFreeProduct(const double* data, double* temp, const double* x, double* y,
                     int accumulate, int bcs);

  // Arithmetic of the element kernels that build the Jacobian of
  // assembleJacobian: 64 (default) or 32.  With 32 the element
  // matrices are computed in FP32 (shells: TACSQuad4Shell's MITC algebra;
  // solids: the isotropic hexahedra) and summed into the block-CSR values in
  // FP64; residuals, functions and sensitivities stay FP64.  No counterpart
  // in the reference (its TacsScalar is double throughout).
  int setElementPrecision(int prec);
  int getElementPrecision() { return elemPrec; }
  // Solid Jacobian kernel: 0 picks by a measured table, 1 the node-centric
  // row kernel, 2 the element blocks + row gather (isotropic material only).
  int setJacobianKernel(int kernel);
  int jacobianKernelInUse();  // 1 or 2: what assembleJacobian launches now

  // Device view of the element-level state used by the kernels
  const double* devNodes() { return d_X; }
  const double* devVars() { return d_u; }
  int getElementOrder() { return order; }
  double lastJacobianMs() { return jac_ms; }
  int status();  // reads and clears the kernel status word

  // internals
  int elementVectorsToNodes(const double* ev, double* out, int accumulate);
  int shellKernel(int prec, int mode, const double* psi, double scale, double ks_weight,
                  double ks_max, double ks_sum, double safety, void* ke, double* ev, double* es);
  void addAuxResidual(double* r, TacsScalar lambda);
  int scatterElementScalarsToDVs(const double* es, double* dfdx);
  int uploadAux();

  enum { FAMILY_SOLID = 0, FAMILY_SHELL = 1 };
  int family;
  int varsPerNode, numNodes, numElements, npe, order, initialized;
  int numDesignVars;
  int* h_ptr;
  int* h_conn;
  TACSElement** elements;
  TACSBcMap* bcMap;
  TACSAuxElements* auxElements;
  int aux_version;
  SolidMaterial mat;
  ShellMaterial smat;
  double* d_ke;  // shell element matrices [nelem][24][24]
  double* d_qgeo;  // solid quadrature geometry [nelem][npe][10] (Jacobian scratch)
  double* d_qstress;  // solid quadrature stresses [nelem][npe][6] (geometric stiffness)
  // device data
  int *d_conn, *d_n2e_ptr, *d_n2e, *d_bcvars, *d_bcnodes, *d_bcmask;
  double* d_bcvals;
  int nbc_nodes;
  int *d_dvnum, *d_dv2e_ptr, *d_dv2e;
  double *d_tfixed, *d_telem;
  void* d_mtab;  // per-element material table (when elements differ), or NULL
  int* d_emat;
  double *d_X, *d_u;
  double *d_ev, *d_es;
  int *d_aux_elem, *d_aux_face, *d_n2a_ptr, *d_n2a;
  double *d_aux_trac, *d_aux_ev, *d_aux_pres;
  int naux;
  int naux_pressure;  // TACSPressure3D aux elements among them (solids)
  int* d_status;
  void* d_red;
  double* d_scalar;
  int *h_rowp, *h_cols;
  long nnzb;
  double jac_ms;
  TACSComm comm;
  int elemPrec, jacKernel;
  // element-block path: every element's npe x npe 3x3 blocks, gathered per
  // row through the node -> element lists
  void* d_keb;
  // Jacobian launch choice for the isotropic solid family: rows-kernel
  // variants 0 .. nv-1 (nv = tacs_hip_solid_jacobian_nvariants(order)), nv
  // for the element-block form; -1 for the tuning table's.  autotune()
  // times every rows variant and (Jacobian kernel setting 0) the element-
  // block form on mat (reps assemblies each), keeps the fastest and returns
  // the number of choices timed (ms: their average times).  Rows variants
  // give identical values; the element-block form agrees to rounding.  No
  // counterpart in the reference.
  int jacVar;
  int autotune(TACSParallelMat* mat, int reps, double* ms, int maxv);
  int jacobianVariant();
};

// src/TACSMatrixFreeMat.h: the Jacobian (or stiffness / mass matrix) of the
// assembler's solid elements applied element by element from quadrature-
// point data, without an assembled matrix.  mult gives K x with the
// Dirichlet rows of the result zeroed (TACSAssembler::
// addMatrixFreeVecProduct with applyBCs).  Shell elements contribute
// nothing, as TACSElement's default getMatVecProductData.
class TACSMatrixFreeMat : public TACSMat {
 public:
  TACSMatrixFreeMat(TACSAssembler* assembler);
  ~TACSMatrixFreeMat();
  void assembleMatrixFreeData(ElementMatrixType matType, double alpha, double beta,
                              double gamma);
  TACSVec* createVec();
  void mult(TACSVec* x, TACSVec* y);
  void multTranspose(TACSVec* x, TACSVec* y);
  void zeroEntries() {}
  void applyBCs(TACSBcMap* bcmap) {}
  const char* getObjectName() { return "TACSMatrixFreeMat"; }
  // y = K x on device vectors of length length()
  void mult64(const double* x, double* y);
  size_t length();
  TACSAssembler* assembler;
  ElementMatrixType matType;
  int data_size, temp_size;
  double* data;
  double* temp;
};

#endif
