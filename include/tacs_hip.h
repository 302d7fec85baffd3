/*
  C ABI of the MI355X implementation of TACS's FE assembly / block-CSR /
  GMRES / adjoint path (library: tacs_amd/lib/libtacs_hip.so).

  Two layers, both plain C (pointers, sizes, status ints; no C++ or torch
  types):

  1. Kernel entry points (tacs_hip_*): stateless launchers over raw device
     pointers on an explicit HIP stream.  Each replaces one CPU loop of the
     reference, cited on its declaration.  They return 0 or a hipError_t /
     TACS_HIP_ERR_* code for launch problems; in-kernel problems are reported
     through the device status word passed to them.

  2. Object entry points (tacs_*): opaque, reference-counted handles to the
     C++ classes of TACSHip.h.  They mirror, one to one, the methods the
     reference's Cython layer binds (tacs/cpp_headers/TACS.pxd:321-428,
     tacs/TACS.pyx, tacs/elements.pyx, tacs/constitutive.pyx,
     tacs/functions.pyx).  Handles start with refcount 0 like TACSObject
     (src/TACSObject.h:112-140); tacs_incref/tacs_decref manage them.
     Functions returning int return 0 on success and non-zero on failure, with
     a message in tacs_last_error() (the reference prints to stderr and
     returns int codes, e.g. TACSAssembler::initialize src/TACSAssembler.cpp:2519).
*/
#ifndef TACS_HIP_C_H
#define TACS_HIP_C_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Packed solid material: the reference's C[21] tangent stiffness
   (TACSMaterialProperties::evalTangentStiffness3D
   src/constitutive/TACSMaterialProperties.cpp:270), density and yield
   stress, plus the Lame constants used when the material is isotropic.
   Per-element isotropic materials (TACSAssembler::setElements with one
   constitutive object per element, src/TACSAssembler.cpp:623): when
   elem_props != NULL (device array, 4 per element: lam, mu, rho, ys),
   element e uses those constants instead of lam, mu, rho, ys. */
typedef struct SolidMaterial {
  double C[21];
  double rho;
  double ys;
  int isotropic;
  double lam, mu;
  const double* elem_props;
} SolidMaterial;

/* Isotropic shell material of TACSIsoShellConstitutive
   (src/constitutive/TACSIsoShellConstitutive.cpp:153-189): plane-stress Q
   (TACSMaterialProperties::evalTangentStiffness2D
   src/constitutive/TACSMaterialProperties.cpp:323), shear correction kcorr,
   drilling regularization (TACSShellConstitutive.cpp:61), yield stress and
   density, and the thickness offset toff (tOffset: B = -toff t^2 Q,
   D = t^3 (1/12 + toff^2) Q, faces at t (1/2 - toff) and -t (1/2 + toff),
   mass moments rho t, -rho t^2 toff, rho t^3 (toff^2 + 1/12)).  The element
   frame: ref_axis = 0 for TACSShellNaturalTransform, 1 for
   TACSShellRefAxisTransform with the unit axis
   (src/elements/shell/TACSShellElementTransform.h:95).  Per-element
   materials: when elem_mat != NULL (device arrays), element e uses
   table[elem_mat[e]] (Q, kcorr, drill_reg, ys, rho, toff; the frame stays
   this struct's), as TACSAssembler::setElements allows one constitutive
   object per element (src/TACSAssembler.cpp:623). */
typedef struct ShellMaterial {
  double Q[6];
  double kcorr;
  double drill_reg;
  double ys;
  double rho;
  double axis[3];
  int ref_axis;
  double toff;
  const struct ShellMaterial* table;
  const int* elem_mat;
} ShellMaterial;

/* ------------------------------------------------------------------------
   1. Kernel entry points
   ------------------------------------------------------------------------ */

/* Jacobian of the solid element family written straight into block-CSR
   (b = 3), Dirichlet rows applied when bcvars != NULL.
   Replaces the element loop of TACSAssembler::assembleJacobian
   (src/TACSAssembler.cpp:4352-4383) -> TACSElement3D::addJacobian
   (src/elements/TACSElement3D.cpp:176) -> TACSAssembler::addMatValues
   (src/TACSAssembler.h:545) -> BCSRMat::addRowValues
   (src/bpmat/BCSRMat.cpp:1800) and TACSParallelMat::applyBCs
   (src/bpmat/TACSParallelMat.cpp:343).
   order: 2 (8-node) or 3 (27-node) hexahedra; n2e: node->element CSR with
   entries (element << 5 | local node); A (FP64) and/or A32 (FP32) outputs;
   G: scratch of 10 * order^3 * nelem doubles (the per-quadrature-point
   inverse Jacobians and weights, computed once per call). */
int tacs_hip_solid_jacobian(int order, int nnodes, int nelem, const int* n2e_ptr,
                            const int* n2e, const int* conn, const double* X,
                            const double* telem, const SolidMaterial* mat, double alpha,
                            double gamma, const int* rowp, const int* cols, const int* bcvars,
                            double* G, double* A, float* A32, int* status, hipStream_t stream);
/* Launch variants of the rows kernel above for the isotropic stiffness /
   mass (wavefronts per workgroup, elements per geometry chunk), all with the
   same sums: 8-node elements 0 (4, 8), 1 (2, 8), 2 (8, 4), 3 (4, 4),
   4 (4, 16); 27-node elements 0 (2, 8), 1 (1, 8), 2 (2, 4), 3 (1, 4).
   tacs_hip_solid_jacobian launches the variant of the tuning table. */
int tacs_hip_solid_jacobian_nvariants(int order);
int tacs_hip_solid_jacobian_variant(int variant, int order, int nnodes, int nelem,
                                    const int* n2e_ptr, const int* n2e, const int* conn,
                                    const double* X, const double* telem,
                                    const SolidMaterial* mat, double alpha, double gamma,
                                    const int* rowp, const int* cols, const int* bcvars,
                                    double* G, double* A, float* A32, int* status,
                                    hipStream_t stream);

/* Element-centric Jacobian of the isotropic solid family, the default path
   of TACSAssembler::assembleJacobian (same reference loop as
   tacs_hip_solid_jacobian).  tacs_hip_solid_elem_blocks writes every
   element's 3x3 blocks K_ab = alpha K + gamma M (local nodes a, b; blocks
   row-major) to ke[e][a*npe + b][9], computed in FP64 (prec 64, ke double)
   or FP32 (prec 32, ke float) arithmetic; tacs_hip_gather_solid_blocks sums
   them into the block-CSR values in FP64 (A and/or its FP32 copy A32): row
   (node) i adds, over its incident elements in the order of the node ->
   element list n2e[n2e_ptr[i] .. n2e_ptr[i+1]) (entries (e << 5) | a,
   ascending e), block b of the element's block row a into the slot of column
   conn[e][b], with the Dirichlet rows of bcvars applied.  order 2 or 3;
   isotropic material; status receives TACS_HIP_ERR_ROW_TOO_LONG (more than
   64 / 125 column blocks) or TACS_HIP_ERR_MISSING_ENTRY. */
int tacs_hip_solid_elem_blocks(int order, int prec, int nelem, const int* conn, const double* X,
                               const double* telem, const SolidMaterial* mat, double alpha,
                               double gamma, void* ke, hipStream_t stream);
int tacs_hip_gather_solid_blocks(int order, int prec, int nrows, const int* rowp,
                                 const int* cols, const int* n2e_ptr, const int* n2e,
                                 const int* conn, const void* ke, const int* bcvars, double* A,
                                 float* A32, int* status, hipStream_t stream);

/* Element-matrix combinations of the solid family straight into block-CSR:
   kc K + mc M + gc G(u) with K the stiffness, M the consistent mass and G
   the geometric stiffness at the state u (replaces the element loops of
   TACSAssembler::assembleMatType src/TACSAssembler.cpp:4418 and
   assembleMatCombo :4519 over TACSElement3D::getMatType
   src/elements/TACSElement3D.cpp:316).  Arguments as
   tacs_hip_solid_jacobian; S: scratch of 6 * order^3 * nelem doubles (the
   per-point stresses), needed with u when gc != 0. */
int tacs_hip_solid_mat_type(int order, int nnodes, int nelem, const int* n2e_ptr,
                            const int* n2e, const int* conn, const double* X, const double* u,
                            const double* telem, const SolidMaterial* mat, double kc, double mc,
                            double gc, const int* rowp, const int* cols, const int* bcvars,
                            double* G, double* S, double* A, float* A32, int* status,
                            hipStream_t stream);

/* Matrix-free Jacobian data (replaces TACSAssembler::assembleMatrixFreeData
   src/TACSAssembler.cpp:5563 over TACSElement3D::getMatVecProductData
   src/elements/TACSElement3D.cpp:379): data[0..3] = [kc, mc, 0, 0], then per
   (element, quadrature point) J^-1 (9 values) and w det(J) t_e; in all
   4 + 10 * order^3 * nelem doubles.  kc multiplies the stiffness, mc the
   consistent mass (alpha and gamma for TACS_JACOBIAN_MATRIX). */
int tacs_hip_solid_matfree_data(int order, int nelem, const int* conn, const double* X,
                                const double* telem, double kc, double mc, double* data,
                                hipStream_t stream);

/* Element part of the matrix-free product (TACSAssembler::
   addMatrixFreeVecProduct src/TACSAssembler.cpp:5645 over
   TACSElementBasis::addMatVecProduct basis/TACSElementBasis.cpp:940):
   ev[nelem * 3*npe] = per-element products of the data above with x; the
   node sums follow with tacs_hip_gather_elem_vectors. */
int tacs_hip_solid_matfree_ev(int order, int nelem, const int* conn, const double* data,
                              const SolidMaterial* mat, const double* x, double* ev,
                              hipStream_t stream);

/* Element vectors ev[nelem * 3*npe]:
   mode 0 residual (TACSElement3D::addResidual src/elements/TACSElement3D.cpp:139),
   mode 1 compliance d/du (TACSCompliance::getElementSVSens
          src/functions/TACSCompliance.cpp:95), coef = alpha,
   mode 2 KS-failure d/du (TACSKSFailure::getElementSVSens
          src/functions/TACSKSFailure.cpp:221). */
int tacs_hip_solid_elem_vectors(int order, int mode, int nelem, const int* conn,
                                const double* X, const double* u, const double* telem,
                                const SolidMaterial* mat, double coef, double ks_weight,
                                double ks_max, double ks_sum, double safety, double* ev,
                                hipStream_t stream);

/* Traction and pressure aux-element vectors: aux element q adds the traction
   aux_trac[3q..3q+2] (TACSTraction3D::addResidual
   src/elements/TACSTraction3D.cpp:87, coordinate components) and the
   pressure aux_pres[q] (TACSPressure3D::addResidual
   src/elements/TACSPressure3D.cpp:60: traction -p n, n the outward unit
   normal of face aux_face[q]); aux_pres may be NULL. */
int tacs_hip_solid_traction_vectors(int order, int naux, const int* aux_elem,
                                    const int* aux_face, const double* aux_trac,
                                    const double* aux_pres, const int* conn, const double* X,
                                    double lambda, double* ev, hipStream_t stream);

/* Node-coordinate sensitivities as element vectors out[nelem * 3*npe]:
   mode 0 compliance dC/dX at fixed u (TACSAssembler::addXptSens
          src/TACSAssembler.cpp:4858 -> TACSCompliance::getElementXptSens
          src/functions/TACSCompliance.cpp:129), coef = scale,
   mode 1 KS failure dKS/dX at fixed u (TACSKSFailure::getElementXptSens),
   mode 2 scale * psi^T dR/dX (TACSAssembler::addAdjointResXptSensProducts
          :5136 -> TACSElement3D::addAdjResXptProduct
          src/elements/TACSElement3D.cpp:260). */
int tacs_hip_solid_xpt_vectors(int order, int mode, int nelem, const int* conn, const double* X,
                               const double* u, const double* psi, const double* telem,
                               const SolidMaterial* mat, double scale, double ks_weight,
                               double ks_max, double ks_sum, double safety, double* out,
                               hipStream_t stream);
/* scale * psi^T dR/dX of traction aux elements, exact derivative (the
   reference differences TACSTraction3D's residual forward,
   TACSElement::addAdjResXptProduct src/elements/TACSElement.cpp:270). */
int tacs_hip_solid_traction_xpt_vectors(int order, int naux, const int* aux_elem,
                                        const int* aux_face, const double* aux_trac,
                                        const int* conn, const double* X, const double* psi,
                                        double scale, double* out, hipStream_t stream);

/* Node gather of element vectors in ascending element order: the reverse of
   TACSBVec::setValues ADD (src/bpmat/TACSBVec.cpp:757). nv = 24 or 81. */
int tacs_hip_gather_elem_vectors(int nv, int nnodes, const int* n2e_ptr, const int* n2e,
                                 const double* ev, int accumulate, double* out,
                                 hipStream_t stream);

/* Element scalars out[nelem]:
   mode 0 compliance integrand (TACSCompliance::elementWiseEval :69),
   mode 1 KS max, mode 2 KS sum (TACSKSFailure::elementWiseEval :125),
   mode 3 dC/dt (TACSCompliance::addElementDVSens :162),
   mode 4 psi^T dR/dt (TACSElement3D::addAdjResProduct
          src/elements/TACSElement3D.cpp:227). */
int tacs_hip_solid_elem_scalars(int order, int mode, int nelem, const int* conn,
                                const double* X, const double* u, const double* psi,
                                const double* telem, const SolidMaterial* mat, double scale,
                                double ks_weight, double ks_max, double safety, double* out,
                                hipStream_t stream);

/* t_e = dv[tNum_e] or the element's fixed t (TACSSolidConstitutive::
   setDesignVars src/constitutive/TACSSolidConstitutive.cpp:70 applied to
   every element, as TACSAssembler::setDesignVars src/TACSAssembler.cpp:3242). */
int tacs_hip_gather_design_vars(int nelem, const int* dvnum, const double* tfixed,
                                const double* dv, double* telem, hipStream_t stream);
/* dfdx[d] += sum of es[e] over the elements e of design variable d, in
   ascending element order (dfdx->setValues ADD in TACSAssembler::addDVSens
   src/TACSAssembler.cpp:4767 and addAdjointResProducts :5033). */
int tacs_hip_scatter_dv_sens(int ndv, const int* dv2e_ptr, const int* dv2e, const double* es,
                             double* dfdx, hipStream_t stream);
/* Dirichlet values on a vector (TACSBVec::applyBCs src/bpmat/TACSBVec.cpp:546). */
int tacs_hip_apply_bc_vec(int b, int nbc, const int* nodes, const int* masks, const double* vals,
                          const double* u, double lambda, double* x, hipStream_t stream);

/* Linear MITC4 shell (TACSQuad4Shell = TACSShellElement<TACSQuadLinearQuadrature,
   TACSShellQuadBasis<2>, TACSLinearizedRotation, TACSShellLinearModel>,
   src/elements/shell/TACSShellElementDefs.h:18-20, natural transform) over
   nelem 4-node elements; node-major variables, 6 per node.  mode:
   0 element stiffness ke[nelem][24][24] (row-major, scaled by `scale`;
     TACSShellElement::addJacobian TACSShellElement.h:435) and, when ev is
     given, the element residual K u (addResidual :295); prec 64 or 32
     (ke in FP64 or FP32; the other modes are FP64 only),
   1 element residual ev (addResidual :295, static state),
   2 compliance integrand es (TACSCompliance::elementWiseEval
     src/functions/TACSCompliance.cpp:69, strain energy density),
   3 / 4 KS max / KS sum es (TACSKSFailure::elementWiseEval
     src/functions/TACSKSFailure.cpp:125; failure of
     TACSIsoShellConstitutive::evalFailure :254),
   5 KS d/du ev (TACSKSFailure::getElementSVSens :221 ->
     TACSShellElement::addPointQuantitySVSens :1353), coef = scale,
   6 KS d/dt es (TACSKSFailure::addElementDVSens -> addPointQuantityDVSens
     :1193 -> TACSIsoShellConstitutive::addFailureDVSens :381),
   7 scale * psi^T dR/dt es (TACSShellElement::addAdjResProduct :917 ->
     TACSIsoShellConstitutive::addStressDVSens :229),
   8 compliance d/du ev, 9 compliance d/dt es (same call chains as 5, 6). */
int tacs_hip_shell_elem_kernel(int prec, int mode, int nelem, const int* conn, const double* X,
                               const double* u, const double* psi, const double* telem,
                               const ShellMaterial* mat, double scale, double ks_weight,
                               double ks_max, double ks_sum, double safety, void* ke, double* ev,
                               double* es, hipStream_t stream);

/* Element matrices ke = kscale K + mscale M + gscale G of the Quad4 shell
   family (24 x 24 per element, FP64 when prec = 64, FP32 when 32): K the
   tangent stiffness, M the consistent mass (mass moments of
   TACSIsoShellConstitutive::evalMassMoments
   src/constitutive/TACSIsoShellConstitutive.cpp:120), G the geometric
   stiffness along the state u (FP64 only, u required: the derivative along
   u of TACSShellNonlinearModel's tangent stiffness at the zero state), and
   the FP64 element residuals ev = K u when ev != NULL (prec 64 only; u =
   NULL is the zero state).  Replaces TACSShellElement::getMatType
   (src/elements/shell/TACSShellElement.h:644: stiffness alpha = 1, mass
   gamma = 1 through addJacobian :447, geometric stiffness :661-908) and the
   gamma term of TACSShellElement::addJacobian :605-617 with
   TACSLinearizedRotation::addDirectorJacobian
   src/elements/shell/TACSDirector.h:369. */
int tacs_hip_shell_mat_type(int prec, int nelem, const int* conn, const double* X,
                            const double* u, const double* telem, const ShellMaterial* mat,
                            double kscale, double mscale, double gscale, void* ke, double* ev,
                            hipStream_t stream);

/* Node-coordinate derivatives of the Quad4 shell family, exact (forward-mode
   dual numbers): es[12 e + k] = d/dX_k of element e's term (coordinate k =
   3 node + component), for node gathering (tacs_hip_gather_elem_vectors_b
   with b = 3, nv = 12).  mode 0: scale psi^T R_e at the state u
   (TACSAssembler::addAdjointResXptSensProducts src/TACSAssembler.cpp:5136
   over TACSElement::addAdjResXptProduct src/elements/TACSElement.cpp:270);
   1: the compliance term scale sum_q detXd e.s (TACSCompliance::
   getElementXptSens src/functions/TACSCompliance.cpp:129); 2: the
   continuous KS failure term with the aggregate's ks_max and ks_sum
   (TACSKSFailure::getElementXptSens src/functions/TACSKSFailure.cpp:289,
   through addPointQuantityXptSens src/elements/TACSElement.cpp:746). */
int tacs_hip_shell_xpt_kernel(int mode, int nelem, const int* conn, const double* X,
                              const double* u, const double* psi, const double* telem,
                              const ShellMaterial* mat, double scale, double ks_weight,
                              double ks_max, double ks_sum, double safety, double* es,
                              hipStream_t stream);
/* out[12 q + k] = scale d/dX_k (psi . F_q) for the pressure loads F_q of the
   listed shell elements (TACSShellPressure through
   TACSElement::addAdjResXptProduct). */
int tacs_hip_shell_pressure_xpt(int naux, const int* aux_elem, const double* pressure,
                                const int* conn, const double* X, const double* psi, double scale,
                                double* out, hipStream_t stream);

/* Sum element matrices ke (FP64 when prec = 64, FP32 when 32) into b x b
   block-CSR rows in ascending element order, Dirichlet rows fused
   (TACSAssembler::addMatValues src/TACSAssembler.h:545 -> BCSRMat::addRowValues
   src/bpmat/BCSRMat.cpp:1800; TACSParallelMat::applyBCs
   src/bpmat/TACSParallelMat.cpp:343).  b = 6, npe = 4. */
int tacs_hip_gather_elem_matrices(int b, int npe, int prec, int nrows, const int* n2e_ptr,
                                  const int* n2e, const int* conn, const void* ke, const int* rowp,
                                  const int* cols, const int* bcvars, double* A, float* A32,
                                  hipStream_t stream);

/* Node gather of element vectors with b values per node (the reverse of
   TACSBVec::setValues ADD src/bpmat/TACSBVec.cpp:757): (b, nv) = (6, 24),
   (3, 24), (3, 81), (3, 12). */
int tacs_hip_gather_elem_vectors_b(int b, int nv, int nnodes, const int* n2e_ptr, const int* n2e,
                                   const double* ev, int accumulate, double* out,
                                   hipStream_t stream);

/* Uniform-pressure shell aux-element vectors ev[naux][24]
   (TACSShellPressure::addResidual src/elements/shell/TACSShellPressure.h:51). */
int tacs_hip_shell_pressure_vectors(int naux, const int* aux_elem, const double* pressure,
                                    const int* conn, const double* X, double lambda, double* ev,
                                    hipStream_t stream);

/* y (+)= A x, b x b block-CSR, b = 3 or 6 (BCSRMat::mult
   src/bpmat/BCSRMat.cpp:1226 -> BCSRMatVecMult3 src/bpmat/BCSRMatMult3.cpp:28,
   BCSRMatVecMult6 src/bpmat/BCSRMatMult6.cpp:82). */
int tacs_hip_bsr_spmv(int b, int nrows, const int* rowp, const int* cols, const double* A,
                      const double* x, double* y, int add, hipStream_t stream);
/* FP32 variant of tacs_hip_bsr_spmv (mixed-precision inner solve). */
int tacs_hip_bsr_spmv_f32(int b, int nrows, const int* rowp, const int* cols, const float* A,
                          const float* x, float* y, int add, hipStream_t stream);
/* y (+)= A32 x with FP32 matrix values and FP64 vectors and arithmetic: the
   fixed FP32-rounded operator of the deflated mixed-precision GMRES. */
int tacs_hip_bsr_spmv_mixed(int b, int nrows, const int* rowp, const int* cols, const float* A,
                            const double* x, double* y, int add, hipStream_t stream);
/* y = D32^{-1} x per block with FP32 inverse blocks and FP64 vectors. */
int tacs_hip_bsr_diag_apply_mixed(int b, int nrows, const float* Dinv, const double* x, double* y,
                                  hipStream_t stream);
/* Y = A X for nrhs right-hand sides stored node-major
   X[(3*node + c)*nrhs + j]; the matrix is read once for all of them
   (the reference multiplies one vector at a time, BCSRMatVecMult3). */
int tacs_hip_bsr3_spmm(int nrows, int nrhs, const int* rowp, const int* cols, const double* A,
                       const double* X, double* Y, hipStream_t stream);
/* Sliced ELL copy of a 3x3 or 6x6 BCSR matrix for the product y = A x (no
   counterpart in the reference; BCSRMat::mult src/bpmat/BCSRMat.cpp:1226 is
   what it computes).  Slices of 64 rows; slice s is width[s] blocks wide and
   starts at slot coff[s] (coff has nslices + 1 entries); slot t of row
   64 s + l holds its column in cs[coff[s] + 64 t + l] and value q of the block
   (row-major) in vs[b^2 coff[s] + 64 (b^2 t + q) + l]; padding slots are zero
   blocks in the row's own column.
   sell_slots:  host: widths and offsets from the host row pointer (either
                output may be NULL); returns the number of slots;
   sell_fill:   values (FP64 or FP32, prec 64/32) and, if cs != NULL, columns
                from the BCSR arrays;
   sell_fill2:  the FP64 copy v64 and/or the FP32 copy v32 (either may be
                NULL) in one pass over the BCSR values;
   sell_spmv:   values prec_vals (64/32), vectors prec_vecs (64/32). */
size_t tacs_hip_sell_slots(int nrows, const int* rowp_host, int* width, long* coff);
int tacs_hip_sell_fill(int b, int prec, int nrows, const int* rowp, const int* cols,
                       const int* width, const long* coff, const double* A, void* vs, int* cs,
                       hipStream_t stream);
int tacs_hip_sell_fill2(int b, int nrows, const int* rowp, const int* cols, const int* width,
                        const long* coff, const double* A, double* v64, float* v32, int* cs,
                        hipStream_t stream);
/* sell_fill3: sell_fill2 plus a bfloat16 copy v16 (the upper 16 bits of the
   FP32 value, rounded to nearest even), any of the three may be NULL. */
int tacs_hip_sell_fill3(int b, int nrows, const int* rowp, const int* cols, const int* width,
                        const long* coff, const double* A, double* v64, float* v32,
                        unsigned short* v16, int* cs, hipStream_t stream);
/* Launch variants of sell_fill3, all writing the same arrays: 3x3 blocks
   0 (one lane per row) and 1-3 (one wave per slice staging 4, 7 or 7
   slots of all 64 rows in LDS: contiguous reads of the BCSR runs, 64-entry
   stores; 16, 32 or 32 rows of loads in flight; non-temporal value loads
   in 1 and 2, non-temporal stores in 2 and 3); 6x6 blocks variant 0 only.
   sell_fill3 launches variant 3 for 3x3 blocks.  No counterpart in the reference. */
int tacs_hip_sell_fill_nvariants(int b);
int tacs_hip_sell_fill_variant(int variant, int b, int nrows, const int* rowp, const int* cols,
                               const int* width, const long* coff, const double* A, double* v64,
                               float* v32, unsigned short* v16, int* cs, hipStream_t stream);
int tacs_hip_sell_spmv(int b, int prec_vals, int prec_vecs, int nrows, const int* width,
                       const long* coff, const int* cs, const void* vs, const void* x, void* y,
                       hipStream_t stream);
/* Launch variants of the sliced ELL product, all with the same sums (slots
   per unrolled pass U, non-temporal value loads on or off): 3x3 blocks U in
   {1, 2, 4, 8} with (variants 0-3) or without (4-7) non-temporal loads, 6x6
   blocks U in {1, 2} with (0, 1) or without (2, 3).  tacs_hip_sell_spmv
   launches the variant of the tuning table. */
int tacs_hip_sell_spmv_nvariants(int b);
/* One Chebyshev step for A z = rhs fused into the sliced ELL product (FP64
   vectors; prec_vals 64, 32 or 16 (bfloat16 values), Dinv the b x b inverse
   diagonal blocks in the same precision, FP32 for bfloat16 values): per
   row, a = (A z)_row, w = D^{-1} (rhs - a),
   d = c1 d + c2 w, zout = z + d.  zout must differ from z.  variant as for
   tacs_hip_sell_spmv_variant. */
int tacs_hip_sell_cheb_step(int variant, int b, int prec_vals, int nrows, const int* width,
                            const long* coff, const int* cs, const void* vs, const double* z,
                            double* zout, const double* rhs, const void* Dinv, double* d,
                            double c1, double c2, hipStream_t stream);
/* The same step in three-term form: zprev holds the iterate before z on
   entry and receives zout = (1 + c1) z - c1 zprev + c2 D^{-1} (rhs - A z)
   (the two-term step with d = z - zprev; no d vector is read or written).
   zprev must differ from z.  No counterpart in the reference. */
int tacs_hip_sell_cheb_step3(int variant, int b, int prec_vals, int nrows, const int* width,
                             const long* coff, const int* cs, const void* vs, const double* z,
                             double* zprev, const double* rhs, const void* Dinv, double c1,
                             double c2, hipStream_t stream);
int tacs_hip_sell_spmv_variant(int variant, int b, int prec_vals, int prec_vecs, int nrows,
                               const int* width, const long* coff, const int* cs, const void* vs,
                               const void* x, void* y, hipStream_t stream);

/* Tuning table (no counterpart in the reference beyond the block-size choice
   of BCSRMat::initBlockImpl, src/bpmat/BCSRMat.cpp:1028): (kernel, shape,
   size class) -> launch variant, -1 for the built-in default.  Every variant
   of a kernel gives the same results; tables are filled by measurement
   (tacs_amd/autotune.py, or the per-object tacs_mat_autotune /
   tacs_assembler_autotune / tacs_batched_autotune_spmm).
   kernel TACS_TUNE_SELL_SPMV: shape 10 b + kind (kind 0: FP64 values and
     vectors, 1: FP32 values with FP64 vectors, 2: FP32 values and vectors);
   kernel TACS_TUNE_SPMM (tacs_hip_bsr3_spmm*): shape R (FP64 values) or
     100 + R (FP32 values);
   kernel TACS_TUNE_JACOBIAN (solid Jacobian rows kernel): shape = order.
   Size class from the number of block rows: 0 below 2^15, 1 below 2^19,
   2 from 2^19.  Returns 0 or TACS_HIP_ERR_ARG. */
#define TACS_TUNE_SELL_SPMV 0
#define TACS_TUNE_SPMM 1
#define TACS_TUNE_JACOBIAN 2
int tacs_tune_set(int kernel, int shape, int size_class, int variant);
int tacs_tune_get(int kernel, int shape, int size_class);
int tacs_tune_size_class(long rows);
int tacs_tune_variant(int kernel, int shape, long rows);
void tacs_tune_reset(void);
/* Diagonal block positions; inverse diagonal blocks (BCSRMat::factorDiag
   src/bpmat/BCSRMat.cpp:1527); y = D^{-1} x. */
/* x[b i + r] = 0 where row r of block row i is an identity row (the
   Dirichlet rows of applyBCs: diagonal value 1, every other value 0). */
int tacs_hip_bsr_zero_identity_rows(int b, int nrows, const int* rowp, const int* cols,
                                    const double* A, double* x, hipStream_t stream);
int tacs_hip_bsr_find_diag(int nrows, const int* rowp, const int* cols, int* diag,
                           hipStream_t stream);
int tacs_hip_bsr_diag_inverse(int b, int nrows, const int* diag, const double* A, double* Dinv,
                              float* Dinv32, int* status, hipStream_t stream);
int tacs_hip_bsr_diag_apply(int b, int nrows, const double* Dinv, const double* x, double* y,
                            hipStream_t stream);
int tacs_hip_bsr_diag_apply_f32(int b, int nrows, const float* Dinv, const float* x, float* y,
                                hipStream_t stream);
/* Dirichlet rows of an assembled matrix (BCSRMat::zeroRow
   src/bpmat/BCSRMat.cpp:2027). */
int tacs_hip_bsr_zero_rows(int b, int nbc, const int* nodes, const int* vars, const int* rowp,
                           const int* cols, double* A, int ident, hipStream_t stream);
int tacs_hip_convert_f64_f32(size_t n, const double* x, float* y, hipStream_t stream);

/* Vector kernels of TACSBVec (src/bpmat/TACSBVec.cpp: norm :196, dot :237,
   axpy :330, scale :228) and GMRES (src/bpmat/KSM.cpp:785; MGS :526,
   CGS :515).  prec is 64 (double) or 32 (float); reductions are
   deterministic and write their result to device memory; scratch is a
   device buffer of tacs_hip_red_scratch_bytes() bytes, zeroed once. */
size_t tacs_hip_red_scratch_bytes(void);
int tacs_hip_dot(int prec, size_t n, const void* wprev, const double* hprev, void* q,
                 const void* w, void* scratch, double* out, int take_sqrt, double* out_inv,
                 hipStream_t stream);
int tacs_hip_mdot(int prec, size_t n, int k, const void* W, size_t ldw, const void* q,
                  void* scratch, double* h, int accumulate, hipStream_t stream);
/* tacs_hip_mdot without the scratch argument, for callers that cannot
   allocate (code under HIP graph capture, serving loops, harnesses that forbid
   hipMalloc).  No counterpart in the reference.  It allocates nothing: the
   block partials live in a static device array of the library (512 KiB on
   every GPU it runs on, held until the process ends), whose address the first
   call made with a GPU current resolves, so make that first call outside any
   capture.  Limit: 1 <= k <= 32 vectors (n is unbounded); other k, or prec
   other than 64 / 32, return TACS_HIP_ERR_ARG with a message in
   tacs_last_error().  Every call on one GPU shares that array: calls that
   overlap on one GPU (several streams or threads at once) race and give wrong
   dots; calls ordered on one stream are safe.  Nothing in the array needs to
   start at zero (each call writes every partial it reads). */
int tacs_hip_mdot_noalloc(int prec, size_t n, int k, const void* W, size_t ldw, const void* q,
                          double* h, int accumulate, hipStream_t stream);
int tacs_hip_cgs_update(int prec, size_t n, int k, const void* W, size_t ldw, const double* c,
                        const double* c2, void* q, hipStream_t stream);
int tacs_hip_lincomb(int prec_w, int prec_y, size_t n, int k, const void* W, size_t ldw,
                     const double* c, void* y, int add, hipStream_t stream);
int tacs_hip_scale(int prec, size_t n, void* x, const double* sp, double sv, int invert,
                   hipStream_t stream);
int tacs_hip_axpby(int px, int py, size_t n, double a, const void* x, double b, void* y,
                   hipStream_t stream);
int tacs_hip_fill(int prec, size_t n, void* x, double v, hipStream_t stream);
/* x[i] in [-1, 1) from a 64-bit hash of (seed, i): a reproducible start
   vector (power iteration). */
int tacs_hip_fill_random(size_t n, double* x, unsigned long long seed, hipStream_t stream);
int tacs_hip_reduce(size_t n, const double* x, int op, void* scratch, double* out,
                    hipStream_t stream);
int tacs_hip_givens(int i, int m, double* H, double* Qc, double* Qs, double* res, double* st,
                    hipStream_t stream);
int tacs_hip_backsolve(int niters, int m, const double* H, double* res, hipStream_t stream);
int tacs_hip_scalar_op(const double* src, double s, double* dst, hipStream_t stream);

/* out[:, j] = sum_i W[:, i] P[i, j] for j < nout (FP32 basis W, n x nin with
   leading dimension ldw; P column-major nin x nout, FP64 on the device;
   nout <= 32, nin <= 512): the basis change of a deflated GMRES restart.  No
   counterpart in the reference. */
int tacs_hip_basis_combine(size_t n, int nin, const float* W, size_t ldw, const double* P,
                           int nout, float* out, size_t ldo, hipStream_t stream);

/* FP64 basis change out[:, j] (+)= sum_i W[:, i] P[i, j] (nin <= 256,
   nout <= 32), and the tall-skinny product out = X^T Y (n x p and n x q,
   p*q <= 4096, p + q <= 256; scratch of tacs_hip_tsgemm_scratch_bytes()),
   used by the recycling GMRES.  No counterpart in the reference. */
int tacs_hip_basis_combine64(size_t n, int nin, const double* W, size_t ldw, const double* P,
                             int nout, double* out, size_t ldo, int add, hipStream_t stream);
size_t tacs_hip_tsgemm_scratch_bytes(void);
int tacs_hip_tsgemm(size_t n, int p, const double* X, int q, const double* Y, double* scratch,
                    double* out, hipStream_t stream);

/* Batched GMRES kernels over interleaved block vectors V[i*R + j] (dof i,
   right-hand side j, R <= 64; the layout of tacs_hip_bsr3_spmm).  Per
   right-hand side j the Hessenberg, rotations and residual vectors are slices
   with strides sH, sQ, sRes.  No counterpart in the reference (one GMRES per
   right-hand side there, KSM.cpp:785).
   bmdot:        out[j*sR + v] (+)= W_v(:, j) . q(:, j), v < k
   *_shared:     the same with single vectors W_v (W[v*ldw + i]) common to every
                 right-hand side (a recycled space)
   bsmall_gemv:  z[j*sz + v] = sum_{i < cnt[j]} B[j*sB + v + i*ldb] y[j*sy + i], v < k
   bcgs_update:  q(:, j) -= sum_v W_v(:, j) c[j*sR + v]
   blincomb:     y(:, j) (+)= sum_{v < cnt[j]} W_v(:, j) c[j*sR + v]
   badd:         y[j*sy + v] += x[j*sx + v], v < k
   bscale:       x(:, j) *= s[j];   bsqrt_inv: nrm = sqrt(sq), inv = 1/nrm
   bgivens / bbacksolve: the GMRES Givens update / back-substitution per j
   bdiag_apply:  b x b block-Jacobi, b = 3 or 6 (FP64 or FP32 inverse blocks)
   bsr3_spmm_mixed: Y = A32 X with FP32 values and FP64 block vectors
   bbasis_combine: out_c(:, j) (+)= sum_{v < nin} W_v(:, j) P[j*sP + v + c*ldp], c < nout <= 32
   bbroadcast:   X(:, j) = x for every j
   bpack / bunpack: copy one right-hand side into / out of a block vector */
size_t tacs_hip_batch_scratch_bytes(int R);
int tacs_hip_bmdot(size_t n, int R, int k, const double* W, size_t ldw, const double* q,
                   double* scratch, double* out, size_t sR, int accumulate, hipStream_t stream);
int tacs_hip_bmdot_shared(size_t n, int R, int k, const double* W, size_t ldw, const double* q,
                          double* scratch, double* out, size_t sR, int accumulate,
                          hipStream_t stream);
int tacs_hip_bcgs_update(size_t n, int R, int k, const double* W, size_t ldw, const double* c,
                         size_t sR, double* q, hipStream_t stream);
int tacs_hip_bcgs_update_shared(size_t n, int R, int k, const double* W, size_t ldw,
                                const double* c, size_t sR, double* q, hipStream_t stream);
int tacs_hip_blincomb_shared(size_t n, int R, int k, const double* W, size_t ldw, const double* c,
                             size_t sR, double* y, int add, hipStream_t stream);
int tacs_hip_bsmall_gemv(int R, int k, const int* cnt, const double* B, size_t sB, int ldb,
                         const double* y, size_t sy, double* z, size_t sz, hipStream_t stream);
int tacs_hip_blincomb(size_t n, int R, int kmax, const double* W, size_t ldw, const double* c,
                      size_t sR, const int* cnt, double* y, int add, hipStream_t stream);
/* Delayed classical Gram-Schmidt (DCGS2, Bielich et al., Parallel Comput.
   112 (2022)) over the basis [W1 (k1 block vectors), W2 (k2)], stride ldw:
   bdotq:  out_qi[j*sR + v] = W_v(:, j) . q_qi(:, j) for nq (1 or 2) vectors
           (scratch: tacs_hip_bdotq_scratch_bytes);
   bdcgs_update: per right-hand side j with cf = coef + j*sC =
           [s (k) | t (k) | 1/omega | tau], k = k1 + k2:
           u <- (u - W s)/omega,  z <- (z - W t - u tau)/omega (z may be NULL). */
size_t tacs_hip_bdotq_scratch_bytes(int R);
int tacs_hip_bdotq(size_t n, int R, int nq, int k1, const double* W1, int k2, const double* W2,
                   size_t ldw, const double* q0, const double* q1, double* scratch, double* out0,
                   double* out1, size_t sR, hipStream_t stream);
int tacs_hip_bdcgs_update(size_t n, int R, int k1, const double* W1, int k2, const double* W2,
                          size_t ldw, const double* coef, size_t sC, double* u, double* z,
                          hipStream_t stream);
/* The per-step small computation of the recycling solver's DCGS2 Arnoldi
   step on the device (one wavefront per right-hand side; no counterpart in
   the reference, whose GMRES::solve src/bpmat/KSM.cpp:785 updates its
   Hessenberg and Givens rotations on the host).  From the dot lists dA, dB
   (slices of stride S) of step i it finishes least-squares column kd_j + i - 1
   (Givens update, stopping test against tgt[j]), forms the first-pass
   coefficients of column kd_j + i and the update coefficients cf = [s | t |
   1/omega | tau] (stride sC).  ls: per right-hand side least-squares state
   (stride sLS >= tacs_hip_bdcgs_state_stride(ld): Hb, Hr (ld x (ld - 1)), g,
   c, rotation cosines, sines, pending column (ld each)); ist: 4 ints per
   right-hand side (converged, failed, finished columns, columns); flags: a
   host-visible copy of ist after the step (may be NULL).  ld <= 512. */
size_t tacs_hip_bdcgs_state_stride(int ld);
int tacs_hip_bdcgs_coeffs(int R, int i, int Kd, int ld, const int* kd, const double* tgt,
                          const double* dA, const double* dB, size_t S, double* ls, size_t sLS,
                          int* ist, double* cf, size_t sC, int* flags, hipStream_t stream);
int tacs_hip_badd(int R, int k, const double* x, size_t sx, double* y, size_t sy,
                  hipStream_t stream);
int tacs_hip_bscale(size_t n, int R, const double* s, double* x, hipStream_t stream);
int tacs_hip_bsqrt_inv(int R, const double* sq, size_t sS, double* nrm, size_t sN, double* inv,
                       hipStream_t stream);
int tacs_hip_bgivens(int R, int i, int m, double* H, size_t sH, double* Qc, double* Qs, size_t sQ,
                     double* res, size_t sRes, double* st, hipStream_t stream);
int tacs_hip_bbacksolve(int R, int m, const int* cnt, const double* H, size_t sH, double* res,
                        size_t sRes, hipStream_t stream);
int tacs_hip_bdiag_apply(int b, int prec, int nrows, int R, const void* Dinv, const double* x,
                         double* y, hipStream_t stream);

/* One Chebyshev step for A z = x with the block-Jacobi inverse D^{-1}
   (prec 64 / 32 blocks, b = 3 or 6) over interleaved block vectors:
   w = D^{-1} (x - Az) (first: D^{-1} x), d = c1 d + c2 w (first: c2 w),
   z += d (first: z = d), and z32 = (float) z when z32 != NULL.  Saad,
   Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1. */
int tacs_hip_bcheb_step(int b, int prec, int nrows, int R, const void* Dinv, const double* x,
                        const double* Az, double* d, double* z, double c1, double c2, int first,
                        float* z32, hipStream_t stream);
/* The same step fused into the 3x3 product over R = 16, 20, ... (R % 4 == 0)
   interleaved right-hand sides with FP32 values A: a = A X (X the FP32 copy
   of z), then w = D^{-1} (x - a) (Dinv FP32), d = c1 d + c2 w, z += d (FP64)
   and Xout = (float) z (Xout != X). */
int tacs_hip_bsr3_spmm_cheb(int nrows, int R, const int* rowp, const int* cols, const float* A,
                            const float* X, const double* x, const float* Dinv, double* d,
                            double* z, float* Xout, double c1, double c2, hipStream_t stream);
/* The same fused step in the forms a polynomial runs after its first steps
   (both read z and d and write d, no counterpart in the reference):
   mode 1, three-term: d holds the iterate before z and receives
   (1 + c1) z - c1 d + c2 w; mode 2, into that form: d holds the last d and
   receives z + c1 d + c2 w.  Mode 0 is tacs_hip_bsr3_spmm_cheb.  d != z. */
int tacs_hip_bsr3_spmm_cheb_mode(int mode, int nrows, int R, const int* rowp, const int* cols,
                                 const float* A, const float* X, const double* x,
                                 const float* Dinv, double* d, double* z, float* Xout, double c1,
                                 double c2, hipStream_t stream);

/* Segmented systems (the R <= 64 diagonal blocks of a block-diagonal 3x3
   BCSR matrix, block j = node range [off[j], off[j+1]), solved as R
   independent systems): a segmented block vector keeps local node i of
   system j at X[(3 i + c) R + j], i < maxn (the largest block); entries
   beyond a system's size are zero.  No counterpart in the reference (one
   assembler and one GMRES per mesh there).
   seg_pack:    X = the segmented form of the union vector x;
   seg_unpack:  x (nnodes = off[R] nodes) from X;
   seg_bdiag_apply: y = D^{-1} x per local node, D^{-1} the union's inverse
                diagonal blocks (prec 64 / 32);
   seg_spmm3:   Y = A X with the union's rows and values (prec 64 / 32),
                every column of a row of block j inside block j. */
int tacs_hip_seg_pack(int maxn, int R, const int* off, const double* x, double* X,
                      hipStream_t stream);
int tacs_hip_seg_unpack(int nnodes, int R, const int* off, const double* X, double* x,
                        hipStream_t stream);
int tacs_hip_seg_bdiag_apply(int prec, int maxn, int R, const int* off, const void* Dinv,
                             const double* x, double* y, hipStream_t stream);
/* seg_to_union / seg_from_union: the same conversions as unpack / pack,
   through LDS tiles of 32 local nodes (coalesced on both sides). */
int tacs_hip_seg_to_union(int maxn, int R, const int* off, const double* X, double* x,
                          hipStream_t stream);
int tacs_hip_seg_from_union(int maxn, int R, const int* off, const double* x, double* X,
                            hipStream_t stream);
int tacs_hip_seg_spmm3(int prec, int maxn, int R, const int* off, const int* rowp,
                       const int* cols, const void* A, const double* X, double* Y,
                       hipStream_t stream);
int tacs_hip_bsr3_spmm_mixed(int nrows, int nrhs, const int* rowp, const int* cols, const float* A,
                             const double* X, double* Y, hipStream_t stream);
/* Launch variants of tacs_hip_bsr3_spmm / _mixed (prec_vals 64 or 32), all
   with the same sums.  nrhs even and at most 12 (two right-hand sides per
   lane): blocks per lane per trip 1 (variant 0) or 2 (1).  nrhs a multiple
   of 4 from 16 (four per lane): blocks per lane per trip and non-temporal
   value loads, variants 0 (1, on), 1 (2, on), 2 (1, off), 3 (2, off),
   4 (4, on).  Other nrhs: one variant.  The two entries above launch the
   variant of the tuning table. */
int tacs_hip_bsr3_spmm_nvariants(int nrhs);
int tacs_hip_bsr3_spmm_variant(int variant, int prec_vals, int nrows, int nrhs, const int* rowp,
                               const int* cols, const void* A, const double* X, double* Y,
                               hipStream_t stream);
int tacs_hip_bbasis_combine(size_t n, int R, int nin, const double* W, size_t ldw, const double* P,
                            size_t sP, int ldp, int nout, double* out, size_t ldo, int add,
                            hipStream_t stream);
int tacs_hip_bbroadcast(size_t n, int R, const double* x, double* X, hipStream_t stream);
/* Per right-hand side tall-skinny products of block vectors:
   out[j*sR + b*ldo + v] = X_v(:, j) . Y_b(:, j), v < p, b < q (X_v, Y_b at
   stride ldw); scratch: tacs_hip_btsgemm_scratch_bytes. */
size_t tacs_hip_btsgemm_scratch_bytes(int R);
int tacs_hip_btsgemm(size_t n, int R, int p, const double* X, int q, const double* Y, size_t ldw,
                     double* scratch, double* out, size_t sR, int ldo, hipStream_t stream);
int tacs_hip_bpack(size_t n, int R, int j, const double* x, double* X, hipStream_t stream);
int tacs_hip_bunpack(size_t n, int R, int j, const double* X, double* x, hipStream_t stream);

/* Host-side dense helper of the deflated-restart GMRES: eigenvalues of the
   row-major n x n matrix G sorted by ascending modulus (wr, wi) and a real
   basis V (column-major n x kout) of the eigenvectors of the k smallest; a
   complex pair contributes its real and imaginary parts (kout <= k + 1).
   Returns kout, or -1 when the QR iteration fails.  No counterpart in the
   reference (its GMRES restarts without deflation). */
int tacs_dense_smallest_eigvecs(int n, const double* G, int k, double* V, double* wr,
                                double* wi);

/* ------------------------------------------------------------------------
   2. Object entry points (mirror tacs/cpp_headers/TACS.pxd and the Cython
      wrappers; all handles are TACSObject*)
   ------------------------------------------------------------------------ */
typedef void* tacs_obj;

const char* tacs_last_error(void);
void tacs_incref(tacs_obj o);
void tacs_decref(tacs_obj o);
int tacs_refcount(tacs_obj o);
/* Stream used by every object operation (default: the null stream). */
void tacs_set_stream(hipStream_t stream);
hipStream_t tacs_get_stream(void);
int tacs_device_synchronize(void);

/* constitutive.pyx: MaterialProperties :89, SolidConstitutive :875 */
tacs_obj tacs_material_create(double rho, double specific_heat, double E, double nu,
                              double ys, double alpha, double kappa);
tacs_obj tacs_solid_constitutive_create(tacs_obj props, double t, int tNum, double tlb,
                                        double tub);
/* elements.pyx: LinearHexaBasis :297, QuadraticHexaBasis :308,
   LinearElasticity3D :498, Element3D :564; TACSTraction3D
   (src/elements/TACSTraction3D.h:22) */
tacs_obj tacs_linear_hexa_basis_create(void);
tacs_obj tacs_quadratic_hexa_basis_create(void);
tacs_obj tacs_linear_elasticity3d_create(tacs_obj stiff, int strain_type);
tacs_obj tacs_element3d_create(tacs_obj model, tacs_obj basis);
tacs_obj tacs_traction3d_create(int vars_per_node, int face, tacs_obj basis,
                                const double* trac, int coord_component);
/* TACSPressure3D (src/elements/TACSPressure3D.h; TACSElement3D::createElementPressure
   src/elements/TACSElement3D.cpp:54): constant pressure p on face (0..5) of a
   solid element, traction -p n with n the outward unit face normal. */
tacs_obj tacs_pressure3d_create(int vars_per_node, int face, tacs_obj basis, double p);
/* TACSIsoShellConstitutive(props, t, tNum, tlb, tub, tOffset, kcorr)
   (src/constitutive/TACSIsoShellConstitutive.h:33; tacs/constitutive.pyx:1109) */
tacs_obj tacs_iso_shell_constitutive_create(tacs_obj props, double t, int tNum, double tlb,
                                            double tub, double tOffset, double kcorr);
/* TACSShellNaturalTransform (src/elements/shell/TACSShellElementTransform.h:21) */
tacs_obj tacs_shell_natural_transform_create(void);
/* TACSShellRefAxisTransform(axis) (TACSShellElementTransform.h:95); the axis
   is normalised as there. */
tacs_obj tacs_shell_ref_axis_transform_create(const double* axis);
/* TACSQuad4Shell(transform, con) (src/elements/shell/TACSShellElementDefs.h:18;
   tacs/elements.pyx:694) */
tacs_obj tacs_quad4_shell_create(tacs_obj transform, tacs_obj con);
/* TACSShellElement::createElementPressure(faceIndex, p)
   (src/elements/shell/TACSShellElement.h:109): a uniform-pressure aux element */
tacs_obj tacs_element_create_pressure(tacs_obj element, int face, double p);
/* TACSAuxElements (src/TACSAuxElements.h:52) */
tacs_obj tacs_aux_elements_create(int size);
int tacs_aux_elements_add(tacs_obj aux, int elem_index, tacs_obj element);

/* TACSAssembler (TACS.pxd:321-428; TACS.pyx Assembler :1595) */
tacs_obj tacs_assembler_create(int vars_per_node, int num_owned_nodes, int num_elements,
                               int num_dependent_nodes);
int tacs_assembler_set_element_connectivity(tacs_obj a, const int* ptr, const int* conn);
int tacs_assembler_set_elements(tacs_obj a, tacs_obj* elements);
int tacs_assembler_add_bcs(tacs_obj a, int nnodes, const int* nodes, int nbcs, const int* vars,
                           const double* vals);
int tacs_assembler_initialize(tacs_obj a);
/* TACSAssembler::setElementPrecision (64 or 32) and setJacobianKernel (0
   measured choice, 1 node-centric rows, 2 element blocks + gather); no
   counterpart in the reference.  tacs_assembler_jacobian_kernel returns the
   solid Jacobian kernel assembleJacobian launches (1 or 2; 0 for shells). */
int tacs_assembler_set_element_precision(tacs_obj a, int prec);
int tacs_assembler_set_jacobian_kernel(tacs_obj a, int kernel);
int tacs_assembler_jacobian_kernel(tacs_obj a);
int tacs_assembler_set_aux_elements(tacs_obj a, tacs_obj aux);
int tacs_assembler_get_info(tacs_obj a, int* vars_per_node, int* num_nodes,
                            int* num_elements, int* num_design_vars);
tacs_obj tacs_assembler_create_vec(tacs_obj a);
tacs_obj tacs_assembler_create_node_vec(tacs_obj a);
tacs_obj tacs_assembler_create_design_vec(tacs_obj a);
tacs_obj tacs_assembler_create_mat(tacs_obj a);
int tacs_assembler_set_nodes(tacs_obj a, tacs_obj X);
int tacs_assembler_get_nodes(tacs_obj a, tacs_obj X);
int tacs_assembler_set_variables(tacs_obj a, tacs_obj q);
int tacs_assembler_get_variables(tacs_obj a, tacs_obj q);
int tacs_assembler_zero_variables(tacs_obj a);
int tacs_assembler_set_design_vars(tacs_obj a, tacs_obj dvs);
int tacs_assembler_get_design_vars(tacs_obj a, tacs_obj dvs);
int tacs_assembler_apply_bcs_vec(tacs_obj a, tacs_obj vec);
int tacs_assembler_apply_bcs_mat(tacs_obj a, tacs_obj mat);
/* TACSAssembler::assembleRes src/TACSAssembler.cpp:4133 */
int tacs_assembler_assemble_res(tacs_obj a, tacs_obj res, double lambda, int apply_bcs);
/* TACSAssembler::assembleJacobian src/TACSAssembler.cpp:4291 */
int tacs_assembler_assemble_jacobian(tacs_obj a, double alpha, double beta, double gamma,
                                     tacs_obj res, tacs_obj mat, int mat_or, double lambda,
                                     int apply_bcs);
/* evalFunctions :4617, addDVSens :4767, addSVSens :4941,
   addAdjointResProducts :5033 */
int tacs_assembler_eval_functions(tacs_obj a, int n, tacs_obj* funcs, double* vals);
int tacs_assembler_add_dv_sens(tacs_obj a, double coef, int n, tacs_obj* funcs,
                               tacs_obj* dfdx);
int tacs_assembler_add_sv_sens(tacs_obj a, double alpha, double beta, double gamma, int n,
                               tacs_obj* funcs, tacs_obj* dfdu);
int tacs_assembler_add_adjoint_res_products(tacs_obj a, double scale, int n, tacs_obj* adjoint,
                                            tacs_obj* dfdx, double lambda);
/* TACSAssembler::addXptSens (src/TACSAssembler.cpp:4858; TACS.pyx :2382) and
   addAdjointResXptSensProducts (:5136; TACS.pyx :2454): node-coordinate
   sensitivities (solid family). */
int tacs_assembler_add_xpt_sens(tacs_obj assembler, double coef, int n, tacs_obj* funcs,
                                tacs_obj* dfdX);
int tacs_assembler_add_adjoint_res_xpt_sens_products(tacs_obj assembler, double scale, int n,
                                                     tacs_obj* adjoint, tacs_obj* dfdX,
                                                     double lambda);

/* TACSBVec (TACS.pyx Vec :594; src/bpmat/TACSBVec.h:67) */
int tacs_vec_get_size(tacs_obj v, int* size, int* bsize);
double* tacs_vec_get_array(tacs_obj v);     /* host mirror, host becomes current */
double* tacs_vec_device_array(tacs_obj v);  /* device array, device becomes current */
/* Device-to-device copies between a vector and caller memory of n doubles
   (on the library stream). */
int tacs_vec_copy_from_device(tacs_obj v, const double* src, size_t n);
int tacs_vec_copy_to_device(tacs_obj v, double* dst, size_t n);
double tacs_vec_norm(tacs_obj v);
double tacs_vec_dot(tacs_obj v, tacs_obj x);
int tacs_vec_axpy(tacs_obj v, double alpha, tacs_obj x);
int tacs_vec_axpby(tacs_obj v, double alpha, double beta, tacs_obj x);
int tacs_vec_scale(tacs_obj v, double alpha);
int tacs_vec_copy_values(tacs_obj v, tacs_obj x);
int tacs_vec_zero_entries(tacs_obj v);
int tacs_vec_set(tacs_obj v, double val);
int tacs_vec_set_rand(tacs_obj v, double lower, double upper);

/* TACSParallelMat (TACS.pyx Mat :960) */
int tacs_mat_mult(tacs_obj m, tacs_obj x, tacs_obj y);

/* TACSAssembler::assembleMatType src/TACSAssembler.cpp:4418 (TACS.pyx
   Assembler.assembleMatType :2228) and assembleMatCombo :4519 (TACS.pyx
   :2256, there with two types): matType 1 stiffness, 2 mass, 3 geometric
   stiffness at the current variables (solids); shells: stiffness only. */
int tacs_assembler_assemble_mat_type(tacs_obj a, int matType, tacs_obj mat, int matOr,
                                     double loadScale, int applyBCs);
int tacs_assembler_assemble_mat_combo(tacs_obj a, int nmats, const int* matTypes,
                                      const double* scale, tacs_obj mat, int matOr,
                                      double loadScale, int applyBCs);

/* Matrix-free Jacobian products.
   tacs_matfree_create: TACSMatrixFreeMat(assembler) src/TACSMatrixFreeMat.cpp:17;
   tacs_matfree_assemble_data: TACSMatrixFreeMat::assembleMatrixFreeData
     src/TACSMatrixFreeMat.cpp:37 (matType: ElementMatrixType of
     src/elements/TACSElementTypes.h:113; JACOBIAN, STIFFNESS and MASS);
   tacs_mat_mult accepts the matrix-free matrix too (TACSMatrixFreeMat::mult
     :55, Dirichlet rows of y zeroed); tacs_matfree_mult_transpose :66;
   tacs_matfree_mult_device: the same product on device arrays;
   tacs_assembler_get_matrix_free_data_size / _assemble_matrix_free_data /
   _add_matrix_free_vec_product: TACSAssembler::getMatrixFreeDataSize,
     assembleMatrixFreeData, addMatrixFreeVecProduct src/TACSAssembler.cpp:
     5502, 5563, 5645 with device arrays data and temp of the returned sizes
     (doubles).  A TACSMatrixFreeMat is also a GMRES operator
     (tacs_gmres_create). */
tacs_obj tacs_matfree_create(tacs_obj assembler);
int tacs_matfree_assemble_data(tacs_obj m, int matType, double alpha, double beta,
                               double gamma);
int tacs_matfree_mult_transpose(tacs_obj m, tacs_obj x, tacs_obj y);
int tacs_matfree_mult_device(tacs_obj m, const double* x, double* y);
int tacs_assembler_get_matrix_free_data_size(tacs_obj a, int matType, int* data_size,
                                             int* temp_size);
int tacs_assembler_assemble_matrix_free_data(tacs_obj a, int matType, double alpha, double beta,
                                             double gamma, double* data);
int tacs_assembler_add_matrix_free_vec_product(tacs_obj a, int matType, const double* data,
                                               double* temp, tacs_obj x, tacs_obj y,
                                               int matOr, double lambda, int applyBCs);
/* y = A x on device vectors in the precisions the solvers use (values
   prec_vals, vectors prec_vecs: 64/64, 32/64 or 32/32), as
   TACSParallelMat::mult src/bpmat/TACSParallelMat.cpp:248 with FP64
   throughout; FP32 values are the FP64 values rounded to nearest. */
int tacs_mat_mult_prec(tacs_obj m, int prec_vals, int prec_vecs, const void* x, void* y);
int tacs_mat_zero_entries(tacs_obj m);
/* Brings the product layouts of an assembled matrix up to date with its
   values now (the sliced ELL copy in FP64, prec 64, or FP32, prec 32), as
   the first product after a change of the values otherwise does; no
   counterpart in the reference. */
int tacs_mat_prepare_products(tacs_obj m, int prec);

/* Measured launch choices (no counterpart in the reference; see the tuning
   table above).  tacs_mat_autotune times every sliced ELL product variant
   of the three kinds on this matrix (reps launches each, after one untimed)
   and keeps the fastest per kind for the matrix; ms[kind * maxv + v]
   receives the average times, *nvariants the variants per kind.  The
   get/set pair reads or forces a kind's variant (-1: the table's).
   tacs_mat_autotune_spmm times the variants of the nrhs right-hand-side
   product (prec 64 or 32 values) and enters the fastest in the tuning table
   for this matrix's size class.  tacs_assembler_autotune times the solid
   Jacobian's rows-kernel variants and, with Jacobian kernel setting 0, the
   element-block form (variant nv), assembling alpha = 1 Jacobians into mat
   (reps each), and keeps the fastest for the assembler; the get/set pair
   reads or forces the choice (-1: the table's).  Every variant of a kernel
   gives identical values; the element-block form agrees with the rows to
   rounding. */
int tacs_mat_autotune(tacs_obj mat, int reps, double* ms, int maxv, int* nvariants);
/* One Chebyshev step of the recycling solvers' preconditioner on device
   vectors (tacs_hip_sell_cheb_step over the matrix's sliced ELL copy in
   precision prec_vals 64, 32 or 16 (bfloat16), with pc's inverse diagonal
   blocks): d = c1 d + c2 D^{-1} (rhs - A z), zout = z + d.  No counterpart
   in the reference. */
int tacs_mat_cheb_step(tacs_obj mat, tacs_obj pc, int prec_vals, const double* z, double* zout,
                       const double* rhs, double* d, double c1, double c2);
/* tacs_hip_sell_cheb_step3 over the matrix's sliced ELL copy with pc's
   inverse diagonal blocks: zprev <- (1 + c1) z - c1 zprev
   + c2 D^{-1} (rhs - A z).  No counterpart in the reference. */
int tacs_mat_cheb_step3(tacs_obj mat, tacs_obj pc, int prec_vals, const double* z, double* zprev,
                        const double* rhs, double c1, double c2);
int tacs_mat_get_sell_variant(tacs_obj mat, int kind, int* variant);
int tacs_mat_set_sell_variant(tacs_obj mat, int kind, int variant);
int tacs_mat_autotune_spmm(tacs_obj mat, int nrhs, int prec, int reps, double* ms, int maxv,
                           int* nvariants);
int tacs_assembler_autotune(tacs_obj assembler, tacs_obj mat, int reps, double* ms, int maxv,
                            int* nchoices);
int tacs_assembler_get_jacobian_variant(tacs_obj assembler, int* variant);
int tacs_assembler_set_jacobian_variant(tacs_obj assembler, int variant);
int tacs_mat_get_info(tacs_obj m, int* bsize, int* nrows, long* nnzb);
/* Host copies of the block-CSR arrays (BCSRMat::getArrays BCSRMat.h:82). */
int tacs_mat_get_arrays(tacs_obj m, const int** rowp, const int** cols, const double** A);
/* Device arrays of the matrix (valid until the matrix is destroyed).  The
   values are read-only here: the FP32 and sliced ELL copies the products use
   are refreshed when assembly, zeroEntries or applyBCs change them. */
int tacs_mat_device_arrays(tacs_obj m, const int** rowp, const int** cols, const double** A);

/* Preconditioner / Krylov (TACS.pyx Pc :1106, KSM :1256) */
tacs_obj tacs_block_jacobi_pc_create(tacs_obj mat);
int tacs_pc_factor(tacs_obj pc);
int tacs_pc_apply(tacs_obj pc, tacs_obj x, tacs_obj y);
/* Chebyshev polynomial of the block-Jacobi iteration as the preconditioner
   of the recycling and batched solvers (degree 1..32, 1 = plain
   block-Jacobi; interval [lmax / ratio, lmax] of D^{-1} A, lmax estimated by
   the next factor()); values 32, or 16 for products inside the polynomial
   that read a bfloat16 copy of the values (single systems with FP32-valued
   operators) or an FP32 copy of the iterate (batched solves of 16 or more
   right-hand sides, 3x3 blocks).  GMRES and pc_apply keep plain
   block-Jacobi.  No counterpart in the reference. */
int tacs_pc_set_chebyshev(tacs_obj pc, int degree, double ratio, int values);
/* the settings and the lmax the last factor() estimated (1.05 x the
   largest Ritz value of 30 Lanczos steps; 0 before a factor() with
   degree > 1) */
int tacs_pc_get_chebyshev(tacs_obj pc, int* degree, double* ratio, double* lmax);
tacs_obj tacs_gmres_create(tacs_obj mat, tacs_obj pc, int m, int nrestart, int is_flexible);
/* Residual monitors (KSMPrintStdout src/bpmat/KSM.h:293, KSMPrintFile :310;
   GMRES::setMonitor KSM.cpp:739, TACS.pyx KSM.setMonitor :1326): GMRES and
   the mixed GMRES print residual norms per step (0 at each restart); the
   recycling and batched solvers print the true residual norm of each
   system at the start of every cycle.  set_monitor takes any of the three
   solvers; monitor NULL removes it. */
tacs_obj tacs_ksm_print_stdout_create(const char* descript, int rank, int freq);
tacs_obj tacs_ksm_print_file_create(const char* filename, const char* descript, int rank,
                                    int freq);
int tacs_ksm_set_monitor(tacs_obj ksm, tacs_obj monitor);
int tacs_gmres_set_ortho_type(tacs_obj ksm, int ortho);
/* GMRES::setCycleGraph: each restart cycle's Arnoldi steps as one captured
   HIP graph (assembled operator; same iterates as the step-by-step loop).  No
   counterpart in the reference. */
int tacs_gmres_set_cycle_graph(tacs_obj ksm, int on);
/* Mixed-precision restarted GMRES(m) (FP32 Arnoldi cycles, FP64 restart
   residual and solution); no counterpart in the reference. */
tacs_obj tacs_mixed_gmres_create(tacs_obj mat, tacs_obj pc, int m, int nrestart);
/* Krylov subspace recycling for the mixed GMRES (GCRO-DR: each cycle keeps
   the k harmonic Ritz vectors of smallest harmonic Ritz value, and the space
   carries over to later solves with the same matrix values; delayed
   classical Gram-Schmidt; 0 <= k <= 31, k < m - 1, m + k <= 250).  No
   counterpart in the reference. */
int tacs_mixed_gmres_set_deflation(tacs_obj ksm, int k);
/* Batched restarted GMRES(m) for nrhs <= 64 right-hand sides of one 3x3
   block matrix: one multi-RHS SpMM per step for all systems, per-system
   Hessenberg, rotations, stopping test and iteration count; prec 64 (FP64
   Arnoldi operator) or 32 (FP32-rounded matrix and block-Jacobi values, FP64
   vectors), every cycle restarted from the FP64 residual.  No counterpart in
   the reference, which solves load cases one GMRES::solve at a time
   (src/bpmat/KSM.cpp:785).  solve_block takes device block vectors with the
   right-hand sides interleaved, B[i*nrhs + j]; get_stats fills per-system
   iteration counts, flags (1: converged), initial and final residual norms
   and returns the number of cycles. */
tacs_obj tacs_batched_gmres_create(tacs_obj mat, tacs_obj pc, int m, int nrestart, int nrhs,
                                   int prec);
int tacs_batched_gmres_set_tolerances(tacs_obj ksm, double rtol, double atol);
/* Krylov subspace recycling per right-hand side for the batched GMRES
   (GCRO-DR: each system keeps a k-dimensional space updated every cycle
   from its harmonic Ritz vectors; 0 <= k <= 31, k < m - 1, m + k <= 250;
   0 turns it off).  No counterpart in the reference. */
int tacs_batched_gmres_set_deflation(tacs_obj ksm, int k);
/* Gram-Schmidt of the recycling Arnoldi steps: 1 (default) delayed
   classical Gram-Schmidt (DCGS2), 0 classical Gram-Schmidt twice. */
int tacs_batched_gmres_set_delayed_reorth(tacs_obj ksm, int delayed);
/* Deflate the batched solves by the recycled space of a mixed GMRES with
   deflation over the same matrix (copied; dropped when the matrix values
   change); with per right-hand side recycling on, it seeds every system's
   space.  Returns its dimension, 0 when src has none, -1 on error. */
int tacs_batched_gmres_set_recycle_space(tacs_obj ksm, tacs_obj src);
int tacs_batched_gmres_solve(tacs_obj ksm, int nrhs, tacs_obj* b, tacs_obj* x, int zero_guess);
int tacs_batched_gmres_solve_block(tacs_obj ksm, int nrhs, const double* B, double* X,
                                   int zero_guess);
int tacs_batched_gmres_get_stats(tacs_obj ksm, int nrhs, int* iters, int* flags, double* r0,
                                 double* rn);
/* Segmented systems (no counterpart in the reference, which builds one
   assembler and one GMRES per mesh): mat block diagonal with nseg <= 64
   diagonal blocks, block j the node range [node_off[j], node_off[j+1])
   (node_off[0] = 0, node_off[nseg] = rows; a row coupling two blocks is
   refused).  Each block is its own system with its own Krylov space,
   recycled space (set_deflation), Hessenberg matrix and stopping test
   |r_j| <= rtol |r_j(0)| (or atol); all advance together and a converged
   block stops counting steps.  solve_segments takes whole-matrix vectors;
   get_stats reports per block.  Returns 1 when every block converged. */
tacs_obj tacs_batched_gmres_create_segments(tacs_obj mat, tacs_obj pc, int m, int nrestart,
                                            int nseg, const int* node_off, int prec);
int tacs_batched_gmres_solve_segments(tacs_obj ksm, tacs_obj b, tacs_obj x, int zero_guess);
int tacs_ksm_set_tolerances(tacs_obj ksm, double rtol, double atol);
int tacs_ksm_solve(tacs_obj ksm, tacs_obj b, tacs_obj x, int zero_guess);
int tacs_ksm_get_iter_count(tacs_obj ksm);
double tacs_ksm_get_res_norm(tacs_obj ksm);
int tacs_mixed_gmres_get_outer_count(tacs_obj ksm);

/* Functions (functions.pyx Compliance :172, KSFailure :358) */
tacs_obj tacs_compliance_create(tacs_obj assembler);
tacs_obj tacs_ksfailure_create(tacs_obj assembler, double ks_weight, double alpha,
                               double safety_factor);
double tacs_function_get_value(tacs_obj f);
double tacs_ksfailure_get_max_failure(tacs_obj f);

#ifdef __cplusplus
}
#endif

#endif
