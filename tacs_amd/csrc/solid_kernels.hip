// gfx950 kernels for TACSElement3D + TACSLinearElasticity3D (linear strain)
// + TACSSolidConstitutive on 8-node (order 2) and 27-node (order 3) hexahedra.
//
// Jacobian assembly is node-centric: one wavefront owns one block row of the
// BCSR matrix, recomputes the 3 x (3*NPE) row strips of the incident elements
// at every quadrature point, accumulates them per column block in LDS in
// ascending element order (the order in which the reference's element loop
// adds them, src/TACSAssembler.cpp:4352-4383), applies the Dirichlet rows and
// writes the row once with coalesced stores.  No atomics, no colouring and no
// zero-fill of the matrix are needed.
//
// Vector-valued element quantities (residual, function state sensitivities,
// traction loads) are computed per element into an element-vector buffer and
// gathered per node in ascending element order (the reference's setValues ADD
// order), again without atomics.
#include "tacs_hip_common.h"

#define WAVE_SYNC()                                    \
  do {                                                 \
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); \
    __builtin_amdgcn_wave_barrier();                   \
  } while (0)

// Gauss points/weights: the reference's literals
// (src/elements/basis/TACSGaussQuadrature.h:26-30).
template <int ORD>
__device__ __forceinline__ double gauss_pt(int i) {
  if (ORD == 2) return i == 0 ? -0.577350269189626 : 0.577350269189626;
  return i == 0 ? -0.774596669241483 : (i == 1 ? 0.0 : 0.774596669241483);
}
template <int ORD>
__device__ __forceinline__ double gauss_wt(int i) {
  if (ORD == 2) return 1.0;
  return i == 1 ? 8.0 / 9.0 : 5.0 / 9.0;
}

// Lagrange functions on the Gauss-Lobatto knots (TACSHexaBasis.cpp).
template <int ORD>
__device__ __forceinline__ void lag1d(double x, double* n, double* d) {
  if (ORD == 2) {
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

// Element of a short register array selected by a runtime index without
// dynamic indexing (which would place the array in private memory).
template <int ORD>
__device__ __forceinline__ double pick(const double* v, int i) {
  // both values are loaded and one is selected: written as i ? v[1] : v[0]
  // the compiler indexed the array with i and kept it in scratch memory
  if (ORD == 2) {
    const double v0 = v[0], v1 = v[1];
    return (i == 1) ? v1 : v0;
  }
  return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

// 1D values at the three coordinates of quadrature point q.
template <int ORD>
struct QP1D {
  double n[3][ORD], d[3][ORD];
  double w;
  __device__ __forceinline__ void eval(int q) {
    int iq = q % ORD, jq = (q / ORD) % ORD, kq = q / (ORD * ORD);
    lag1d<ORD>(gauss_pt<ORD>(iq), n[0], d[0]);
    lag1d<ORD>(gauss_pt<ORD>(jq), n[1], d[1]);
    lag1d<ORD>(gauss_pt<ORD>(kq), n[2], d[2]);
    w = gauss_wt<ORD>(iq) * gauss_wt<ORD>(jq) * gauss_wt<ORD>(kq);
  }
  __device__ __forceinline__ void basis(int a, double& N, double g[3]) const {
    int ia = a % ORD, ja = (a / ORD) % ORD, ka = a / (ORD * ORD);
    double n0 = pick<ORD>(n[0], ia), n1 = pick<ORD>(n[1], ja), n2 = pick<ORD>(n[2], ka);
    double d0 = pick<ORD>(d[0], ia), d1 = pick<ORD>(d[1], ja), d2 = pick<ORD>(d[2], ka);
    N = n0 * n1 * n2;
    g[0] = d0 * n1 * n2;
    g[1] = n0 * d1 * n2;
    g[2] = n0 * n1 * d2;
  }
};

// inv3x3 of src/elements/TACSElementAlgebra.h:1980; returns det.
__device__ __forceinline__ double inv3(const double A[9], double B[9]) {
  double det = A[8] * (A[0] * A[4] - A[3] * A[1]) -
               A[7] * (A[0] * A[5] - A[3] * A[2]) +
               A[6] * (A[1] * A[5] - A[2] * A[4]);
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

__device__ __forceinline__ int voigt(int i, int j) {
  return i == j ? i : (i + j == 3 ? 3 : (i + j == 2 ? 4 : 5));
}
__device__ __forceinline__ double c21(const SolidMaterial& m, int i, int j) {
  if (i > j) {
    int t = i;
    i = j;
    j = t;
  }
  // packed row offsets of the upper triangle: 0,6,11,15,18,20
  const int off[6] = {0, 6, 11, 15, 18, 20};
  return m.C[off[i] + (j - i)];
}

// Element e's material: mat, with (lam, mu, rho, ys) from the per-element
// table when the elements carry different isotropic materials
// (TACSAssembler::setElements allows one constitutive object per element,
// src/TACSAssembler.cpp:623).
__device__ __forceinline__ SolidMaterial elem_mat(const SolidMaterial& mat, int e) {
  SolidMaterial m = mat;
  if (mat.elem_props) {
    const double* q = mat.elem_props + 4 * (size_t)e;
    m.lam = q[0];
    m.mu = q[1];
    m.rho = q[2];
    m.ys = q[3];
  }
  return m;
}

// s = C e (no thickness scaling), TACSSolidConstitutive::evalStress :142.
__device__ __forceinline__ void stress6(const SolidMaterial& m, const double e[6],
                                        double s[6]) {
  if (m.isotropic) {
    double tr = e[0] + e[1] + e[2];
    s[0] = m.lam * tr + 2.0 * m.mu * e[0];
    s[1] = m.lam * tr + 2.0 * m.mu * e[1];
    s[2] = m.lam * tr + 2.0 * m.mu * e[2];
    s[3] = m.mu * e[3];
    s[4] = m.mu * e[4];
    s[5] = m.mu * e[5];
  } else {
#pragma unroll
    for (int i = 0; i < 6; i++) {
      double v = 0.0;
#pragma unroll
      for (int j = 0; j < 6; j++) v += c21(m, i, j) * e[j];
      s[i] = v;
    }
  }
}

// Element geometry at one quadrature point: J = Xd^{-1}, returns detXd.
template <int ORD>
__device__ __forceinline__ double elem_geometry(const QP1D<ORD>& qp, const int* __restrict__ c,
                                                const double* __restrict__ X, double J[9]) {
  constexpr int NPE = ORD * ORD * ORD;
  double Xd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int b = 0; b < NPE; b++) {
    double N, g[3];
    qp.basis(b, N, g);
    int nb = c[b];
    double x0 = X[3 * nb], x1 = X[3 * nb + 1], x2 = X[3 * nb + 2];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      Xd[k] += x0 * g[k];
      Xd[3 + k] += x1 * g[k];
      Xd[6 + k] += x2 * g[k];
    }
  }
  return inv3(Xd, J);
}

__device__ __forceinline__ void grad_x(const double J[9], const double g[3], double gx[3]) {
  gx[0] = g[0] * J[0] + g[1] * J[3] + g[2] * J[6];
  gx[1] = g[0] * J[1] + g[1] * J[4] + g[2] * J[7];
  gx[2] = g[0] * J[2] + g[1] * J[5] + g[2] * J[8];
}

// ---------------------------------------------------------------------------
// Geometry of every (element, quadrature point): G[(e NQP + q) 10 + k] =
// J^-1 (k < 9, J = dX/dxi) and w det(J) t_e (k = 9), computed once per
// assembly; the row kernel copies the entries of a row's elements instead of
// recomputing them for each of the element's nodes.
// ---------------------------------------------------------------------------
template <int ORD>
__global__ __launch_bounds__(256) void solid_qp_geometry(int nelem, const int* __restrict__ conn,
                                                         const double* __restrict__ X,
                                                         const double* __restrict__ telem,
                                                         double* __restrict__ G) {
  constexpr int NPE = ORD * ORD * ORD, NQP = NPE;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nelem * NQP) return;
  const int elem = (int)(t / NQP), q = (int)(t - (size_t)elem * NQP);
  QP1D<ORD> qp;
  qp.eval(q);
  double J[9];
  const double det = elem_geometry<ORD>(qp, &conn[(size_t)elem * NPE], X, J);
  double* g = &G[t * 10];
#pragma unroll
  for (int k = 0; k < 9; k++) g[k] = J[k];
  g[9] = qp.w * det * telem[elem];
}

// Stress C e(u) (no thickness factor; the table's w det t carries it) at
// every (element, quadrature point), for the geometric stiffness:
// S[(e NQP + q) 6 + k], Voigt order of TACSSolidConstitutive::evalStress.
template <int ORD>
__global__ __launch_bounds__(256) void solid_qp_stress(int nelem, const int* __restrict__ conn,
                                                       const double* __restrict__ u,
                                                       const double* __restrict__ G,
                                                       SolidMaterial mat,
                                                       double* __restrict__ S) {
  constexpr int NPE = ORD * ORD * ORD, NQP = NPE;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nelem * NQP) return;
  const int elem = (int)(t / NQP), q = (int)(t - (size_t)elem * NQP);
  const SolidMaterial me = elem_mat(mat, elem);
  QP1D<ORD> qp;
  qp.eval(q);
  double J[9];
#pragma unroll
  for (int k = 0; k < 9; k++) J[k] = G[t * 10 + k];
  double Ux[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int* c = &conn[(size_t)elem * NPE];
  for (int a = 0; a < NPE; a++) {
    double N, g0[3], g[3];
    qp.basis(a, N, g0);
    grad_x(J, g0, g);
    const int na = c[a];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double uj = u[3 * (size_t)na + j];
      Ux[3 * j] += uj * g[0];
      Ux[3 * j + 1] += uj * g[1];
      Ux[3 * j + 2] += uj * g[2];
    }
  }
  const double e[6] = {Ux[0], Ux[4], Ux[8], Ux[5] + Ux[7], Ux[2] + Ux[6], Ux[1] + Ux[3]};
  double s[6];
  stress6(me, e, s);
#pragma unroll
  for (int k = 0; k < 6; k++) S[t * 6 + k] = s[k];
}

// ---------------------------------------------------------------------------
// Jacobian: one wavefront per block row.
// ---------------------------------------------------------------------------
template <int ORD>
struct JacCfg;
template <>
struct JacCfg<2> {
  static constexpr int MAXC = 64, WPB = 4;
};
template <>
struct JacCfg<3> {
  static constexpr int MAXC = 125, WPB = 2;
};

// ISO: isotropic tangent (lambda, mu closed form) or the general C[21]; two
// instantiations keep the general path's registers out of the isotropic one.
// GEO: adds geo times the geometric stiffness from the per-point stresses Sq
// (TACSLinearElasticity3D::evalWeakMatrix, TACS_GEOMETRIC_STIFFNESS_MATRIX,
// src/elements/TACSLinearElasticity.cpp:1704: coefficient sigma_de of
// du_j/dx_d du_j/dx_e).
// WPB wavefronts (rows) per workgroup and ECH elements per LDS geometry
// chunk are launch variants: the sums do not depend on them.
template <int ORD, bool ISO, bool GEO, int WPB = JacCfg<ORD>::WPB, int ECH = 8>
__global__ __launch_bounds__(64 * WPB) void solid_jacobian_rows(
    int nnodes, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
    const int* __restrict__ conn, const double* __restrict__ Gq, const double* __restrict__ Sq,
    SolidMaterial mat, double alpha, double gamma, double geo, const int* __restrict__ rowp,
    const int* __restrict__ cols,
    const int* __restrict__ bcvars, double* __restrict__ A,
    float* __restrict__ A32, int* __restrict__ status) {
  constexpr int NPE = ORD * ORD * ORD, NQP = NPE;
  constexpr int MAXC = JacCfg<ORD>::MAXC;
  __shared__ double geom[WPB][ECH * NQP][10];
  __shared__ double sgeo[WPB][GEO ? ECH * NQP : 1][6];
  __shared__ double rowbuf[WPB][MAXC * 9];
  __shared__ int colbuf[WPB][MAXC];

  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  double (*G)[10] = geom[wid];
  double (*SG)[6] = sgeo[wid];
  double* R = rowbuf[wid];
  int* CB = colbuf[wid];

  for (int row = blockIdx.x * WPB + wid; row < nnodes; row += gridDim.x * WPB) {
    const int rs = rowp[row];
    const int nc = rowp[row + 1] - rs;
    if (nc > MAXC) {
      if (lane == 0) atomicExch(status, TACS_HIP_ERR_ROW_TOO_LONG);
      continue;
    }
    for (int k = lane; k < nc; k += 64) CB[k] = cols[rs + k];
    for (int k = lane; k < nc * 9; k += 64) R[k] = 0.0;
    const int es = n2e_ptr[row];
    const int ne = n2e_ptr[row + 1] - es;
    WAVE_SYNC();

    for (int c0 = 0; c0 < ne; c0 += ECH) {
      const int nce = min(ECH, ne - c0);
      // Geometry of every (element, quadrature point) pair of the chunk,
      // copied from the per-assembly table (10 values each).
      for (int p = lane; p < nce * NQP * 10; p += 64) {
        int eq = p / 10, k = p - 10 * eq;
        int e = eq / NQP, q = eq - e * NQP;
        int elem = n2e[es + c0 + e] >> 5;
        G[eq][k] = Gq[((size_t)elem * NQP + q) * 10 + k];
      }
      if (GEO) {
        for (int p = lane; p < nce * NQP * 6; p += 64) {
          int eq = p / 6, k = p - 6 * eq;
          int e = eq / NQP, q = eq - e * NQP;
          int elem = n2e[es + c0 + e] >> 5;
          SG[eq][k] = Sq[((size_t)elem * NQP + q) * 6 + k];
        }
      }
      WAVE_SYNC();
      // (element, column node) pairs: accumulate the 3x3 block over qp.
      for (int p0 = 0; p0 < nce * NPE; p0 += 64) {
        int p = p0 + lane;
        bool valid = p < nce * NPE;
        int e = valid ? p / NPE : 0;
        int b = valid ? p - e * NPE : 0;
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        int slot = -1;
        double lam = mat.lam, mu = mat.mu, rho = mat.rho;
        if (valid) {
          int packed = n2e[es + c0 + e];
          int elem = packed >> 5, a = packed & 31;
          if (mat.elem_props) {
            const double* mp = mat.elem_props + 4 * (size_t)elem;
            lam = mp[0];
            mu = mp[1];
            rho = mp[2];
          }
          int cn = conn[(size_t)elem * NPE + b];
          int lo = 0, hi = nc - 1;
          while (lo <= hi) {
            int mid = (lo + hi) >> 1;
            int v = CB[mid];
            if (v == cn) {
              slot = mid;
              break;
            }
            if (v < cn) lo = mid + 1;
            else hi = mid - 1;
          }
          for (int q = 0; q < NQP; q++) {
            QP1D<ORD> qp;
            qp.eval(q);
            const double* Jq = G[e * NQP + q];
            double J[9];
#pragma unroll
            for (int k = 0; k < 9; k++) J[k] = Jq[k];
            double wdt = Jq[9];
            double Na, Nb, ga0[3], gb0[3], ga[3], gb[3];
            qp.basis(a, Na, ga0);
            qp.basis(b, Nb, gb0);
            grad_x(J, ga0, ga);
            grad_x(J, gb0, gb);
            double f = alpha * wdt;
            if (ISO) {
              double gab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
              for (int j = 0; j < 3; j++) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                  double v = lam * ga[j] * gb[k] + mu * gb[j] * ga[k];
                  if (j == k) v += mu * gab;
                  acc[3 * j + k] += f * v;
                }
              }
            } else {
#pragma unroll
              for (int j = 0; j < 3; j++) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                  double v = 0.0;
#pragma unroll
                  for (int d = 0; d < 3; d++) {
#pragma unroll
                    for (int ee = 0; ee < 3; ee++)
                      v += c21(mat, voigt(j, d), voigt(k, ee)) * ga[d] * gb[ee];
                  }
                  acc[3 * j + k] += f * v;
                }
              }
            }
            if (gamma != 0.0) {
              double m = gamma * rho * wdt * Na * Nb;
              acc[0] += m;
              acc[4] += m;
              acc[8] += m;
            }
            if (GEO) {
              // ga . sigma . gb, sigma = [s0 s5 s4; s5 s1 s3; s4 s3 s2]
              const double* sq = SG[e * NQP + q];
              const double t0 = sq[0] * gb[0] + sq[5] * gb[1] + sq[4] * gb[2];
              const double t1 = sq[5] * gb[0] + sq[1] * gb[1] + sq[3] * gb[2];
              const double t2 = sq[4] * gb[0] + sq[3] * gb[1] + sq[2] * gb[2];
              const double gg = geo * wdt * (ga[0] * t0 + ga[1] * t1 + ga[2] * t2);
              acc[0] += gg;
              acc[4] += gg;
              acc[8] += gg;
            }
          }
          if (slot < 0) atomicExch(status, TACS_HIP_ERR_MISSING_ENTRY);
        }
        // Add the blocks into the row in ascending element order.
        int efirst = p0 / NPE;
        int elast = min(nce * NPE, p0 + 64) - 1;
        elast = elast / NPE;
        for (int ee = efirst; ee <= elast; ee++) {
          if (valid && e == ee && slot >= 0) {
#pragma unroll
            for (int k = 0; k < 9; k++) R[slot * 9 + k] += acc[k];
          }
          WAVE_SYNC();
        }
      }
    }
    // Dirichlet rows (BCSRMat::zeroRow with ident = 1, BCSRMat.cpp:2027) and
    // the coalesced write of the row.
    const int bc = bcvars ? bcvars[row] : 0;
    double* Ar = A + (size_t)rs * 9;
    for (int k = lane; k < nc * 9; k += 64) {
      int s = k / 9, rc = k - 9 * s, r = rc / 3, cc = rc - 3 * r;
      double v = R[k];
      if (bc & (1 << r)) v = (CB[s] == row && r == cc) ? 1.0 : 0.0;
      if (A) Ar[k] = v;
      if (A32) A32[(size_t)rs * 9 + k] = (float)v;
    }
    WAVE_SYNC();
  }
}

// ---------------------------------------------------------------------------
// Per-element vectors: residual, compliance / KS state sensitivities.
// One thread per element.
// ---------------------------------------------------------------------------
enum { EV_RESIDUAL = 0, EV_COMPLIANCE_SV = 1, EV_KS_SV = 2 };

template <int ORD, int MODE>
__global__ __launch_bounds__(128) void solid_elem_vectors(
    int nelem, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ u, const double* __restrict__ telem,
    SolidMaterial mat, double coef, double ks_weight, double ks_max,
    double ks_sum, double safety, double* __restrict__ out) {
  constexpr int NPE = ORD * ORD * ORD, NQP = NPE, NV = 3 * NPE;
  int el = blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= nelem) return;
  const SolidMaterial me = elem_mat(mat, el);
  const int* c = &conn[(size_t)el * NPE];
  double* o = &out[(size_t)el * NV];
  double r[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) r[k] = 0.0;
  const double t = telem ? telem[el] : 1.0;
  for (int q = 0; q < NQP; q++) {
    QP1D<ORD> qp;
    qp.eval(q);
    double J[9];
    double det = elem_geometry<ORD>(qp, c, X, J);
    double Ux[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < NPE; b++) {
      double N, g0[3], g[3];
      qp.basis(b, N, g0);
      grad_x(J, g0, g);
      int nb = c[b];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        double ub = u[3 * nb + j];
        Ux[3 * j] += ub * g[0];
        Ux[3 * j + 1] += ub * g[1];
        Ux[3 * j + 2] += ub * g[2];
      }
    }
    double e[6] = {Ux[0], Ux[4], Ux[8], Ux[5] + Ux[7], Ux[2] + Ux[6], Ux[1] + Ux[3]};
    double s[6];
    stress6(me, e, s);
    double f;
    double sv[6];
    if (MODE == EV_RESIDUAL) {
      f = coef * qp.w * det * t;  // coef = 1
#pragma unroll
      for (int i = 0; i < 6; i++) sv[i] = s[i];
    } else if (MODE == EV_COMPLIANCE_SV) {
      f = 2.0 * coef * qp.w * det * t;  // coef = alpha
#pragma unroll
      for (int i = 0; i < 6; i++) sv[i] = s[i];
    } else {
      // KS failure: d KS/d e = dfdq * C * dvm/ds (TACSSolidConstitutive::
      // evalFailureStrainSens :301, vonMisesFailure3DStressSens
      // TACSMaterialProperties.cpp:449)
      double vm = sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[0] - s[2]) * (s[0] - s[2]) +
                              (s[1] - s[2]) * (s[1] - s[2]) +
                              6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
      double fail = safety * vm / me.ys;
      double dfdq = exp(ks_weight * (fail - ks_max)) / ks_sum * qp.w * det * safety;
      double ds[6] = {0, 0, 0, 0, 0, 0};
      if (vm != 0.0) {
        double fact = 0.5 / (me.ys * vm);
        ds[0] = fact * (2.0 * s[0] - s[1] - s[2]);
        ds[1] = fact * (2.0 * s[1] - s[0] - s[2]);
        ds[2] = fact * (2.0 * s[2] - s[0] - s[1]);
        ds[3] = 6.0 * fact * s[3];
        ds[4] = 6.0 * fact * s[4];
        ds[5] = 6.0 * fact * s[5];
      }
      stress6(me, ds, sv);
      f = coef * dfdq;
    }
#pragma unroll
    for (int a = 0; a < NPE; a++) {
      double N, g0[3], g[3];
      qp.basis(a, N, g0);
      grad_x(J, g0, g);
      r[3 * a] += f * (sv[0] * g[0] + sv[5] * g[1] + sv[4] * g[2]);
      r[3 * a + 1] += f * (sv[5] * g[0] + sv[1] * g[1] + sv[3] * g[2]);
      r[3 * a + 2] += f * (sv[4] * g[0] + sv[3] * g[1] + sv[2] * g[2]);
    }
  }
#pragma unroll
  for (int k = 0; k < NV; k++) o[k] = r[k];
}

// ---------------------------------------------------------------------------
// Node-coordinate sensitivities (one thread per element).  For an integrand
// w detXd W(Ux), Ux = Ud Xd^{-1}, the node derivative is
// w detXd (W I - sum_p Ux_p^T dW/dUx_p) grad_x N_a:
//   XV_COMPLIANCE  W = e.(tCe), dW/dUx = 2 sigma  (TACSCompliance::
//                  getElementXptSens src/functions/TACSCompliance.cpp:129 ->
//                  TACSElement3D::addPointQuantityXptSens TACSElement3D.cpp:618)
//   XV_KS          W = exp(rho (f - max)) / (rho S), f = safety vm / ys
//                  (TACSKSFailure::getElementXptSens)
//   XV_ADJ         W = e(psi).sigma(u), dW/dUx_u = sigma(psi), dW/dUx_psi =
//                  sigma(u) (TACSElement3D::addAdjResXptProduct :260)
// ---------------------------------------------------------------------------
enum { XV_COMPLIANCE = 0, XV_KS = 1, XV_ADJ = 2 };

__device__ __forceinline__ void sig3(const double s[6], double S[9]) {
  S[0] = s[0];
  S[4] = s[1];
  S[8] = s[2];
  S[5] = S[7] = s[3];
  S[2] = S[6] = s[4];
  S[1] = S[3] = s[5];
}

template <int ORD, int MODE>
__global__ __launch_bounds__(128) void solid_xpt_vectors(
    int nelem, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ u, const double* __restrict__ psi,
    const double* __restrict__ telem, SolidMaterial mat, double scale, double ks_weight,
    double ks_max, double ks_sum, double safety, double* __restrict__ out) {
  constexpr int NPE = ORD * ORD * ORD, NQP = NPE, NV = 3 * NPE;
  int el = blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= nelem) return;
  const SolidMaterial me = elem_mat(mat, el);
  const int* c = &conn[(size_t)el * NPE];
  double r[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) r[k] = 0.0;
  const double t = telem ? telem[el] : 1.0;
  for (int q = 0; q < NQP; q++) {
    QP1D<ORD> qp;
    qp.eval(q);
    double J[9];
    double det = elem_geometry<ORD>(qp, c, X, J);
    double Ux[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Px[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < NPE; b++) {
      double N, g0[3], g[3];
      qp.basis(b, N, g0);
      grad_x(J, g0, g);
      int nb = c[b];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        double ub = u[3 * nb + j];
        Ux[3 * j] += ub * g[0];
        Ux[3 * j + 1] += ub * g[1];
        Ux[3 * j + 2] += ub * g[2];
        if (MODE == XV_ADJ) {
          double pb = psi[3 * nb + j];
          Px[3 * j] += pb * g[0];
          Px[3 * j + 1] += pb * g[1];
          Px[3 * j + 2] += pb * g[2];
        }
      }
    }
    double e[6] = {Ux[0], Ux[4], Ux[8], Ux[5] + Ux[7], Ux[2] + Ux[6], Ux[1] + Ux[3]};
    double s[6];
    stress6(me, e, s);
    double W, M[9];
    if (MODE == XV_COMPLIANCE) {
      W = 0.0;
#pragma unroll
      for (int i = 0; i < 6; i++) {
        s[i] *= t;
        W += e[i] * s[i];
      }
      double S[9];
      sig3(s, S);
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int l = 0; l < 3; l++) {
          double v = (i == l) ? W : 0.0;
#pragma unroll
          for (int k = 0; k < 3; k++) v -= 2.0 * Ux[3 * k + i] * S[3 * k + l];
          M[3 * i + l] = v;
        }
    } else if (MODE == XV_KS) {
      double vm = sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[0] - s[2]) * (s[0] - s[2]) +
                              (s[1] - s[2]) * (s[1] - s[2]) +
                              6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
      double fail = safety * vm / me.ys;
      W = exp(ks_weight * (fail - ks_max)) / (ks_weight * ks_sum);
      double ds[6] = {0, 0, 0, 0, 0, 0}, de[6], S[9];
      if (vm != 0.0) {
        double fact = 0.5 / (me.ys * vm);
        ds[0] = fact * (2.0 * s[0] - s[1] - s[2]);
        ds[1] = fact * (2.0 * s[1] - s[0] - s[2]);
        ds[2] = fact * (2.0 * s[2] - s[0] - s[1]);
        ds[3] = 6.0 * fact * s[3];
        ds[4] = 6.0 * fact * s[4];
        ds[5] = 6.0 * fact * s[5];
      }
      stress6(me, ds, de);
      sig3(de, S);
      const double f = ks_weight * W * safety;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int l = 0; l < 3; l++) {
          double v = (i == l) ? W : 0.0;
#pragma unroll
          for (int k = 0; k < 3; k++) v -= f * Ux[3 * k + i] * S[3 * k + l];
          M[3 * i + l] = v;
        }
    } else {
      double ep[6] = {Px[0], Px[4], Px[8], Px[5] + Px[7], Px[2] + Px[6], Px[1] + Px[3]};
      double sp[6];
      stress6(me, ep, sp);
      W = 0.0;
#pragma unroll
      for (int i = 0; i < 6; i++) {
        s[i] *= t;
        sp[i] *= t;
        W += ep[i] * s[i];
      }
      double Su[9], Sp[9];
      sig3(s, Su);
      sig3(sp, Sp);
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int l = 0; l < 3; l++) {
          double v = (i == l) ? W : 0.0;
#pragma unroll
          for (int k = 0; k < 3; k++) v -= Ux[3 * k + i] * Sp[3 * k + l] + Px[3 * k + i] * Su[3 * k + l];
          M[3 * i + l] = v;
        }
    }
    const double f = scale * qp.w * det;
#pragma unroll
    for (int a = 0; a < NPE; a++) {
      double N, g0[3], g[3];
      qp.basis(a, N, g0);
      grad_x(J, g0, g);
#pragma unroll
      for (int i = 0; i < 3; i++)
        r[3 * a + i] += f * (M[3 * i] * g[0] + M[3 * i + 1] * g[1] + M[3 * i + 2] * g[2]);
    }
  }
  double* o = &out[(size_t)el * NV];
#pragma unroll
  for (int k = 0; k < NV; k++) o[k] = r[k];
}

// scale * psi^T dR/dX of coordinate-component traction aux elements: the
// exact derivative of -sum_q w A N_a t, A = |t1 x t2|:
// dA/dX_a = (t2 x n)(grad N_a . tg1) + (n x t1)(grad N_a . tg2).
template <int ORD>
__global__ void solid_traction_xpt_vectors(int naux, const int* __restrict__ aux_elem,
                                           const int* __restrict__ aux_face,
                                           const double* __restrict__ aux_trac,
                                           const int* __restrict__ conn,
                                           const double* __restrict__ X,
                                           const double* __restrict__ psi, double scale,
                                           double* __restrict__ out) {
  constexpr int NPE = ORD * ORD * ORD, NV = 3 * NPE;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= naux) return;
  const int* c = &conn[(size_t)aux_elem[q] * NPE];
  const int face = aux_face[q];
  const double tr[3] = {aux_trac[3 * q], aux_trac[3 * q + 1], aux_trac[3 * q + 2]};
  double* o = &out[(size_t)q * NV];
  for (int k = 0; k < NV; k++) o[k] = 0.0;
  const double tang[6][6] = {{0, 0, 1, 0, 1, 0}, {0, 1, 0, 0, 0, 1}, {1, 0, 0, 0, 0, 1},
                             {0, 0, 1, 1, 0, 0}, {0, 1, 0, 1, 0, 0}, {1, 0, 0, 0, 1, 0}};
  for (int n = 0; n < ORD * ORD; n++) {
    int ia = n % ORD, ib = n / ORD;
    double fixed = -1.0 + 2.0 * (face % 2);
    double pt[3];
    if (face / 2 == 0) {
      pt[0] = fixed; pt[1] = gauss_pt<ORD>(ia); pt[2] = gauss_pt<ORD>(ib);
    } else if (face / 2 == 1) {
      pt[0] = gauss_pt<ORD>(ia); pt[1] = fixed; pt[2] = gauss_pt<ORD>(ib);
    } else {
      pt[0] = gauss_pt<ORD>(ia); pt[1] = gauss_pt<ORD>(ib); pt[2] = fixed;
    }
    double w = gauss_wt<ORD>(ia) * gauss_wt<ORD>(ib);
    QP1D<ORD> qp;
    lag1d<ORD>(pt[0], qp.n[0], qp.d[0]);
    lag1d<ORD>(pt[1], qp.n[1], qp.d[1]);
    lag1d<ORD>(pt[2], qp.n[2], qp.d[2]);
    double Xd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, P = 0.0;
    for (int b = 0; b < NPE; b++) {
      double N, g[3];
      qp.basis(b, N, g);
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) Xd[3 * i + k] += X[3 * c[b] + i] * g[k];
      P += N * (psi[3 * c[b]] * tr[0] + psi[3 * c[b] + 1] * tr[1] + psi[3 * c[b] + 2] * tr[2]);
    }
    const double* tg = tang[face];
    double t1[3], t2[3];
    for (int i = 0; i < 3; i++) {
      t1[i] = Xd[3 * i] * tg[0] + Xd[3 * i + 1] * tg[1] + Xd[3 * i + 2] * tg[2];
      t2[i] = Xd[3 * i] * tg[3] + Xd[3 * i + 1] * tg[4] + Xd[3 * i + 2] * tg[5];
    }
    double nv[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2],
                    t1[0] * t2[1] - t1[1] * t2[0]};
    double A = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    double nh[3] = {nv[0] / A, nv[1] / A, nv[2] / A};
    double a1[3] = {t2[1] * nh[2] - t2[2] * nh[1], t2[2] * nh[0] - t2[0] * nh[2],
                    t2[0] * nh[1] - t2[1] * nh[0]};
    double a2[3] = {nh[1] * t1[2] - nh[2] * t1[1], nh[2] * t1[0] - nh[0] * t1[2],
                    nh[0] * t1[1] - nh[1] * t1[0]};
    const double f = -scale * w * P;
    for (int a = 0; a < NPE; a++) {
      double N, g[3];
      qp.basis(a, N, g);
      double g1 = g[0] * tg[0] + g[1] * tg[1] + g[2] * tg[2];
      double g2 = g[0] * tg[3] + g[1] * tg[4] + g[2] * tg[5];
      for (int i = 0; i < 3; i++) o[3 * a + i] += f * (a1[i] * g1 + a2[i] * g2);
    }
  }
}

// Traction aux elements (TACSTraction3D::addResidual, TACSTraction3D.cpp:87)
// and pressure aux elements (TACSPressure3D::addResidual, TACSPressure3D.cpp:60,
// traction -p n with n the outward unit normal): one thread per aux element;
// element vector out[aux * NV].
template <int ORD>
__global__ void solid_traction_vectors(int naux, const int* __restrict__ aux_elem,
                                       const int* __restrict__ aux_face,
                                       const double* __restrict__ aux_trac,
                                       const double* __restrict__ aux_pres,
                                       const int* __restrict__ conn,
                                       const double* __restrict__ X, double lambda,
                                       double* __restrict__ out) {
  constexpr int NPE = ORD * ORD * ORD, NV = 3 * NPE;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= naux) return;
  const int* c = &conn[(size_t)aux_elem[q] * NPE];
  const int face = aux_face[q];
  const double tr0[3] = {aux_trac[3 * q], aux_trac[3 * q + 1], aux_trac[3 * q + 2]};
  const double pres = aux_pres ? aux_pres[q] : 0.0;
  double* o = &out[(size_t)q * NV];
  for (int k = 0; k < NV; k++) o[k] = 0.0;
  // tangents ordered so that t1 x t2 is the outward normal
  const double tang[6][6] = {{0, 0, 1, 0, 1, 0}, {0, 1, 0, 0, 0, 1}, {1, 0, 0, 0, 0, 1},
                             {0, 0, 1, 1, 0, 0}, {0, 1, 0, 1, 0, 0}, {1, 0, 0, 0, 1, 0}};
  for (int n = 0; n < ORD * ORD; n++) {
    int ia = n % ORD, ib = n / ORD;
    double fixed = -1.0 + 2.0 * (face % 2);
    double pt[3];
    if (face / 2 == 0) {
      pt[0] = fixed; pt[1] = gauss_pt<ORD>(ia); pt[2] = gauss_pt<ORD>(ib);
    } else if (face / 2 == 1) {
      pt[0] = gauss_pt<ORD>(ia); pt[1] = fixed; pt[2] = gauss_pt<ORD>(ib);
    } else {
      pt[0] = gauss_pt<ORD>(ia); pt[1] = gauss_pt<ORD>(ib); pt[2] = fixed;
    }
    double w = gauss_wt<ORD>(ia) * gauss_wt<ORD>(ib);
    QP1D<ORD> qp;
    lag1d<ORD>(pt[0], qp.n[0], qp.d[0]);
    lag1d<ORD>(pt[1], qp.n[1], qp.d[1]);
    lag1d<ORD>(pt[2], qp.n[2], qp.d[2]);
    double Xd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < NPE; b++) {
      double N, g[3];
      qp.basis(b, N, g);
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) Xd[3 * i + k] += X[3 * c[b] + i] * g[k];
    }
    const double* tg = tang[face];
    double t1[3], t2[3];
    for (int i = 0; i < 3; i++) {
      t1[i] = Xd[3 * i] * tg[0] + Xd[3 * i + 1] * tg[1] + Xd[3 * i + 2] * tg[2];
      t2[i] = Xd[3 * i] * tg[3] + Xd[3 * i + 1] * tg[4] + Xd[3 * i + 2] * tg[5];
    }
    double nx = t1[1] * t2[2] - t1[2] * t2[1];
    double ny = t1[2] * t2[0] - t1[0] * t2[2];
    double nz = t1[0] * t2[1] - t1[1] * t2[0];
    const double area = sqrt(nx * nx + ny * ny + nz * nz);
    const double wa = lambda * w * area;
    // traction plus pressure: t - (-p n) enters the residual as -(t + p n)
    const double pa = area > 0.0 ? pres / area : 0.0;
    const double tr[3] = {tr0[0] + pa * nx, tr0[1] + pa * ny, tr0[2] + pa * nz};
    for (int a = 0; a < NPE; a++) {
      double N, g[3];
      qp.basis(a, N, g);
      for (int k = 0; k < 3; k++) o[3 * a + k] -= wa * N * tr[k];
    }
  }
}

// ---------------------------------------------------------------------------
// Matrix-free Jacobian-vector product, element part (TACSElement3D::
// addMatVecProduct src/elements/TACSElement3D.cpp:428 through
// TACSElementBasis::addMatVecProduct basis/TACSElementBasis.cpp:940):
//   ev[e] = sum_q  Bq^T (kc w det t C) Bq x_e  +  mc rho w det t Nq^T Nq x_e
// from the quadrature-point geometry table of solid_qp_geometry (J^-1 and
// w det t, 10 values per point) behind a 4-value header [kc, mc, 0, 0].
// A group of L lanes owns one element: lane q interpolates the displacement
// gradient at quadrature point q and turns it into the stress, pulled back
// to reference coordinates, in LDS; lane a then contracts it with the basis
// gradients of node a.  No element matrix is formed.
// ---------------------------------------------------------------------------
template <int ORD>
struct MFCfg;
template <>
struct MFCfg<2> {
  static constexpr int L = 8;
};
template <>
struct MFCfg<3> {
  static constexpr int L = 32;
};

template <int ORD, bool ISO>
__global__ __launch_bounds__(256) void solid_matfree_ev(int nelem, const int* __restrict__ conn,
                                                        const double* __restrict__ data,
                                                        SolidMaterial mat,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ ev) {
  constexpr int NPE = ORD * ORD * ORD, NQP = NPE;
  constexpr int L = MFCfg<ORD>::L, EPB = 256 / L;
  constexpr int UK = ORD == 2 ? 8 : 1;  // unrolling of the quadrature / node loops
  __shared__ double xs[EPB][NPE * 3];
  __shared__ double ss[EPB][NQP][12];
  const int g = threadIdx.x / L;
  const int l = threadIdx.x - g * L;
  const double kc = data[0], mc = data[1];
  const double* __restrict__ Gq = data + 4;
  for (int e0 = blockIdx.x * EPB; e0 < nelem; e0 += gridDim.x * EPB) {
    const int elem = e0 + g;
    const bool ok = elem < nelem;
    const SolidMaterial me = elem_mat(mat, ok ? elem : 0);
    if (ok && l < NPE) {
      const int n = conn[(size_t)elem * NPE + l];
      xs[g][3 * l] = x[3 * (size_t)n];
      xs[g][3 * l + 1] = x[3 * (size_t)n + 1];
      xs[g][3 * l + 2] = x[3 * (size_t)n + 2];
    }
    WAVE_SYNC();
    if (ok && l < NQP) {
      QP1D<ORD> qp;
      qp.eval(l);
      // U and dU/dxi by sum factorisation over the tensor-product nodes:
      // for each (ja, ka) line, t0 = sum_ia x n_ia(xi), t1 = sum_ia x n'_ia(xi)
      double U[3] = {0, 0, 0}, Ud[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll UK
      for (int ka = 0; ka < ORD; ka++) {
        const double nz = pick<ORD>(qp.n[2], ka), dz = pick<ORD>(qp.d[2], ka);
#pragma unroll
        for (int ja = 0; ja < ORD; ja++) {
          double t0[3] = {0, 0, 0}, t1[3] = {0, 0, 0};
#pragma unroll
          for (int ia = 0; ia < ORD; ia++) {
            const double* xa = &xs[g][3 * (ia + ORD * (ja + ORD * ka))];
#pragma unroll
            for (int c = 0; c < 3; c++) {
              t0[c] += qp.n[0][ia] * xa[c];
              t1[c] += qp.d[0][ia] * xa[c];
            }
          }
          const double w = qp.n[1][ja] * nz, wy = qp.d[1][ja] * nz, wz = qp.n[1][ja] * dz;
#pragma unroll
          for (int c = 0; c < 3; c++) {
            U[c] += w * t0[c];
            Ud[3 * c] += w * t1[c];
            Ud[3 * c + 1] += wy * t0[c];
            Ud[3 * c + 2] += wz * t0[c];
          }
        }
      }
      const double2* G2 = reinterpret_cast<const double2*>(Gq + ((size_t)elem * NQP + l) * 10);
      double J[10];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        const double2 v = G2[k];
        J[2 * k] = v.x;
        J[2 * k + 1] = v.y;
      }
      const double wdt = J[9];
      double Ux[9];
#pragma unroll
      for (int c = 0; c < 3; c++) grad_x(J, &Ud[3 * c], &Ux[3 * c]);
      const double e[6] = {Ux[0], Ux[4], Ux[8], Ux[5] + Ux[7], Ux[2] + Ux[6], Ux[1] + Ux[3]};
      double s[6];
      if (ISO) {
        const double tr = e[0] + e[1] + e[2];
        s[0] = me.lam * tr + 2.0 * me.mu * e[0];
        s[1] = me.lam * tr + 2.0 * me.mu * e[1];
        s[2] = me.lam * tr + 2.0 * me.mu * e[2];
        s[3] = me.mu * e[3];
        s[4] = me.mu * e[4];
        s[5] = me.mu * e[5];
      } else {
        stress6(me, e, s);
      }
      const double f = kc * wdt;
      // physical stress tensor rows [s0 s5 s4; s5 s1 s3; s4 s3 s2], scaled
      const double P[9] = {f * s[0], f * s[5], f * s[4], f * s[5], f * s[1],
                           f * s[3], f * s[4], f * s[3], f * s[2]};
      // reference-coordinate pull-back: S[c][k] = sum_d P[c][d] J[3k + d]
#pragma unroll
      for (int c = 0; c < 3; c++) {
#pragma unroll
        for (int k = 0; k < 3; k++)
          ss[g][l][3 * c + k] =
              P[3 * c] * J[3 * k] + P[3 * c + 1] * J[3 * k + 1] + P[3 * c + 2] * J[3 * k + 2];
      }
      const double m = mc * me.rho * wdt;
      ss[g][l][9] = m * U[0];
      ss[g][l][10] = m * U[1];
      ss[g][l][11] = m * U[2];
    }
    WAVE_SYNC();
    if (ok && l < NPE) {
      // node l's 1D Lagrange values and slopes at the Gauss coordinates,
      // evaluated once; the quadrature loop then only multiplies them
      const int ia = l % ORD, ja = (l / ORD) % ORD, ka = l / (ORD * ORD);
      double nI[ORD], dI[ORD], nJ[ORD], dJ[ORD], nK[ORD], dK[ORD];
#pragma unroll
      for (int r = 0; r < ORD; r++) {
        double n1[ORD], d1[ORD];
        lag1d<ORD>(gauss_pt<ORD>(r), n1, d1);
        nI[r] = pick<ORD>(n1, ia);
        dI[r] = pick<ORD>(d1, ia);
        nJ[r] = pick<ORD>(n1, ja);
        dJ[r] = pick<ORD>(d1, ja);
        nK[r] = pick<ORD>(n1, ka);
        dK[r] = pick<ORD>(d1, ka);
      }
      double o0 = 0.0, o1 = 0.0, o2 = 0.0;
#pragma unroll UK
      for (int kq = 0; kq < ORD; kq++) {
        const double nk = pick<ORD>(nK, kq), dk = pick<ORD>(dK, kq);
#pragma unroll
        for (int jq = 0; jq < ORD; jq++) {
          const double njk = nJ[jq] * nk, djk = dJ[jq] * nk, nzk = nJ[jq] * dk;
#pragma unroll
          for (int iq = 0; iq < ORD; iq++) {
            const double* S = ss[g][iq + ORD * (jq + ORD * kq)];
            const double g0 = dI[iq] * njk, g1 = nI[iq] * djk, g2 = nI[iq] * nzk;
            const double N = nI[iq] * njk;
            o0 += g0 * S[0] + g1 * S[1] + g2 * S[2] + N * S[9];
            o1 += g0 * S[3] + g1 * S[4] + g2 * S[5] + N * S[10];
            o2 += g0 * S[6] + g1 * S[7] + g2 * S[8] + N * S[11];
          }
        }
      }
      double* o = &ev[((size_t)elem * NPE + l) * 3];
      o[0] = o0;
      o[1] = o1;
      o[2] = o2;
    }
    WAVE_SYNC();
  }
}

// Header of the matrix-free data: [kc, mc, 0, 0].
__global__ void matfree_header(double* data, double kc, double mc) {
  if (threadIdx.x == 0) {
    data[0] = kc;
    data[1] = mc;
    data[2] = 0.0;
    data[3] = 0.0;
  }
}

// Node gather of element vectors in ascending element order:
// out[3*i + j] (+)= sum_{(e,a) in n2e(i)} ev[e*NV + 3a + j].
template <int NV>
__global__ void gather_elem_vectors(int nnodes, const int* __restrict__ n2e_ptr,
                                    const int* __restrict__ n2e,
                                    const double* __restrict__ ev, int accumulate,
                                    double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (accumulate) {
    s0 = out[3 * i];
    s1 = out[3 * i + 1];
    s2 = out[3 * i + 2];
  }
  for (int p = n2e_ptr[i]; p < n2e_ptr[i + 1]; p++) {
    int packed = n2e[p];
    const double* v = &ev[(size_t)(packed >> 5) * NV + 3 * (packed & 31)];
    s0 += v[0];
    s1 += v[1];
    s2 += v[2];
  }
  out[3 * i] = s0;
  out[3 * i + 1] = s1;
  out[3 * i + 2] = s2;
}

// ---------------------------------------------------------------------------
// Per-element scalars: compliance integrand, KS max / sum, DV sensitivity
// (e.Ce) and adjoint-residual product (e(u).C e(psi)).
// ---------------------------------------------------------------------------
enum { ES_COMPLIANCE = 0, ES_KS_MAX = 1, ES_KS_SUM = 2, ES_DV_COMPLIANCE = 3, ES_ADJ_PRODUCT = 4 };

template <int ORD, int MODE>
__global__ __launch_bounds__(128) void solid_elem_scalars(
    int nelem, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ u, const double* __restrict__ psi,
    const double* __restrict__ telem, SolidMaterial mat, double scale,
    double ks_weight, double ks_max, double safety, double* __restrict__ out) {
  constexpr int NPE = ORD * ORD * ORD, NQP = NPE;
  int el = blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= nelem) return;
  const SolidMaterial me = elem_mat(mat, el);
  const int* c = &conn[(size_t)el * NPE];
  const double t = telem ? telem[el] : 1.0;
  double acc = (MODE == ES_KS_MAX) ? -1e20 : 0.0;
  for (int q = 0; q < NQP; q++) {
    QP1D<ORD> qp;
    qp.eval(q);
    double J[9];
    double det = elem_geometry<ORD>(qp, c, X, J);
    double Ux[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double Px[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < NPE; b++) {
      double N, g0[3], g[3];
      qp.basis(b, N, g0);
      grad_x(J, g0, g);
      int nb = c[b];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        double ub = u[3 * nb + j];
        Ux[3 * j] += ub * g[0];
        Ux[3 * j + 1] += ub * g[1];
        Ux[3 * j + 2] += ub * g[2];
        if (MODE == ES_ADJ_PRODUCT) {
          double pb = psi[3 * nb + j];
          Px[3 * j] += pb * g[0];
          Px[3 * j + 1] += pb * g[1];
          Px[3 * j + 2] += pb * g[2];
        }
      }
    }
    double e[6] = {Ux[0], Ux[4], Ux[8], Ux[5] + Ux[7], Ux[2] + Ux[6], Ux[1] + Ux[3]};
    double s[6];
    stress6(me, e, s);
    double wd = qp.w * det;
    if (MODE == ES_COMPLIANCE) {
      acc += wd * t * (e[0] * s[0] + e[1] * s[1] + e[2] * s[2] + e[3] * s[3] + e[4] * s[4] + e[5] * s[5]);
    } else if (MODE == ES_DV_COMPLIANCE) {
      acc += scale * wd * (e[0] * s[0] + e[1] * s[1] + e[2] * s[2] + e[3] * s[3] + e[4] * s[4] + e[5] * s[5]);
    } else if (MODE == ES_ADJ_PRODUCT) {
      double ph[6] = {Px[0], Px[4], Px[8], Px[5] + Px[7], Px[2] + Px[6], Px[1] + Px[3]};
      acc += scale * wd * (ph[0] * s[0] + ph[1] * s[1] + ph[2] * s[2] + ph[3] * s[3] + ph[4] * s[4] + ph[5] * s[5]);
    } else {
      double vm = sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[0] - s[2]) * (s[0] - s[2]) +
                              (s[1] - s[2]) * (s[1] - s[2]) +
                              6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
      double fail = safety * vm / me.ys;
      if (MODE == ES_KS_MAX) acc = fmax(acc, fail);
      else acc += wd * exp(ks_weight * (fail - ks_max));
    }
  }
  out[el] = acc;
}

// ---------------------------------------------------------------------------
// Design variables: t_e = dv[tNum_e] (or the fixed t when tNum_e < 0), and
// the reverse reduction of per-element DV derivatives into the design vector
// in ascending element order (dfdx->setValues ADD, TACSAssembler.cpp:4811).
// ---------------------------------------------------------------------------
__global__ void gather_design_vars(int nelem, const int* __restrict__ dvnum,
                                   const double* __restrict__ tfixed,
                                   const double* __restrict__ dv, double* __restrict__ telem) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  int d = dvnum[e];
  telem[e] = (d >= 0 && dv) ? dv[d] : tfixed[e];
}

__global__ void scatter_dv_sens(int ndv, const int* __restrict__ dv2e_ptr,
                                const int* __restrict__ dv2e, const double* __restrict__ es,
                                double* __restrict__ dfdx) {
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndv) return;
  double s = 0.0;
  for (int p = dv2e_ptr[d]; p < dv2e_ptr[d + 1]; p++) s += es[dv2e[p]];
  dfdx[d] += s;
}

// Dirichlet values on a vector (TACSBVec::applyBCs src/bpmat/TACSBVec.cpp:546):
// constrained entries become u - lambda*value (u given) or 0.
__global__ void apply_bc_vec(int b, int nbc, const int* __restrict__ nodes,
                             const int* __restrict__ masks, const double* __restrict__ vals,
                             const double* __restrict__ u, double lambda, double* __restrict__ x) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nbc) return;
  int node = nodes[q], m = masks[q];
  for (int k = 0; k < b; k++) {
    if (m & (1 << k)) {
      size_t i = (size_t)b * node + k;
      x[i] = u ? u[i] - lambda * (vals ? vals[b * q + k] : 0.0) : 0.0;
    }
  }
}

// Host-side launchers (C ABI, raw device pointers, explicit stream).
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Element-centric Jacobian.  A team of threads owns one element: it stages
// the element's nodes, the inverse Jacobian and w det(J) t of every
// quadrature point, and the physical gradients (and values) of every basis
// function at every point in LDS, then each thread accumulates whole 3x3
// blocks K_ab (local nodes a, b) over the points in registers.  T is the
// arithmetic type of the element (double, or float for the FP32 element
// kernels).  The blocks go to ke[e][a NPE + b][9] (row-major blocks); the
// row gather below sums them into block-CSR in a fixed order.  Isotropic
// tangent (lambda, mu) plus the consistent mass; the general C[21] and the
// geometric stiffness stay with the row kernel above.
// ---------------------------------------------------------------------------
template <int ORD>
struct EBCfg;
template <>
struct EBCfg<2> {  // 64 blocks: one per thread of a wavefront, 4 elements per group
  static constexpr int TEAM = 64, EPB = 4;
};
template <>
struct EBCfg<3> {  // 729 blocks over 256 threads (2 or 3 each), one element per group
  static constexpr int TEAM = 256, EPB = 1;
};

template <int ORD, typename T>
struct EBLds {
  static constexpr int NPE = ORD * ORD * ORD, NQP = NPE;
  T X[NPE][3];
  T J[NQP][9];
  T wdt[NQP];
  T g[NQP][3][NPE];  // component k of the gradient of basis function a at point q
  T N[NQP][NPE];
};

template <int ORD, typename T>
__device__ __forceinline__ void lag1d_t(T x, T* n, T* d) {
  if (ORD == 2) {
    n[0] = T(0.5) * (T(1) - x);
    n[1] = T(0.5) * (T(1) + x);
    d[0] = T(-0.5);
    d[1] = T(0.5);
  } else {
    n[0] = T(-0.5) * x * (T(1) - x);
    n[1] = (T(1) - x) * (T(1) + x);
    n[2] = T(0.5) * (T(1) + x) * x;
    d[0] = T(-0.5) + x;
    d[1] = T(-2) * x;
    d[2] = T(0.5) + x;
  }
}

// value and reference gradient of basis function a at point q
template <int ORD, typename T>
__device__ __forceinline__ void basis_t(int q, int a, T& N, T g[3]) {
  T n[3][ORD], d[3][ORD];
  lag1d_t<ORD, T>(T(gauss_pt<ORD>(q % ORD)), n[0], d[0]);
  lag1d_t<ORD, T>(T(gauss_pt<ORD>((q / ORD) % ORD)), n[1], d[1]);
  lag1d_t<ORD, T>(T(gauss_pt<ORD>(q / (ORD * ORD))), n[2], d[2]);
  const int ia = a % ORD, ja = (a / ORD) % ORD, ka = a / (ORD * ORD);
  T n0 = n[0][0], n1 = n[1][0], n2 = n[2][0], d0 = d[0][0], d1 = d[1][0], d2 = d[2][0];
#pragma unroll
  for (int i = 1; i < ORD; i++) {
    if (ia == i) { n0 = n[0][i]; d0 = d[0][i]; }
    if (ja == i) { n1 = n[1][i]; d1 = d[1][i]; }
    if (ka == i) { n2 = n[2][i]; d2 = d[2][i]; }
  }
  N = n0 * n1 * n2;
  g[0] = d0 * n1 * n2;
  g[1] = n0 * d1 * n2;
  g[2] = n0 * n1 * d2;
}

template <typename T>
__device__ __forceinline__ T inv3_t(const T A[9], T B[9]) {
  T det = A[8] * (A[0] * A[4] - A[3] * A[1]) - A[7] * (A[0] * A[5] - A[3] * A[2]) +
          A[6] * (A[1] * A[5] - A[2] * A[4]);
  T r = T(1) / det;
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

template <int ORD, typename T>
__global__ __launch_bounds__(EBCfg<ORD>::TEAM* EBCfg<ORD>::EPB) void solid_elem_blocks(
    int nelem, const int* __restrict__ conn, const double* __restrict__ Xg,
    const double* __restrict__ telem, SolidMaterial mat, double alpha, double gamma,
    T* __restrict__ ke) {
  constexpr int NPE = ORD * ORD * ORD, NQP = NPE, NB = NPE * NPE;
  constexpr int TEAM = EBCfg<ORD>::TEAM, EPB = EBCfg<ORD>::EPB;
  constexpr int BPT = (NB + TEAM - 1) / TEAM;
  __shared__ EBLds<ORD, T> lds[EPB];
  const int team = threadIdx.x / TEAM, tid = threadIdx.x % TEAM;
  const int el = blockIdx.x * EPB + team;
  const bool active = el < nelem;
  const SolidMaterial me = elem_mat(mat, active ? el : 0);
  EBLds<ORD, T>& L = lds[team];
  const int* c = &conn[(size_t)(active ? el : 0) * NPE];

  if (active)
    for (int k = tid; k < NPE * 3; k += TEAM) L.X[k / 3][k % 3] = T(Xg[3 * (size_t)c[k / 3] + k % 3]);
  __syncthreads();
  // inverse Jacobian and w det(J) t at every point
  if (active) {
    const T te = T(telem[el]);
    for (int q = tid; q < NQP; q += TEAM) {
      T Xd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int b = 0; b < NPE; b++) {
        T N, g[3];
        basis_t<ORD, T>(q, b, N, g);
#pragma unroll
        for (int r = 0; r < 3; r++) {
          const T x = L.X[b][r];
          Xd[3 * r] += x * g[0];
          Xd[3 * r + 1] += x * g[1];
          Xd[3 * r + 2] += x * g[2];
        }
      }
      T Ji[9];
      const T det = inv3_t<T>(Xd, Ji);
#pragma unroll
      for (int k = 0; k < 9; k++) L.J[q][k] = Ji[k];
      const T w = T(gauss_wt<ORD>(q % ORD) * gauss_wt<ORD>((q / ORD) % ORD) *
                    gauss_wt<ORD>(q / (ORD * ORD)));
      L.wdt[q] = w * det * te;
    }
  }
  __syncthreads();
  // physical gradients g = J^-T dN/dxi and values of every basis function
  if (active) {
    for (int p = tid; p < NQP * NPE; p += TEAM) {
      const int q = p / NPE, a = p - q * NPE;
      T N, g0[3];
      basis_t<ORD, T>(q, a, N, g0);
      const T* J = L.J[q];
      L.g[q][0][a] = g0[0] * J[0] + g0[1] * J[3] + g0[2] * J[6];
      L.g[q][1][a] = g0[0] * J[1] + g0[1] * J[4] + g0[2] * J[7];
      L.g[q][2][a] = g0[0] * J[2] + g0[1] * J[5] + g0[2] * J[8];
      L.N[q][a] = N;
    }
  }
  __syncthreads();
  if (!active) return;
  // K_ab = sum_q alpha w det t [lam g_a g_b^T + mu g_b g_a^T + mu (g_a.g_b) I]
  //        + gamma rho w det t N_a N_b I
  // (TACSLinearElasticity3D::evalWeakMatrix, TACSLinearElasticity.cpp:1158)
  T acc[BPT][9];
#pragma unroll
  for (int r = 0; r < BPT; r++)
#pragma unroll
    for (int k = 0; k < 9; k++) acc[r][k] = T(0);
  const T lam = T(me.lam), mu = T(me.mu), al = T(alpha), gr = T(gamma * me.rho);
  const bool mass = gamma != 0.0;
  for (int q = 0; q < NQP; q++) {
    const T f = al * L.wdt[q];
#pragma unroll
    for (int r = 0; r < BPT; r++) {
      const int p = tid + r * TEAM;
      if (BPT * TEAM == NB || p < NB) {
        const int a = p / NPE, b = p - a * NPE;
        const T ga[3] = {f * L.g[q][0][a], f * L.g[q][1][a], f * L.g[q][2][a]};
        const T gb[3] = {L.g[q][0][b], L.g[q][1][b], L.g[q][2][b]};
        const T la[3] = {lam * ga[0], lam * ga[1], lam * ga[2]};
        const T mb[3] = {mu * gb[0], mu * gb[1], mu * gb[2]};
        T d = mu * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
        if (mass) d += gr * L.wdt[q] * L.N[q][a] * L.N[q][b];
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int k = 0; k < 3; k++) acc[r][3 * j + k] += la[j] * gb[k] + mb[j] * ga[k];
        acc[r][0] += d;
        acc[r][4] += d;
        acc[r][8] += d;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < BPT; r++) {
    const int p = tid + r * TEAM;
    if (BPT * TEAM == NB || p < NB) {
      T* o = &ke[((size_t)el * NB + p) * 9];
#pragma unroll
      for (int k = 0; k < 9; k++) o[k] = acc[r][k];
    }
  }
}

// Block-CSR rows from element blocks: one wavefront per row (node a).  The
// row's incident (element e, local node a) pairs are taken in ascending
// element order (the order the reference's element loop adds them,
// src/TACSAssembler.cpp:4352-4383), CH pairs at a time: the wave loads the
// CH element block rows ke[e][a npe .. a npe + npe) (npe contiguous 3x3
// blocks each: coalesced runs, all loads in flight together), finds the
// row's slot of every column conn[e][b], then adds the pairs one after the
// other into the row accumulated in FP64 in LDS.  Dirichlet rows fused
// (BCSRMat::zeroRow ident = 1, src/bpmat/BCSRMat.cpp:2027); the row is
// written once, coalesced.
template <int NPE>
struct GatherCfg {
  static constexpr int MAXC = NPE == 8 ? 64 : 125, WPB = 4, CH = 8;
  static constexpr int NV = NPE * 9, PER = (NV + 63) / 64;  // values of a block row, per lane
};
template <int NPE, typename T>
__global__ __launch_bounds__(64 * GatherCfg<NPE>::WPB) void gather_solid_rows(
    int nrows, const int* __restrict__ rowp, const int* __restrict__ cols,
    const int* __restrict__ n2e_ptr, const int* __restrict__ n2e, const int* __restrict__ conn,
    const T* __restrict__ ke, const int* __restrict__ bcvars, double* __restrict__ A,
    float* __restrict__ A32, int* __restrict__ status) {
  constexpr int MAXC = GatherCfg<NPE>::MAXC, WPB = GatherCfg<NPE>::WPB, CH = GatherCfg<NPE>::CH;
  constexpr int NV = GatherCfg<NPE>::NV, PER = GatherCfg<NPE>::PER;
  __shared__ double rowbuf[WPB][MAXC * 9];
  __shared__ int colbuf[WPB][MAXC];
  __shared__ short smap[WPB][CH * NPE];  // slots < MAXC <= 125
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double* R = rowbuf[wid];
  int* CB = colbuf[wid];
  short* SM = smap[wid];
  for (int row = blockIdx.x * WPB + wid; row < nrows; row += gridDim.x * WPB) {
    const int rs = rowp[row];
    const int nc = rowp[row + 1] - rs;
    if (nc > MAXC) {
      if (lane == 0) atomicExch(status, TACS_HIP_ERR_ROW_TOO_LONG);
      continue;
    }
    for (int k = lane; k < nc; k += 64) CB[k] = cols[rs + k];
    for (int k = lane; k < nc * 9; k += 64) R[k] = 0.0;
    const int es = n2e_ptr[row];
    const int ne = n2e_ptr[row + 1] - es;
    WAVE_SYNC();
    for (int p0 = 0; p0 < ne; p0 += CH) {
      const int np = min(CH, ne - p0);
      // the chunk's block rows, every load issued before the first use
      T v[CH][PER];
#pragma unroll
      for (int c = 0; c < CH; c++) {
        const int packed = c < np ? n2e[es + p0 + c] : 0;
        const T* src = ke + ((size_t)(packed >> 5) * NPE + (packed & 31)) * NPE * 9;
#pragma unroll
        for (int r = 0; r < PER; r++) {
          const int q = lane + 64 * r;
          v[c][r] = (c < np && (NV % 64 == 0 || q < NV)) ? src[q] : T(0);
        }
      }
      // slot of column conn[e][b] for (pair c, block b) = t: t = c npe + b
      for (int t = lane; t < np * NPE; t += 64) {
        const int c = t / NPE, b = t - c * NPE;
        const int cn = conn[(size_t)(n2e[es + p0 + c] >> 5) * NPE + b];
        int lo = 0, hi = nc - 1, sl = -1;
        while (lo <= hi) {
          const int mid = (lo + hi) >> 1, cv = CB[mid];
          if (cv == cn) {
            sl = mid;
            break;
          }
          if (cv < cn) lo = mid + 1;
          else hi = mid - 1;
        }
        if (sl < 0) atomicExch(status, TACS_HIP_ERR_MISSING_ENTRY);
        SM[t] = (short)sl;
      }
      WAVE_SYNC();
#pragma unroll
      for (int c = 0; c < CH; c++) {
        if (c < np) {
#pragma unroll
          for (int r = 0; r < PER; r++) {
            const int q = lane + 64 * r;
            if (NV % 64 == 0 || q < NV) {
              const int b = q / 9, k = q - 9 * b;
              const int sl = SM[c * NPE + b];
              if (sl >= 0) R[sl * 9 + k] += (double)v[c][r];
            }
          }
        }
        WAVE_SYNC();
      }
    }
    const int bc = bcvars ? bcvars[row] : 0;
    for (int k = lane; k < nc * 9; k += 64) {
      const int s = k / 9, rc = k - 9 * s, r = rc / 3, cc = rc - 3 * r;
      double val = R[k];
      if (bc & (1 << r)) val = (CB[s] == row && r == cc) ? 1.0 : 0.0;
      if (A) A[(size_t)rs * 9 + k] = val;
      if (A32) A32[(size_t)rs * 9 + k] = (float)val;
    }
    WAVE_SYNC();
  }
}

// Launch variants (WPB, ECH) of the isotropic stiffness / mass rows kernel:
// 8-node elements 0 (4, 8), 1 (2, 8), 2 (8, 4), 3 (4, 4), 4 (4, 16);
// 27-node elements 0 (2, 8), 1 (1, 8), 2 (2, 4), 3 (1, 4).
static int jac_rows_nvariants(int order) { return order == 2 ? 5 : order == 3 ? 4 : 0; }

template <int O>
static void jac_rows_iso_launch(int variant, int nnodes, const int* n2e_ptr, const int* n2e,
                                const int* conn, const double* G, const SolidMaterial* mat,
                                double kc, double mc, const int* rowp, const int* cols,
                                const int* bcvars, double* A, float* A32, int* status,
                                hipStream_t stream) {
#define JAC_V(W, E)                                                                              \
  hipLaunchKernelGGL((solid_jacobian_rows<O, true, false, W, E>),                                \
                     dim3(tacs_grid_for(nnodes, W, 1 << 20)), dim3(64 * W), 0, stream, nnodes,   \
                     n2e_ptr, n2e, conn, G, (const double*)NULL, *mat, kc, mc, 0.0, rowp, cols,  \
                     bcvars, A, A32, status)
  if constexpr (O == 2) {
    switch (variant) {
      case 1: JAC_V(2, 8); break;
      case 2: JAC_V(8, 4); break;
      case 3: JAC_V(4, 4); break;
      case 4: JAC_V(4, 16); break;
      default: JAC_V(4, 8); break;
    }
  } else {
    switch (variant) {
      case 1: JAC_V(1, 8); break;
      case 2: JAC_V(2, 4); break;
      case 3: JAC_V(1, 4); break;
      default: JAC_V(2, 8); break;
    }
  }
#undef JAC_V
}

static int solid_mat_type_impl(int rows_variant, int order, int nnodes, int nelem,
                               const int* n2e_ptr, const int* n2e, const int* conn,
                               const double* X, const double* u, const double* telem,
                               const SolidMaterial* mat, double kc, double mc, double gc,
                               const int* rowp, const int* cols, const int* bcvars, double* G,
                               double* S, double* A, float* A32, int* status,
                               hipStream_t stream) {
  if (order != 2 && order != 3) return TACS_HIP_ERR_ARG;
  if (gc != 0.0 && (!u || !S)) return TACS_HIP_ERR_ARG;
  const int nqp = order * order * order;
  const size_t nq = (size_t)nelem * nqp;
  const dim3 qgrid((unsigned)((nq + 255) / 256));
  if (nq > 0) {
    if (order == 2)
      hipLaunchKernelGGL(solid_qp_geometry<2>, qgrid, dim3(256), 0, stream, nelem, conn, X, telem,
                         G);
    else
      hipLaunchKernelGGL(solid_qp_geometry<3>, qgrid, dim3(256), 0, stream, nelem, conn, X, telem,
                         G);
    if (gc != 0.0) {
      if (order == 2)
        hipLaunchKernelGGL(solid_qp_stress<2>, qgrid, dim3(256), 0, stream, nelem, conn, u, G,
                           *mat, S);
      else
        hipLaunchKernelGGL(solid_qp_stress<3>, qgrid, dim3(256), 0, stream, nelem, conn, u, G,
                           *mat, S);
    }
  }
#define JAC_LAUNCH(O, I, GE)                                                                   \
  hipLaunchKernelGGL((solid_jacobian_rows<O, I, GE>),                                          \
                     dim3(tacs_grid_for(nnodes, JacCfg<O>::WPB, 1 << 20)),                     \
                     dim3(64 * JacCfg<O>::WPB), 0, stream, nnodes, n2e_ptr, n2e, conn, G, S,   \
                     *mat, kc, mc, gc, rowp, cols, bcvars, A, A32, status)
#define JAC_ISO(O, GE)            \
  if (mat->isotropic)             \
    JAC_LAUNCH(O, true, GE);      \
  else                            \
    JAC_LAUNCH(O, false, GE)
  const bool ge = gc != 0.0;
  if (!ge && mat->isotropic) {
    if (order == 2)
      jac_rows_iso_launch<2>(rows_variant, nnodes, n2e_ptr, n2e, conn, G, mat, kc, mc, rowp, cols,
                             bcvars, A, A32, status, stream);
    else
      jac_rows_iso_launch<3>(rows_variant, nnodes, n2e_ptr, n2e, conn, G, mat, kc, mc, rowp, cols,
                             bcvars, A, A32, status, stream);
    return (int)hipGetLastError();
  }
  if (order == 2) {
    if (ge) {
      JAC_ISO(2, true);
    } else {
      JAC_ISO(2, false);
    }
  } else {
    if (ge) {
      JAC_ISO(3, true);
    } else {
      JAC_ISO(3, false);
    }
  }
#undef JAC_ISO
#undef JAC_LAUNCH
  return (int)hipGetLastError();
}

extern "C" {

int tacs_hip_solid_mat_type(int order, int nnodes, int nelem, const int* n2e_ptr,
                            const int* n2e, const int* conn, const double* X, const double* u,
                            const double* telem, const SolidMaterial* mat, double kc, double mc,
                            double gc, const int* rowp, const int* cols, const int* bcvars,
                            double* G, double* S, double* A, float* A32, int* status,
                            hipStream_t stream) {
  int v = tacs_tune_variant(TACS_TUNE_JACOBIAN, order, nnodes);
  if (v >= jac_rows_nvariants(order)) v = 0;  // the element-block form: rows default
  return solid_mat_type_impl(v, order, nnodes, nelem, n2e_ptr, n2e, conn, X, u, telem, mat, kc,
                             mc, gc, rowp, cols, bcvars, G, S, A, A32, status, stream);
}

int tacs_hip_solid_jacobian_nvariants(int order) { return jac_rows_nvariants(order); }

int tacs_hip_solid_jacobian_variant(int variant, int order, int nnodes, int nelem,
                                    const int* n2e_ptr, const int* n2e, const int* conn,
                                    const double* X, const double* telem,
                                    const SolidMaterial* mat, double alpha, double gamma,
                                    const int* rowp, const int* cols, const int* bcvars,
                                    double* G, double* A, float* A32, int* status,
                                    hipStream_t stream) {
  if (variant < 0 || variant >= jac_rows_nvariants(order)) return TACS_HIP_ERR_ARG;
  return solid_mat_type_impl(variant, order, nnodes, nelem, n2e_ptr, n2e, conn, X, NULL, telem,
                             mat, alpha, gamma, 0.0, rowp, cols, bcvars, G, NULL, A, A32, status,
                             stream);
}

int tacs_hip_solid_elem_blocks(int order, int prec, int nelem, const int* conn, const double* X,
                               const double* telem, const SolidMaterial* mat, double alpha,
                               double gamma, void* ke, hipStream_t stream) {
  if ((order != 2 && order != 3) || (prec != 64 && prec != 32) || !mat || !mat->isotropic)
    return TACS_HIP_ERR_ARG;
  if (nelem <= 0) return 0;
#define EB_LAUNCH(O, T)                                                                         \
  hipLaunchKernelGGL((solid_elem_blocks<O, T>),                                                 \
                     dim3((nelem + EBCfg<O>::EPB - 1) / EBCfg<O>::EPB),                         \
                     dim3(EBCfg<O>::TEAM * EBCfg<O>::EPB), 0, stream, nelem, conn, X, telem, *mat, \
                     alpha, gamma, (T*)ke)
  if (order == 2 && prec == 64)
    EB_LAUNCH(2, double);
  else if (order == 2)
    EB_LAUNCH(2, float);
  else if (prec == 64)
    EB_LAUNCH(3, double);
  else
    EB_LAUNCH(3, float);
#undef EB_LAUNCH
  return (int)hipGetLastError();
}

int tacs_hip_gather_solid_blocks(int order, int prec, int nrows, const int* rowp,
                                 const int* cols, const int* n2e_ptr, const int* n2e,
                                 const int* conn, const void* ke, const int* bcvars, double* A,
                                 float* A32, int* status, hipStream_t stream) {
  if ((order != 2 && order != 3) || (prec != 64 && prec != 32)) return TACS_HIP_ERR_ARG;
  if (nrows <= 0) return 0;
#define GS_LAUNCH(NPE, T)                                                                       \
  hipLaunchKernelGGL((gather_solid_rows<NPE, T>),                                               \
                     dim3(tacs_grid_for(nrows, GatherCfg<NPE>::WPB, 1 << 20)),                  \
                     dim3(64 * GatherCfg<NPE>::WPB), 0, stream, nrows, rowp, cols, n2e_ptr, n2e, \
                     conn, (const T*)ke, bcvars, A, A32, status)
  if (order == 2 && prec == 64)
    GS_LAUNCH(8, double);
  else if (order == 2)
    GS_LAUNCH(8, float);
  else if (prec == 64)
    GS_LAUNCH(27, double);
  else
    GS_LAUNCH(27, float);
#undef GS_LAUNCH
  return (int)hipGetLastError();
}

int tacs_hip_solid_jacobian(int order, int nnodes, int nelem, const int* n2e_ptr,
                            const int* n2e, const int* conn, const double* X,
                            const double* telem, const SolidMaterial* mat, double alpha,
                            double gamma, const int* rowp, const int* cols, const int* bcvars,
                            double* G, double* A, float* A32, int* status, hipStream_t stream) {
  return tacs_hip_solid_mat_type(order, nnodes, nelem, n2e_ptr, n2e, conn, X, NULL, telem, mat,
                                 alpha, gamma, 0.0, rowp, cols, bcvars, G, NULL, A, A32, status,
                                 stream);
}

int tacs_hip_solid_matfree_data(int order, int nelem, const int* conn, const double* X,
                                const double* telem, double kc, double mc, double* data,
                                hipStream_t stream) {
  if (order != 2 && order != 3) return TACS_HIP_ERR_ARG;
  hipLaunchKernelGGL(matfree_header, dim3(1), dim3(64), 0, stream, data, kc, mc);
  const size_t nq = (size_t)nelem * order * order * order;
  if (nq > 0) {
    if (order == 2)
      hipLaunchKernelGGL(solid_qp_geometry<2>, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0,
                         stream, nelem, conn, X, telem, data + 4);
    else
      hipLaunchKernelGGL(solid_qp_geometry<3>, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0,
                         stream, nelem, conn, X, telem, data + 4);
  }
  return (int)hipGetLastError();
}

int tacs_hip_solid_matfree_ev(int order, int nelem, const int* conn, const double* data,
                              const SolidMaterial* mat, const double* x, double* ev,
                              hipStream_t stream) {
  if (order != 2 && order != 3) return TACS_HIP_ERR_ARG;
  if (nelem <= 0) return 0;
  const int epb = 256 / (order == 2 ? MFCfg<2>::L : MFCfg<3>::L);
  const int grid = tacs_grid_for(nelem, epb, 1 << 20);
#define MF_LAUNCH(O, I)                                                                     \
  hipLaunchKernelGGL((solid_matfree_ev<O, I>), dim3(grid), dim3(256), 0, stream, nelem, conn, \
                     data, *mat, x, ev)
  if (order == 2 && mat->isotropic)
    MF_LAUNCH(2, true);
  else if (order == 2)
    MF_LAUNCH(2, false);
  else if (mat->isotropic)
    MF_LAUNCH(3, true);
  else
    MF_LAUNCH(3, false);
#undef MF_LAUNCH
  return (int)hipGetLastError();
}

// mode: 0 residual (coef = 1), 1 compliance dfdu (coef = alpha), 2 KS dfdu
int tacs_hip_solid_elem_vectors(int order, int mode, int nelem, const int* conn,
                                const double* X, const double* u, const double* telem,
                                const SolidMaterial* mat, double coef, double ks_weight,
                                double ks_max, double ks_sum, double safety, double* ev,
                                hipStream_t stream) {
  int grid = (nelem + 127) / 128;
  if (nelem <= 0) return 0;
#define EV_LAUNCH(O, M)                                                                       \
  hipLaunchKernelGGL((solid_elem_vectors<O, M>), dim3(grid), dim3(128), 0, stream, nelem, conn, \
                     X, u, telem, *mat, coef, ks_weight, ks_max, ks_sum, safety, ev)
  if (order == 2) {
    if (mode == 0) EV_LAUNCH(2, EV_RESIDUAL);
    else if (mode == 1) EV_LAUNCH(2, EV_COMPLIANCE_SV);
    else EV_LAUNCH(2, EV_KS_SV);
  } else if (order == 3) {
    if (mode == 0) EV_LAUNCH(3, EV_RESIDUAL);
    else if (mode == 1) EV_LAUNCH(3, EV_COMPLIANCE_SV);
    else EV_LAUNCH(3, EV_KS_SV);
  } else {
    return TACS_HIP_ERR_ARG;
  }
#undef EV_LAUNCH
  return (int)hipGetLastError();
}

int tacs_hip_solid_traction_vectors(int order, int naux, const int* aux_elem,
                                    const int* aux_face, const double* aux_trac,
                                    const double* aux_pres, const int* conn, const double* X,
                                    double lambda, double* ev, hipStream_t stream) {
  if (naux <= 0) return 0;
  int grid = (naux + 63) / 64;
  if (order == 2)
    hipLaunchKernelGGL(solid_traction_vectors<2>, dim3(grid), dim3(64), 0, stream, naux,
                       aux_elem, aux_face, aux_trac, aux_pres, conn, X, lambda, ev);
  else if (order == 3)
    hipLaunchKernelGGL(solid_traction_vectors<3>, dim3(grid), dim3(64), 0, stream, naux,
                       aux_elem, aux_face, aux_trac, aux_pres, conn, X, lambda, ev);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_solid_xpt_vectors(int order, int mode, int nelem, const int* conn, const double* X,
                               const double* u, const double* psi, const double* telem,
                               const SolidMaterial* mat, double scale, double ks_weight,
                               double ks_max, double ks_sum, double safety, double* out,
                               hipStream_t stream) {
  if (nelem <= 0) return 0;
  if ((order != 2 && order != 3) || mode < 0 || mode > 2 || (mode == 2 && !psi))
    return TACS_HIP_ERR_ARG;
  dim3 g((nelem + 127) / 128), tb(128);
#define XV_LAUNCH(O, M)                                                                          \
  hipLaunchKernelGGL((solid_xpt_vectors<O, M>), g, tb, 0, stream, nelem, conn, X, u, psi, telem, \
                     *mat, scale, ks_weight, ks_max, ks_sum, safety, out)
  if (order == 2) {
    if (mode == 0) XV_LAUNCH(2, XV_COMPLIANCE);
    else if (mode == 1) XV_LAUNCH(2, XV_KS);
    else XV_LAUNCH(2, XV_ADJ);
  } else {
    if (mode == 0) XV_LAUNCH(3, XV_COMPLIANCE);
    else if (mode == 1) XV_LAUNCH(3, XV_KS);
    else XV_LAUNCH(3, XV_ADJ);
  }
#undef XV_LAUNCH
  return (int)hipGetLastError();
}

int tacs_hip_solid_traction_xpt_vectors(int order, int naux, const int* aux_elem,
                                        const int* aux_face, const double* aux_trac,
                                        const int* conn, const double* X, const double* psi,
                                        double scale, double* out, hipStream_t stream) {
  if (naux <= 0) return 0;
  dim3 g((naux + 127) / 128), tb(128);
  if (order == 2)
    hipLaunchKernelGGL(solid_traction_xpt_vectors<2>, g, tb, 0, stream, naux, aux_elem, aux_face,
                       aux_trac, conn, X, psi, scale, out);
  else if (order == 3)
    hipLaunchKernelGGL(solid_traction_xpt_vectors<3>, g, tb, 0, stream, naux, aux_elem, aux_face,
                       aux_trac, conn, X, psi, scale, out);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_gather_elem_vectors(int nv, int nnodes, const int* n2e_ptr, const int* n2e,
                                 const double* ev, int accumulate, double* out,
                                 hipStream_t stream) {
  int grid = (nnodes + 255) / 256;
  if (nnodes <= 0) return 0;
  if (nv == 24)
    hipLaunchKernelGGL(gather_elem_vectors<24>, dim3(grid), dim3(256), 0, stream, nnodes,
                       n2e_ptr, n2e, ev, accumulate, out);
  else if (nv == 81)
    hipLaunchKernelGGL(gather_elem_vectors<81>, dim3(grid), dim3(256), 0, stream, nnodes,
                       n2e_ptr, n2e, ev, accumulate, out);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

// mode: 0 compliance integrand, 1 KS max, 2 KS sum, 3 compliance dC/dt (e.Ce),
// 4 adjoint-residual product
int tacs_hip_solid_elem_scalars(int order, int mode, int nelem, const int* conn,
                                const double* X, const double* u, const double* psi,
                                const double* telem, const SolidMaterial* mat, double scale,
                                double ks_weight, double ks_max, double safety, double* out,
                                hipStream_t stream) {
  if (nelem <= 0) return 0;
  int grid = (nelem + 127) / 128;
#define ES_LAUNCH(O, M)                                                                        \
  hipLaunchKernelGGL((solid_elem_scalars<O, M>), dim3(grid), dim3(128), 0, stream, nelem, conn, \
                     X, u, psi, telem, *mat, scale, ks_weight, ks_max, safety, out)
  if (order == 2) {
    switch (mode) {
      case 0: ES_LAUNCH(2, ES_COMPLIANCE); break;
      case 1: ES_LAUNCH(2, ES_KS_MAX); break;
      case 2: ES_LAUNCH(2, ES_KS_SUM); break;
      case 3: ES_LAUNCH(2, ES_DV_COMPLIANCE); break;
      case 4: ES_LAUNCH(2, ES_ADJ_PRODUCT); break;
      default: return TACS_HIP_ERR_ARG;
    }
  } else if (order == 3) {
    switch (mode) {
      case 0: ES_LAUNCH(3, ES_COMPLIANCE); break;
      case 1: ES_LAUNCH(3, ES_KS_MAX); break;
      case 2: ES_LAUNCH(3, ES_KS_SUM); break;
      case 3: ES_LAUNCH(3, ES_DV_COMPLIANCE); break;
      case 4: ES_LAUNCH(3, ES_ADJ_PRODUCT); break;
      default: return TACS_HIP_ERR_ARG;
    }
  } else {
    return TACS_HIP_ERR_ARG;
  }
#undef ES_LAUNCH
  return (int)hipGetLastError();
}

int tacs_hip_gather_design_vars(int nelem, const int* dvnum, const double* tfixed,
                                const double* dv, double* telem, hipStream_t stream) {
  if (nelem <= 0) return 0;
  hipLaunchKernelGGL(gather_design_vars, dim3((nelem + 255) / 256), dim3(256), 0, stream, nelem,
                     dvnum, tfixed, dv, telem);
  return (int)hipGetLastError();
}

int tacs_hip_scatter_dv_sens(int ndv, const int* dv2e_ptr, const int* dv2e, const double* es,
                             double* dfdx, hipStream_t stream) {
  if (ndv <= 0) return 0;
  hipLaunchKernelGGL(scatter_dv_sens, dim3((ndv + 255) / 256), dim3(256), 0, stream, ndv,
                     dv2e_ptr, dv2e, es, dfdx);
  return (int)hipGetLastError();
}

int tacs_hip_apply_bc_vec(int b, int nbc, const int* nodes, const int* masks, const double* vals,
                          const double* u, double lambda, double* x, hipStream_t stream) {
  if (nbc <= 0) return 0;
  hipLaunchKernelGGL(apply_bc_vec, dim3((nbc + 255) / 256), dim3(256), 0, stream, b, nbc, nodes,
                     masks, vals, u, lambda, x);
  return (int)hipGetLastError();
}

}  // extern "C"
