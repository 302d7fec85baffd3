// Linear MITC4 shell (TACSQuad4Shell: TACSShellElement<TACSQuadLinearQuadrature,
// TACSShellQuadBasis<2>, TACSLinearizedRotation, TACSShellLinearModel>,
// src/elements/shell/TACSShellElementDefs.h:18-20) with the natural transform
// and TACSIsoShellConstitutive, for gfx950.
//
// Work layout.  One 32-lane team per element, four elements per 128-thread
// workgroup.  The team stages the element geometry in LDS (node normals and
// drill frames, tying-point frames, quadrature-point transforms), then lane c
// (c < 24: node c/6, dof c%6) builds column c of the linear strain map
// B_q (9 x 24 per quadrature point) -- the strain of the unit vector e_c.  The
// map is the reference's forward strain computation (TACSShellElement::
// addResidual, TACSShellElement.h:295-420) specialised to one nonzero
// variable, which makes every term a rank-1 update.  The epilogue depends on
// MODE: the element stiffness K = sum_q detXd_q B_q^T Cs B_q (lane c owns
// column c), the residual B^T Cs B u, or the per-element function values and
// sensitivities.  Element matrices go to an element-major buffer that
// gather_elem_matrices sums into block-CSR rows in ascending element order
// (deterministic, no atomics, no zero fill).
#include "tacs_hip_common.h"

namespace {

constexpr int SH_TEAM = 32;
constexpr int SH_EPB = 4;  // elements (teams) per workgroup
constexpr int SH_NV = 24;

// Forward-mode dual number: value and derivative along one direction.  The
// element routines instantiated with it give the exact directional
// derivatives in the node coordinates, which the reference's shell element
// approximates by differencing its residual and point quantities in each
// coordinate (TACSElement::addAdjResXptProduct / addPointQuantityXptSens,
// src/elements/TACSElement.cpp:270, :746; exact to rounding in its complex
// build).
struct ShDual {
  double v, d;
  ShDual() = default;
  __host__ __device__ constexpr ShDual(double a) : v(a), d(0.0) {}
  __host__ __device__ constexpr ShDual(double a, double b) : v(a), d(b) {}
  __device__ explicit operator double() const { return v; }
  // hidden friends: found by argument-dependent lookup only, so they never
  // capture float or double arguments
  friend __device__ ShDual operator+(ShDual a, ShDual b) { return ShDual(a.v + b.v, a.d + b.d); }
  friend __device__ ShDual operator-(ShDual a, ShDual b) { return ShDual(a.v - b.v, a.d - b.d); }
  friend __device__ ShDual operator-(ShDual a) { return ShDual(-a.v, -a.d); }
  friend __device__ ShDual operator*(ShDual a, ShDual b) {
    return ShDual(a.v * b.v, a.d * b.v + a.v * b.d);
  }
  friend __device__ ShDual operator/(ShDual a, ShDual b) {
    const double q = a.v / b.v;
    return ShDual(q, (a.d - q * b.d) / b.v);
  }
  friend __device__ ShDual& operator+=(ShDual& a, ShDual b) { return a = a + b; }
  friend __device__ ShDual& operator-=(ShDual& a, ShDual b) { return a = a - b; }
  friend __device__ ShDual& operator*=(ShDual& a, ShDual b) { return a = a * b; }
  friend __device__ ShDual& operator/=(ShDual& a, ShDual b) { return a = a / b; }
  friend __device__ bool operator<(ShDual a, ShDual b) { return a.v < b.v; }
  friend __device__ bool operator>(ShDual a, ShDual b) { return a.v > b.v; }
  friend __device__ bool operator==(ShDual a, ShDual b) { return a.v == b.v; }
  friend __device__ bool operator!=(ShDual a, ShDual b) { return a.v != b.v; }
  friend __device__ ShDual sqrt(ShDual a) {
    const double r = ::sqrt(a.v);
    return ShDual(r, r > 0.0 ? 0.5 * a.d / r : 0.0);
  }
  friend __device__ ShDual exp(ShDual a) {
    const double e = ::exp(a.v);
    return ShDual(e, e * a.d);
  }
  friend __device__ ShDual log(ShDual a) { return ShDual(::log(a.v), a.d / a.v); }
};
// node coordinate x, seeded with derivative 1 when it is the direction
template <typename T>
__device__ __forceinline__ T sh_seed(double x, bool dir) {
  return T(x);
}
template <>
__device__ __forceinline__ ShDual sh_seed<ShDual>(double x, bool dir) {
  return ShDual(x, dir ? 1.0 : 0.0);
}
__device__ __forceinline__ double sh_dpart(double) { return 0.0; }
__device__ __forceinline__ double sh_dpart(float) { return 0.0; }
__device__ __forceinline__ double sh_dpart(ShDual a) { return a.d; }

template <typename T>
struct ShellLDS {
  T X[12];
  T nXxi[4][6];   // X,xi at the nodes
  T fn[4][3];     // unit node normals
  T Tn[4][9];     // node transforms
  T XTn[4][9];    // Xdn^-1 Tn
  T tXxi[9][6];   // X,xi at the tying points
  T tn0[9][3];    // interpolated normal at the tying points
  T qT[4][9];     // T at the quadrature points
  T qXT[4][9];    // Xd^-1 T
  T qXzT[4][9];   // -Xd^-1 Xdz Xd^-1 T
  T detw[4];      // w * det(Xd)
  T B[4][9][SH_NV];
  T u[SH_NV], psi[SH_NV];
  T e[4][9], ep[4][9], de[4][9];
  T val[4];
  T red[SH_NV];
};

template <typename T>
struct ShellMat {
  T Q0, Q1, Q5;  // plane-stress Q (Q2 = Q4 = 0, Q3 = Q0)
  T kcorr, drill_reg, ys;
  T toff;  // tOffset: the mid-surface sits t (1/2 - toff) below the top face
};

// Entry (p, q) of the cross-product matrix a^x (a^x v = a x v)
template <typename T>
__device__ __forceinline__ T sh_skew(const T* a, int p, int q) {
  if (p == q) return T(0);
  const int r = 3 - p - q;  // the third index
  const T v = a[r];
  // (p, q) in cyclic order (0,1), (1,2), (2,0) -> -a_r, otherwise +a_r
  return ((p + 1) % 3 == q) ? -v : v;
}

// Gauss points of TacsGaussQuadPts2 (src/elements/basis/TACSGaussQuadrature.h:26)
__device__ __forceinline__ double sh_gauss(int i) { return i ? 0.577350269189626 : -0.577350269189626; }

// Tying points in storage order G11(2), G22(2), G12(1), G23(2), G13(2)
// (TACSShellElementQuadBasis.h:487-560).
__device__ __forceinline__ void sh_tying_pt(int t, double& x, double& y) {
  x = (t == 2 || t == 5) ? -1.0 : ((t == 3 || t == 6) ? 1.0 : 0.0);
  y = (t == 0 || t == 7) ? -1.0 : ((t == 1 || t == 8) ? 1.0 : 0.0);
}

// Shape function of node a = i + 2j and its parametric gradient at (x, y)
template <typename T>
__device__ __forceinline__ void sh_shape(int a, T x, T y, T& N, T& Nx, T& Ny) {
  T na = (a & 1) ? T(0.5) * (T(1) + x) : T(0.5) * (T(1) - x);
  T nb = (a & 2) ? T(0.5) * (T(1) + y) : T(0.5) * (T(1) - y);
  T da = (a & 1) ? T(0.5) : T(-0.5);
  T db = (a & 2) ? T(0.5) : T(-0.5);
  N = na * nb;
  Nx = da * nb;
  Ny = na * db;
}

// X,xi of a 3-vector field at (x, y): g[2k] = d/dxi, g[2k+1] = d/deta
template <typename T>
__device__ __forceinline__ void sh_grad3(const T* v, T x, T y, T g[6]) {
#pragma unroll
  for (int k = 0; k < 6; k++) g[k] = T(0);
#pragma unroll
  for (int a = 0; a < 4; a++) {
    T N, Nx, Ny;
    sh_shape<T>(a, x, y, N, Nx, Ny);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      g[2 * k] += Nx * v[3 * a + k];
      g[2 * k + 1] += Ny * v[3 * a + k];
    }
  }
}

template <typename T>
__device__ __forceinline__ void sh_interp3(const T* v, T x, T y, T f[3]) {
  f[0] = f[1] = f[2] = T(0);
#pragma unroll
  for (int a = 0; a < 4; a++) {
    T N, Nx, Ny;
    sh_shape<T>(a, x, y, N, Nx, Ny);
#pragma unroll
    for (int k = 0; k < 3; k++) f[k] += N * v[3 * a + k];
  }
}

template <typename T>
__device__ __forceinline__ void sh_cross(const T a[3], const T b[3], T c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
__device__ __forceinline__ T sh_inv3(const T A[9], T B[9]) {
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

template <typename T>
__device__ __forceinline__ void sh_mm(const T A[9], const T B[9], T C[9]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// TACSShellNaturalTransform::computeTransform (TACSShellElementTransform.h:
// 25-92), including its removal of the normal component from t1[0] only,
// three times over.  Columns of T: t1, t2 = n x t1, n.
// Element frame T = [t1 t2 n].  Natural transform
// (TACSShellNaturalTransform::computeTransform, TACSShellElementTransform.h:
// 25-60): t1 from X,xi with the reference's projection step, which subtracts
// d n[0] from t1[0] three times and leaves t1[1], t1[2] as they are, so t1
// is not orthogonal to n on a warped element; kept as it is, since it
// decides the element matrices.  Reference-axis transform
// (TACSShellRefAxisTransform::computeTransform :116-160): t1 = the unit
// axis minus its normal component, normalised.  t2 = n x t1.
template <typename T>
__device__ __forceinline__ void sh_transform(const T Xxi[6], const T n0[3], T Tm[9],
                                             int ref_axis = 0, const double* axis = nullptr) {
  T inv = T(1) / sqrt(n0[0] * n0[0] + n0[1] * n0[1] + n0[2] * n0[2]);
  T n[3] = {n0[0] * inv, n0[1] * inv, n0[2] * inv};
  T t1[3];
  if (ref_axis) {
    const T a[3] = {T(axis[0]), T(axis[1]), T(axis[2])};
    T an = a[0] * n[0] + a[1] * n[1] + a[2] * n[2];
    t1[0] = a[0] - an * n[0];
    t1[1] = a[1] - an * n[1];
    t1[2] = a[2] - an * n[2];
  } else {
    t1[0] = Xxi[0];
    t1[1] = Xxi[2];
    t1[2] = Xxi[4];
    T d = n[0] * t1[0] + n[1] * t1[1] + n[2] * t1[2];
    t1[0] = t1[0] - d * n[0];
    t1[0] = t1[0] - d * n[0];
    t1[0] = t1[0] - d * n[0];
  }
  inv = T(1) / sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
  t1[0] *= inv;
  t1[1] *= inv;
  t1[2] *= inv;
  T t2[3];
  sh_cross<T>(n, t1, t2);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    Tm[3 * k] = t1[k];
    Tm[3 * k + 1] = t2[k];
    Tm[3 * k + 2] = n[k];
  }
}

// Stress of TACSIsoShellConstitutive (TACSIsoShellConstitutive.cpp:153-189 with
// tOffset = 0; computeStress TACSShellConstitutive.h:133)
template <typename T>
__device__ __forceinline__ void sh_stress(const ShellMat<T>& m, T t, const T e[9], T s[9]) {
  // A = t Q, B = -toff t^2 Q, D = t^3 (1/12 + toff^2) Q
  T I = t * t * t / T(12);
  T As = m.kcorr * t * m.Q5;
  T a0 = m.Q0 * e[0] + m.Q1 * e[1], a1 = m.Q1 * e[0] + m.Q0 * e[1], a2 = m.Q5 * e[2];
  T b0 = m.Q0 * e[3] + m.Q1 * e[4], b1 = m.Q1 * e[3] + m.Q0 * e[4], b2 = m.Q5 * e[5];
  s[0] = t * a0;
  s[1] = t * a1;
  s[2] = t * a2;
  s[3] = I * b0;
  s[4] = I * b1;
  s[5] = I * b2;
  if (m.toff != T(0)) {
    const T Bc = -m.toff * t * t, Dc = m.toff * m.toff * t * t * t;
    s[0] += Bc * b0;
    s[1] += Bc * b1;
    s[2] += Bc * b2;
    s[3] += Dc * b0 + Bc * a0;
    s[4] += Dc * b1 + Bc * a1;
    s[5] += Dc * b2 + Bc * a2;
  }
  s[6] = As * e[6];
  s[7] = As * e[7];
  s[8] = T(0.5) * m.drill_reg * (As + As) * e[8];
}

// von Mises of a plane-stress state and its gradient (vonMisesFailure2D and
// vonMisesFailure2DStressSens, TACSMaterialProperties.cpp:481-500)
template <typename T>
__device__ __forceinline__ T sh_vm2(const T s[3], T ys, T sens[3]) {
  T f = sqrt(s[0] * s[0] + s[1] * s[1] - s[0] * s[1] + T(3) * s[2] * s[2]);
  if (f != T(0)) {
    sens[0] = (s[0] - T(0.5) * s[1]) / (f * ys);
    sens[1] = (s[1] - T(0.5) * s[0]) / (f * ys);
    sens[2] = T(3) * s[2] / (f * ys);
  } else {
    sens[0] = sens[1] = sens[2] = T(0);
  }
  return f / ys;
}

// Failure index (KS of top/bottom von Mises, weight 100;
// TACSIsoShellConstitutive::evalFailure :254), strain gradient
// (evalFailureStrainSens :300) and thickness derivative (addFailureDVSens :381)
template <typename T>
__device__ __forceinline__ T sh_failure(const ShellMat<T>& m, T t, const T e[9], T de[9], T& dt) {
  const T ksw = T(100);
  T ht = (T(0.5) - m.toff) * t, hb = -(T(0.5) + m.toff) * t;
  T et[3], eb[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    et[k] = e[k] + ht * e[3 + k];
    eb[k] = e[k] + hb * e[3 + k];
  }
  T st[3] = {m.Q0 * et[0] + m.Q1 * et[1], m.Q1 * et[0] + m.Q0 * et[1], m.Q5 * et[2]};
  T sb[3] = {m.Q0 * eb[0] + m.Q1 * eb[1], m.Q1 * eb[0] + m.Q0 * eb[1], m.Q5 * eb[2]};
  T pt[3], pb[3];
  T top = sh_vm2<T>(st, m.ys, pt), bot = sh_vm2<T>(sb, m.ys, pb);
  T mx = top > bot ? top : bot;
  T sum = exp(ksw * (top - mx)) + exp(ksw * (bot - mx));
  T ft = exp(ksw * (top - mx)) / sum, fb = exp(ksw * (bot - mx)) / sum;
  T phit[3] = {m.Q0 * pt[0] + m.Q1 * pt[1], m.Q1 * pt[0] + m.Q0 * pt[1], m.Q5 * pt[2]};
  T phib[3] = {m.Q0 * pb[0] + m.Q1 * pb[1], m.Q1 * pb[0] + m.Q0 * pb[1], m.Q5 * pb[2]};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    de[k] = ft * phit[k] + fb * phib[k];
    de[3 + k] = ft * ht * phit[k] + fb * hb * phib[k];
  }
  de[6] = de[7] = de[8] = T(0);
  dt = ft * (T(0.5) - m.toff) * (phit[0] * e[3] + phit[1] * e[4] + phit[2] * e[5]) -
       fb * (T(0.5) + m.toff) * (phib[0] * e[3] + phib[1] * e[4] + phib[2] * e[5]);
  return mx + log(sum) / ksw;
}

}  // namespace

enum {
  SH_MATRIX = 0,
  SH_RESIDUAL = 1,
  SH_COMPLIANCE = 2,
  SH_KS_MAX = 3,
  SH_KS_SUM = 4,
  SH_KS_SV = 5,
  SH_KS_DV = 6,
  SH_ADJ_PRODUCT = 7,
  SH_COMPLIANCE_SV = 8,
  SH_COMPLIANCE_DV = 9,
  // node-coordinate derivatives (T = ShDual, one team per element and
  // coordinate direction; es[12 el + k] receives d/dX_k):
  SH_XPT_ADJ = 10,         // scale psi^T R
  SH_XPT_COMPLIANCE = 11,  // scale sum_q detXd_q e.s
  SH_XPT_KS = 12,          // scale sum_q detXd_q exp(w (f_q - max)) / (w sum)
  SH_GEO = 13,             // ke += scale G (geometric stiffness along the state u)
};

// Work arrays of the geometric stiffness (SH_GEO), per element
struct ShellGeoLDS {
  double sp[4][9];      // stress of the linear strain of the path at the points
  double W[9];          // tying-point weights of sp (see SH_GEO)
  double PU[9][6];      // path: dU/dxi at the tying points
  double PD[9][3];      // path: director at the tying points
  double PU0[4][9];     // path: u0x at the points
  double PU1[4][9];     // path: u1x at the points
  double B1[4][9][24];  // 2 Q(path, e_c): the path derivative of the strain map
  double PW[4][24][9];  // unit column c at point q: p (3), w0 (3), w1 (3)
};

struct ShellParams {
  double Q0, Q1, Q5, kcorr, drill_reg, ys;
  double scale, ks_weight, ks_max, ks_sum, safety;
  double mass, rho;  // coefficient of the consistent mass (matrix mode), density
  double axis[3];    // reference axis of TACSShellRefAxisTransform (ref_axis = 1)
  int ref_axis;
  double toff;
  const ShellMaterial* mtab;  // per-element materials (device), or NULL
  const int* emat;
};

template <typename T, int MODE>
__global__ __launch_bounds__(SH_TEAM* SH_EPB) void shell_elem_kernel(
    int nelem, const int* __restrict__ conn, const double* __restrict__ Xg,
    const double* __restrict__ ug, const double* __restrict__ psig,
    const double* __restrict__ telem, ShellParams P, T* __restrict__ ke,
    double* __restrict__ ev, double* __restrict__ es) {
  __shared__ ShellLDS<T> lds_all[SH_EPB];
  const int team = threadIdx.x / SH_TEAM, lane = threadIdx.x % SH_TEAM;
  constexpr bool XPT = MODE == SH_XPT_ADJ || MODE == SH_XPT_COMPLIANCE || MODE == SH_XPT_KS;
  const int gt = blockIdx.x * SH_EPB + team;
  const int el = XPT ? gt / 12 : gt;
  const int dir = XPT ? gt % 12 : -1;
  const bool active = el < nelem;
  ShellLDS<T>& L = lds_all[team];
  ShellMat<T> M = {T(P.Q0), T(P.Q1), T(P.Q5), T(P.kcorr), T(P.drill_reg), T(P.ys), T(P.toff)};
  double rho = P.rho;
  if (P.emat && active) {
    // the element's own material (per-element constitutive objects)
    const ShellMaterial& me = P.mtab[P.emat[el]];
    M = {T(me.Q[0]), T(me.Q[1]), T(me.Q[5]), T(me.kcorr), T(me.drill_reg), T(me.ys), T(me.toff)};
    rho = me.rho;
  }
  const T t = active ? T(telem ? telem[el] : 1.0) : T(1);

  // ---- load the element's nodes and variables
  if (active) {
    const int* c = &conn[(size_t)el * 4];
    if (lane < 12) L.X[lane] = sh_seed<T>(Xg[3 * (size_t)c[lane / 3] + lane % 3], lane == dir);
    if (lane < 24) {
      L.u[lane] = ug ? T(ug[6 * (size_t)c[lane / 6] + lane % 6]) : T(0);
      if (MODE == SH_ADJ_PRODUCT || MODE == SH_XPT_ADJ)
        L.psi[lane] = T(psig[6 * (size_t)c[lane / 6] + lane % 6]);
    }
  }
  __syncthreads();

  // ---- node normals (TacsShellComputeNodeNormals, TACSShellUtilities.h:301)
  if (active && lane < 4) {
    T x = (lane & 1) ? T(1) : T(-1), y = (lane & 2) ? T(1) : T(-1);
    T g[6];
    sh_grad3<T>(L.X, x, y, g);
    T a[3] = {g[0], g[2], g[4]}, b[3] = {g[1], g[3], g[5]}, f[3];
    sh_cross<T>(a, b, f);
    T nrm = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (nrm != T(0)) {
      T r = T(1) / nrm;
      f[0] *= r;
      f[1] *= r;
      f[2] *= r;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) L.nXxi[lane][k] = g[k];
#pragma unroll
    for (int k = 0; k < 3; k++) L.fn[lane][k] = f[k];
  }
  __syncthreads();

  // ---- frames: node drill frames, tying points, quadrature points
  if (active && lane < 4) {
    T Tm[9], Xd[9], Xinv[9];
    sh_transform<T>(L.nXxi[lane], L.fn[lane], Tm, P.ref_axis, P.axis);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      Xd[3 * k] = L.nXxi[lane][2 * k];
      Xd[3 * k + 1] = L.nXxi[lane][2 * k + 1];
      Xd[3 * k + 2] = L.fn[lane][k];
    }
    sh_inv3<T>(Xd, Xinv);
    sh_mm<T>(Xinv, Tm, L.XTn[lane]);
#pragma unroll
    for (int k = 0; k < 9; k++) L.Tn[lane][k] = Tm[k];
  } else if (active && lane >= 4 && lane < 13) {
    int ty = lane - 4;
    double xd, yd;
    sh_tying_pt(ty, xd, yd);
    T g[6], n0[3];
    sh_grad3<T>(L.X, T(xd), T(yd), g);
    sh_interp3<T>(&L.fn[0][0], T(xd), T(yd), n0);
#pragma unroll
    for (int k = 0; k < 6; k++) L.tXxi[ty][k] = g[k];
#pragma unroll
    for (int k = 0; k < 3; k++) L.tn0[ty][k] = n0[k];
  } else if (active && lane >= 13 && lane < 17) {
    int q = lane - 13;
    T x = T(sh_gauss(q & 1)), y = T(sh_gauss(q >> 1));
    T g[6], n0[3], nxi[6], Tm[9], Xd[9], Xdz[9], Xinv[9], XT[9], tmp[9];
    sh_grad3<T>(L.X, x, y, g);
    sh_interp3<T>(&L.fn[0][0], x, y, n0);
    sh_grad3<T>(&L.fn[0][0], x, y, nxi);
    sh_transform<T>(g, n0, Tm, P.ref_axis, P.axis);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      Xd[3 * k] = g[2 * k];
      Xd[3 * k + 1] = g[2 * k + 1];
      Xd[3 * k + 2] = n0[k];
      Xdz[3 * k] = nxi[2 * k];
      Xdz[3 * k + 1] = nxi[2 * k + 1];
      Xdz[3 * k + 2] = T(0);
    }
    T det = sh_inv3<T>(Xd, Xinv);
    sh_mm<T>(Xinv, Tm, XT);
    sh_mm<T>(Xinv, Xdz, tmp);
#pragma unroll
    for (int k = 0; k < 9; k++) tmp[k] = -tmp[k];
    sh_mm<T>(tmp, XT, L.qXzT[q]);
#pragma unroll
    for (int k = 0; k < 9; k++) {
      L.qT[q][k] = Tm[k];
      L.qXT[q][k] = XT[k];
    }
    L.detw[q] = det;  // quadrature weights are 1
  }
  __syncthreads();

  // ---- column c of B_q: the strains of the unit vector e_c
  if (active && lane < SH_NV) {
    const int a = lane / 6, k = lane % 6;
    T ety[9], etn[4];
    if (k < 3) {
      // translation u_k of node a: Uxi[2k] = N_a,xi, Uxi[2k+1] = N_a,eta, d = 0
#pragma unroll
      for (int ty = 0; ty < 9; ty++) {
        double xd, yd;
        sh_tying_pt(ty, xd, yd);
        T N, Nx, Ny;
        sh_shape<T>(a, T(xd), T(yd), N, Nx, Ny);
        const T* g = L.tXxi[ty];
        if (ty < 2) ety[ty] = Nx * g[2 * k];
        else if (ty < 4) ety[ty] = Ny * g[2 * k + 1];
        else if (ty < 5) ety[ty] = T(0.5) * (Nx * g[2 * k + 1] + Ny * g[2 * k]);
        else if (ty < 7) ety[ty] = T(0.5) * (L.tn0[ty][k] * Ny);
        else ety[ty] = T(0.5) * (L.tn0[ty][k] * Nx);
      }
      // drill strain at node i: u0x = Tn^T [u0xi | 0] Xdn^-1 Tn (rank 1)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        T N, Nx, Ny;
        sh_shape<T>(a, (i & 1) ? T(1) : T(-1), (i & 2) ? T(1) : T(-1), N, Nx, Ny);
        const T* X = L.XTn[i];
        T m0 = Nx * X[0] + Ny * X[3];  // row k of [u0xi|0] Xdn^-1 Tn, column 0
        T m1 = Nx * X[1] + Ny * X[4];  // column 1
        // u0x[1] = Tn[k][0] m1, u0x[3] = Tn[k][1] m0
        etn[i] = T(0.5) * (L.Tn[i][3 * k + 1] * m0 - L.Tn[i][3 * k] * m1);
      }
    } else {
      // rotation q_r of node a: d_a = e_r x fn_a, Uxi = 0
      const int r = k - 3;
      const T* f = L.fn[a];
      T da[3];
      da[0] = (r == 0) ? T(0) : ((r == 1) ? f[2] : -f[1]);
      da[1] = (r == 0) ? -f[2] : ((r == 1) ? T(0) : f[0]);
      da[2] = (r == 0) ? f[1] : ((r == 1) ? -f[0] : T(0));
#pragma unroll
      for (int ty = 0; ty < 9; ty++) {
        double xd, yd;
        sh_tying_pt(ty, xd, yd);
        T N, Nx, Ny;
        sh_shape<T>(a, T(xd), T(yd), N, Nx, Ny);
        const T* g = L.tXxi[ty];
        if (ty < 5) ety[ty] = T(0);
        else if (ty < 7) ety[ty] = T(0.5) * (N * (g[1] * da[0] + g[3] * da[1] + g[5] * da[2]));
        else ety[ty] = T(0.5) * (N * (g[0] * da[0] + g[2] * da[1] + g[4] * da[2]));
      }
      // drill strain at node a: Ct - I = Tn^T (C(q) - I) Tn with C = I - q^x
      // (TACSLinearizedRotation::computeRotationMat, TACSDirector.h:25-34)
      T Cm[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
      if (r == 0) { Cm[5] = T(1); Cm[7] = T(-1); }
      if (r == 1) { Cm[2] = T(-1); Cm[6] = T(1); }
      if (r == 2) { Cm[1] = T(1); Cm[3] = T(-1); }
#pragma unroll
      for (int i = 0; i < 4; i++) etn[i] = T(0);
      const T* Tn = L.Tn[a];
      T c10 = T(0), c01 = T(0);
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int s = 0; s < 3; s++) {
          c10 += Tn[3 * p + 1] * Cm[3 * p + s] * Tn[3 * s];
          c01 += Tn[3 * p] * Cm[3 * p + s] * Tn[3 * s + 1];
        }
      etn[a] = T(0.5) * (c10 - c01);
    }

    // quadrature points
#pragma unroll
    for (int q = 0; q < 4; q++) {
      T x = T(sh_gauss(q & 1)), y = T(sh_gauss(q >> 1));
      T N, Nx, Ny;
      sh_shape<T>(a, x, y, N, Nx, Ny);
      const T* Tm = L.qT[q];
      const T* XT = L.qXT[q];
      const T* XzT = L.qXzT[q];
      // u1x = p w^T (rank 1)
      T p[3], w[3];
      if (k < 3) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
          p[j] = Tm[3 * k + j];
          w[j] = Nx * XzT[j] + Ny * XzT[3 + j];
        }
      } else {
        const int r = k - 3;
        const T* f = L.fn[a];
        T da[3];
        da[0] = (r == 0) ? T(0) : ((r == 1) ? f[2] : -f[1]);
        da[1] = (r == 0) ? -f[2] : ((r == 1) ? T(0) : f[0]);
        da[2] = (r == 0) ? f[1] : ((r == 1) ? -f[0] : T(0));
#pragma unroll
        for (int j = 0; j < 3; j++) {
          p[j] = Tm[j] * da[0] + Tm[3 + j] * da[1] + Tm[6 + j] * da[2];
          w[j] = Nx * XT[j] + Ny * XT[3 + j] + N * XzT[6 + j];
        }
      }
      // tying strain interpolation (interpTyingStrain, :651-675)
      T nax0 = T(0.5) * (T(1) - x), nax1 = T(0.5) * (T(1) + x);
      T nby0 = T(0.5) * (T(1) - y), nby1 = T(0.5) * (T(1) + y);
      T g11 = nby0 * ety[0] + nby1 * ety[1];
      T g22 = nax0 * ety[2] + nax1 * ety[3];
      T g12 = ety[4];
      T g23 = nax0 * ety[5] + nax1 * ety[6];
      T g13 = nby0 * ety[7] + nby1 * ety[8];
      // e0ty = XT^T G XT (mat3x3SymmTransformTranspose, TACSElementAlgebra.h:1094)
      T G[9] = {g11, g12, g13, g12, g22, g23, g13, g23, T(0)};
      T W[9], E0[9];
      sh_mm<T>(G, XT, W);
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
          E0[3 * i + j] = XT[i] * W[j] + XT[3 + i] * W[3 + j] + XT[6 + i] * W[6 + j];
      T et = T(0);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        T Ni, Nxi, Nyi;
        sh_shape<T>(i, x, y, Ni, Nxi, Nyi);
        et += Ni * etn[i];
      }
      L.B[q][0][lane] = E0[0];
      L.B[q][1][lane] = E0[4];
      L.B[q][2][lane] = T(2) * E0[1];
      L.B[q][3][lane] = p[0] * w[0];
      L.B[q][4][lane] = p[1] * w[1];
      L.B[q][5][lane] = p[0] * w[1] + p[1] * w[0];
      L.B[q][6][lane] = T(2) * E0[5];
      L.B[q][7][lane] = T(2) * E0[2];
      L.B[q][8][lane] = et;
    }
  }
  __syncthreads();

  if constexpr (MODE == SH_GEO) {
    // Geometric stiffness of TACSShellElement::getMatType(TACS_GEOMETRIC_
    // STIFFNESS_MATRIX) (TACSShellElement.h:661-908): the derivative, along
    // the path p = the element state, of the tangent stiffness of
    // TACSShellNonlinearModel at the zero state.  The nonlinear strain is
    // e(u) = L u + Q(u, u) (L: the linear map above, B[q][k][c] = (L e_c)_k;
    // Q: the quadratic terms of computeTyingStrain and evalStrain,
    // TACSShellElementModel.h:1088-1140, :1996-2012), so with s_p = C L p
    //   G[r][c] = sum_q detw_q (B1[:, r] . C B[:, c] + B[:, r] . C B1[:, c]
    //                           + s_p . 2 Q(e_r, e_c)),   B1[:, c] = 2 Q(p, e_c).
    // The tying part of s_p . Q(e_r, e_c) is sum_t W_t g_t(e_r, e_c) over the
    // nine tying-point bilinear values: the interpolation and the transform
    // E0 = XT^T G XT are linear, so s_p contracts with them once per element
    // (M_q = XT S XT^T, S the symmetric matrix of s_p's in-plane and shear
    // parts).
    __shared__ ShellGeoLDS geo_all[SH_EPB];
    ShellGeoLDS& Gs = geo_all[team];
    // (1) s_p at the points (lanes 0-3) and the path's tying-point fields
    //     (lanes 4-12) and point gradients (lanes 13-16)
    if (active && lane < 4) {
      const int q = lane;
      T ep[9];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        T acc = T(0);
        for (int c = 0; c < SH_NV; c++) acc += L.B[q][k][c] * L.u[c];
        ep[k] = acc;
      }
      T sq[9];
      sh_stress<T>(M, t, ep, sq);
#pragma unroll
      for (int k = 0; k < 9; k++) Gs.sp[q][k] = double(sq[k]);
    } else if (active && lane >= 4 && lane < 13) {
      const int ty = lane - 4;
      double xd, yd;
      sh_tying_pt(ty, xd, yd);
      T U[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, D[3] = {T(0), T(0), T(0)};
#pragma unroll
      for (int a = 0; a < 4; a++) {
        T N, Nx, Ny;
        sh_shape<T>(a, T(xd), T(yd), N, Nx, Ny);
        const T* ua = &L.u[6 * a];
        const T* f = L.fn[a];
        // director of node a: q_a x fn_a
        const T da[3] = {ua[4] * f[2] - ua[5] * f[1], ua[5] * f[0] - ua[3] * f[2],
                         ua[3] * f[1] - ua[4] * f[0]};
#pragma unroll
        for (int k = 0; k < 3; k++) {
          U[2 * k] += Nx * ua[k];
          U[2 * k + 1] += Ny * ua[k];
          D[k] += N * da[k];
        }
      }
#pragma unroll
      for (int k = 0; k < 6; k++) Gs.PU[ty][k] = double(U[k]);
#pragma unroll
      for (int k = 0; k < 3; k++) Gs.PD[ty][k] = double(D[k]);
    } else if (active && lane >= 13 && lane < 17) {
      const int q = lane - 13;
      const T x = T(sh_gauss(q & 1)), y = T(sh_gauss(q >> 1));
      T ux[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, d0[3] = {T(0), T(0), T(0)},
        dx[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
#pragma unroll
      for (int a = 0; a < 4; a++) {
        T N, Nx, Ny;
        sh_shape<T>(a, x, y, N, Nx, Ny);
        const T* ua = &L.u[6 * a];
        const T* f = L.fn[a];
        const T da[3] = {ua[4] * f[2] - ua[5] * f[1], ua[5] * f[0] - ua[3] * f[2],
                         ua[3] * f[1] - ua[4] * f[0]};
#pragma unroll
        for (int k = 0; k < 3; k++) {
          ux[2 * k] += Nx * ua[k];
          ux[2 * k + 1] += Ny * ua[k];
          d0[k] += N * da[k];
          dx[2 * k] += Nx * da[k];
          dx[2 * k + 1] += Ny * da[k];
        }
      }
      // u0d = [u0,xi | d0], d0d = [d0,xi | 0] (rows: components)
      T u0d[9], d0d[9];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        u0d[3 * k] = ux[2 * k];
        u0d[3 * k + 1] = ux[2 * k + 1];
        u0d[3 * k + 2] = d0[k];
        d0d[3 * k] = dx[2 * k];
        d0d[3 * k + 1] = dx[2 * k + 1];
        d0d[3 * k + 2] = T(0);
      }
      const T* Tm = L.qT[q];
      T a1[9], a2[9], u0x[9], u1x[9];
      sh_mm<T>(u0d, L.qXT[q], a1);  // u0x = T^T u0d XT
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
          u0x[3 * i + j] = Tm[i] * a1[j] + Tm[3 + i] * a1[3 + j] + Tm[6 + i] * a1[6 + j];
      sh_mm<T>(d0d, L.qXT[q], a1);  // u1x = T^T (d0d XT + u0d XzT)
      sh_mm<T>(u0d, L.qXzT[q], a2);
#pragma unroll
      for (int k = 0; k < 9; k++) a1[k] += a2[k];
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
          u1x[3 * i + j] = Tm[i] * a1[j] + Tm[3 + i] * a1[3 + j] + Tm[6 + i] * a1[6 + j];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        Gs.PU0[q][k] = double(u0x[k]);
        Gs.PU1[q][k] = double(u1x[k]);
      }
    }
    // (1b) unit columns: p, w0, w1 at every point (u0x = p w0^T, u1x = p w1^T)
    if (active && lane < SH_NV) {
      const int a = lane / 6, k = lane % 6;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const T x = T(sh_gauss(q & 1)), y = T(sh_gauss(q >> 1));
        T N, Nx, Ny;
        sh_shape<T>(a, x, y, N, Nx, Ny);
        const T* Tm = L.qT[q];
        const T* XT = L.qXT[q];
        const T* XzT = L.qXzT[q];
        T pv[3], w0[3], w1[3];
        if (k < 3) {
#pragma unroll
          for (int j = 0; j < 3; j++) {
            pv[j] = Tm[3 * k + j];
            w0[j] = Nx * XT[j] + Ny * XT[3 + j];
            w1[j] = Nx * XzT[j] + Ny * XzT[3 + j];
          }
        } else {
          const int r = k - 3;
          const T* f = L.fn[a];
          T da[3];
          da[0] = (r == 0) ? T(0) : ((r == 1) ? f[2] : -f[1]);
          da[1] = (r == 0) ? -f[2] : ((r == 1) ? T(0) : f[0]);
          da[2] = (r == 0) ? f[1] : ((r == 1) ? -f[0] : T(0));
#pragma unroll
          for (int j = 0; j < 3; j++) {
            pv[j] = Tm[j] * da[0] + Tm[3 + j] * da[1] + Tm[6 + j] * da[2];
            w0[j] = N * XT[6 + j];
            w1[j] = Nx * XT[j] + Ny * XT[3 + j] + N * XzT[6 + j];
          }
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
          Gs.PW[q][lane][j] = double(pv[j]);
          Gs.PW[q][lane][3 + j] = double(w0[j]);
          Gs.PW[q][lane][6 + j] = double(w1[j]);
        }
      }
    }
    __syncthreads();
    // (2) tying weights W_t of s_p (lane 0)
    if (active && lane == 0) {
      double W[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int q = 0; q < 4; q++) {
        const double* sq = Gs.sp[q];
        const double S[9] = {sq[0], sq[2], sq[7], sq[2], sq[1], sq[6], sq[7], sq[6], 0.0};
        double XT[9], A1[9], Mq[9];
#pragma unroll
        for (int k = 0; k < 9; k++) XT[k] = double(L.qXT[q][k]);
        // M = XT S XT^T
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) A1[3 * i + j] = XT[3 * i] * S[j] + XT[3 * i + 1] * S[3 + j] + XT[3 * i + 2] * S[6 + j];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            Mq[3 * i + j] = A1[3 * i] * XT[3 * j] + A1[3 * i + 1] * XT[3 * j + 1] + A1[3 * i + 2] * XT[3 * j + 2];
        const double x = sh_gauss(q & 1), y = sh_gauss(q >> 1);
        const double na0 = 0.5 * (1.0 - x), na1 = 0.5 * (1.0 + x);
        const double nb0 = 0.5 * (1.0 - y), nb1 = 0.5 * (1.0 + y);
        const double dw = double(L.detw[q]);
        W[0] += dw * Mq[0] * nb0;
        W[1] += dw * Mq[0] * nb1;
        W[2] += dw * Mq[4] * na0;
        W[3] += dw * Mq[4] * na1;
        W[4] += dw * 2.0 * Mq[1];
        W[5] += dw * 2.0 * Mq[5] * na0;
        W[6] += dw * 2.0 * Mq[5] * na1;
        W[7] += dw * 2.0 * Mq[2] * nb0;
        W[8] += dw * 2.0 * Mq[2] * nb1;
      }
#pragma unroll
      for (int k = 0; k < 9; k++) Gs.W[k] = W[k];
    }
    // (3) B1[:, c] = 2 Q(p, e_c) (lanes 0-23)
    if (active && lane < SH_NV) {
      const int c = lane, a = c / 6, k = c % 6;
      double gq[9];  // 2 g_t(p, e_c)
#pragma unroll
      for (int ty = 0; ty < 9; ty++) {
        double xd, yd;
        sh_tying_pt(ty, xd, yd);
        double N, Nx, Ny;
        sh_shape<double>(a, xd, yd, N, Nx, Ny);
        double Uc[6] = {0, 0, 0, 0, 0, 0}, Dc[3] = {0, 0, 0};
        if (k < 3) {
          Uc[2 * k] = Nx;
          Uc[2 * k + 1] = Ny;
        } else {
          const int r = k - 3;
          const double f0 = double(L.fn[a][0]), f1 = double(L.fn[a][1]), f2 = double(L.fn[a][2]);
          Dc[0] = N * ((r == 0) ? 0.0 : ((r == 1) ? f2 : -f1));
          Dc[1] = N * ((r == 0) ? -f2 : ((r == 1) ? 0.0 : f0));
          Dc[2] = N * ((r == 0) ? f1 : ((r == 1) ? -f0 : 0.0));
        }
        const double* Up = Gs.PU[ty];
        const double* Dp = Gs.PD[ty];
        double g;
        if (ty < 2) g = 0.5 * (Up[0] * Uc[0] + Up[2] * Uc[2] + Up[4] * Uc[4]);
        else if (ty < 4) g = 0.5 * (Up[1] * Uc[1] + Up[3] * Uc[3] + Up[5] * Uc[5]);
        else if (ty < 5)
          g = 0.25 * (Up[0] * Uc[1] + Up[2] * Uc[3] + Up[4] * Uc[5] + Uc[0] * Up[1] + Uc[2] * Up[3] +
                      Uc[4] * Up[5]);
        else if (ty < 7)
          g = 0.25 * (Dp[0] * Uc[1] + Dp[1] * Uc[3] + Dp[2] * Uc[5] + Dc[0] * Up[1] + Dc[1] * Up[3] +
                      Dc[2] * Up[5]);
        else
          g = 0.25 * (Dp[0] * Uc[0] + Dp[1] * Uc[2] + Dp[2] * Uc[4] + Dc[0] * Up[0] + Dc[1] * Up[2] +
                      Dc[2] * Up[4]);
        gq[ty] = 2.0 * g;
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const double x = sh_gauss(q & 1), y = sh_gauss(q >> 1);
        const double na0 = 0.5 * (1.0 - x), na1 = 0.5 * (1.0 + x);
        const double nb0 = 0.5 * (1.0 - y), nb1 = 0.5 * (1.0 + y);
        const double g11 = nb0 * gq[0] + nb1 * gq[1], g22 = na0 * gq[2] + na1 * gq[3];
        const double g12 = gq[4], g23 = na0 * gq[5] + na1 * gq[6], g13 = nb0 * gq[7] + nb1 * gq[8];
        const double Gm[9] = {g11, g12, g13, g12, g22, g23, g13, g23, 0.0};
        double XT[9], Wm[9], E0[9];
#pragma unroll
        for (int i = 0; i < 9; i++) XT[i] = double(L.qXT[q][i]);
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            Wm[3 * i + j] = Gm[3 * i] * XT[j] + Gm[3 * i + 1] * XT[3 + j] + Gm[3 * i + 2] * XT[6 + j];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
            E0[3 * i + j] = XT[i] * Wm[j] + XT[3 + i] * Wm[3 + j] + XT[6 + i] * Wm[6 + j];
        // bending: u0x_c = pc w0c^T, u1x_c = pc w1c^T; path u0x, u1x full
        const double* pw = Gs.PW[q][c];
        const double* P0 = Gs.PU0[q];
        const double* P1 = Gs.PU1[q];
        double u0c[9], u1c[9];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            u0c[3 * i + j] = pw[i] * pw[3 + j];
            u1c[3 * i + j] = pw[i] * pw[6 + j];
          }
        double b3 = 0.0, b4 = 0.0, b5 = 0.0;
        for (int i = 0; i < 3; i++) {
          b3 += P0[3 * i] * u1c[3 * i] + u0c[3 * i] * P1[3 * i];
          b4 += P0[3 * i + 1] * u1c[3 * i + 1] + u0c[3 * i + 1] * P1[3 * i + 1];
          b5 += P0[3 * i] * u1c[3 * i + 1] + u0c[3 * i] * P1[3 * i + 1] + P1[3 * i] * u0c[3 * i + 1] +
                u1c[3 * i] * P0[3 * i + 1];
        }
        // 2 Q = the symmetric sums above (each the sum of both orders)
        double* b1 = &Gs.B1[q][0][0];
        b1[0 * 24 + c] = E0[0];
        b1[1 * 24 + c] = E0[4];
        b1[2 * 24 + c] = 2.0 * E0[1];
        b1[3 * 24 + c] = b3;
        b1[4 * 24 + c] = b4;
        b1[5 * 24 + c] = b5;
        b1[6 * 24 + c] = 2.0 * E0[5];
        b1[7 * 24 + c] = 2.0 * E0[2];
        b1[8 * 24 + c] = 0.0;
      }
    }
    __syncthreads();
    // (4) column c of G (lanes 0-23)
    if (active && lane < SH_NV) {
      const int c = lane, ac = c / 6, kc = c % 6;
      double CB0[4][9], CB1[4][9];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        T b[9], sb[9];
#pragma unroll
        for (int k = 0; k < 9; k++) b[k] = L.B[q][k][c];
        sh_stress<T>(M, t, b, sb);
#pragma unroll
        for (int k = 0; k < 9; k++) CB0[q][k] = double(sb[k]) * double(L.detw[q]);
#pragma unroll
        for (int k = 0; k < 9; k++) b[k] = T(Gs.B1[q][k][c]);
        sh_stress<T>(M, t, b, sb);
#pragma unroll
        for (int k = 0; k < 9; k++) CB1[q][k] = double(sb[k]) * double(L.detw[q]);
      }
      double fc[3] = {double(L.fn[ac][0]), double(L.fn[ac][1]), double(L.fn[ac][2])};
      T* kout = &ke[(size_t)el * SH_NV * SH_NV + c];
      for (int r = 0; r < SH_NV; r++) {
        const int ar = r / 6, kr = r % 6;
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
          for (int k = 0; k < 9; k++)
            acc += Gs.B1[q][k][r] * CB0[q][k] + double(L.B[q][k][r]) * CB1[q][k];
        // tying part of s_p . 2 Q(e_r, e_c)
        double h = 0.0;
        const bool tr = kr < 3, tc = kc < 3;
        if ((tr && tc && kr == kc) || (tr != tc)) {
#pragma unroll
          for (int ty = 0; ty < 9; ty++) {
            double xd, yd;
            sh_tying_pt(ty, xd, yd);
            double Nr, Nxr, Nyr, Nc, Nxc, Nyc;
            sh_shape<double>(ar, xd, yd, Nr, Nxr, Nyr);
            sh_shape<double>(ac, xd, yd, Nc, Nxc, Nyc);
            double g = 0.0;
            if (tr && tc) {
              if (ty < 2) g = 0.5 * Nxr * Nxc;
              else if (ty < 4) g = 0.5 * Nyr * Nyc;
              else if (ty < 5) g = 0.25 * (Nxr * Nyc + Nxc * Nyr);
            } else if (ty >= 5) {
              // one translation (node at, component kt) and one rotation
              // (node ad, axis rd): 0.25 N_ad (e_rd x fn_ad)_kt dN_at
              const int at = tr ? ar : ac, kt = tr ? kr : kc;
              const int ad = tr ? ac : ar, rd = (tr ? kc : kr) - 3;
              const double Nd = tr ? Nc : Nr;
              const double Ntx = tr ? Nxr : Nxc, Nty = tr ? Nyr : Nyc;
              const double f0 = double(L.fn[ad][0]), f1 = double(L.fn[ad][1]), f2 = double(L.fn[ad][2]);
              const double dk = (kt == 0) ? ((rd == 0) ? 0.0 : ((rd == 1) ? f2 : -f1))
                                          : ((kt == 1) ? ((rd == 0) ? -f2 : ((rd == 1) ? 0.0 : f0))
                                                       : ((rd == 0) ? f1 : ((rd == 1) ? -f0 : 0.0)));
              (void)at;
              g = 0.25 * Nd * dk * (ty < 7 ? Nty : Ntx);
            }
            h += Gs.W[ty] * 2.0 * g;
          }
        }
        // bending part: 2 Q(e_r, e_c) at the points, weighted by detw s_p
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const double* pr = Gs.PW[q][r];
          const double* pc = Gs.PW[q][c];
          const double pp = pr[0] * pc[0] + pr[1] * pc[1] + pr[2] * pc[2];
          if (pp != 0.0) {
            const double q3 = pp * (pr[3] * pc[6] + pc[3] * pr[6]);
            const double q4 = pp * (pr[4] * pc[7] + pc[4] * pr[7]);
            const double q5 = pp * (pr[3] * pc[7] + pc[3] * pr[7] + pr[6] * pc[4] + pc[6] * pr[4]);
            const double* sq = Gs.sp[q];
            h += double(L.detw[q]) * (sq[3] * q3 + sq[4] * q4 + sq[5] * q5);
          }
        }
        kout[(size_t)r * SH_NV] += T(P.scale * (acc + h));
      }
    }
    return;
  }

  if (MODE == SH_MATRIX) {
    if (active && lane < SH_NV) {
      // column c of K = sum_q detw_q B_q^T Cs B_q; residual r_c = K[:,c].u (K symmetric)
      T s[4][9];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        T b[9];
#pragma unroll
        for (int kk = 0; kk < 9; kk++) b[kk] = L.B[q][kk][lane];
        sh_stress<T>(M, t, b, s[q]);
#pragma unroll
        for (int kk = 0; kk < 9; kk++) s[q][kk] *= L.detw[q];
      }
      // consistent mass (TACSShellElement::addJacobian, TACSShellElement.h:
      // 605-617, and TACSLinearizedRotation::addDirectorJacobian,
      // TACSDirector.h:381-467): moments m0 = rho t, m1 = -rho t^2 toff,
      // m2 = rho t^3 (toff^2 + 1/12) (TACSIsoShellConstitutive.cpp:120-129).
      // Node pair (b row, a column), cN = sum_q detw_q N_b N_a:
      //   u-u block  m0 cN I,
      //   q-q block  -m2 cN fn_b^x fn_a^x = m2 cN ((fn_b . fn_a) I - fn_a fn_b^T),
      //   q-u block  m1 cN fn_b^x, u-q block its transpose.
      const int ca = lane / 6, ck = lane % 6;
      T cN[4] = {T(0), T(0), T(0), T(0)};
      const bool with_mass = P.mass != 0.0;
      if (with_mass) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          T x = T(sh_gauss(q & 1)), y = T(sh_gauss(q >> 1));
          T Na, Nx, Ny;
          sh_shape<T>(ca, x, y, Na, Nx, Ny);
#pragma unroll
          for (int b = 0; b < 4; b++) {
            T Nb;
            sh_shape<T>(b, x, y, Nb, Nx, Ny);
            cN[b] += L.detw[q] * Nb * Na;
          }
        }
      }
      const T m0 = T(rho) * t, m1 = -T(rho) * t * t * M.toff,
              m2 = T(rho) * t * t * t * (M.toff * M.toff + T(1) / T(12));
      T racc = T(0);
      T* kout = &ke[(size_t)el * SH_NV * SH_NV + lane];
      for (int r = 0; r < SH_NV; r++) {
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
          for (int kk = 0; kk < 9; kk++) acc += L.B[q][kk][r] * s[q][kk];
        T kv = acc * T(P.scale);
        if (with_mass) {
          const int rb = r / 6, rk = r % 6;
          T mv = T(0);
          if (rk < 3 && ck < 3) {
            if (rk == ck) mv = m0 * cN[rb];
          } else if (rk >= 3 && ck >= 3) {
            const T* fb = L.fn[rb];
            const T* fa = L.fn[ca];
            T dot = fb[0] * fa[0] + fb[1] * fa[1] + fb[2] * fa[2];
            mv = m2 * cN[rb] * ((rk == ck ? dot : T(0)) - fa[rk - 3] * fb[ck - 3]);
          } else if (rk >= 3) {
            // rotation row, displacement column: m1 cN (fn_b^x)[rk-3][ck]
            // (the d2Tdotu terms of addDirectorJacobian, TACSDirector.h:431-467)
            mv = m1 * cN[rb] * sh_skew(L.fn[rb], rk - 3, ck);
          } else {
            // displacement row, rotation column: the transposed entry
            mv = m1 * cN[rb] * sh_skew(L.fn[ca], ck - 3, rk);
          }
          kv += T(P.mass) * mv;
        }
        kout[(size_t)r * SH_NV] = kv;
        racc += acc * L.u[r];
      }
      if (ev) ev[(size_t)el * SH_NV + lane] = double(racc);
    }
    return;
  }

  // ---- strains of u (and psi) at the quadrature points: lanes 0..35
  if (active) {
    for (int j = lane; j < 36; j += SH_TEAM) {
      int q = j / 9, kk = j % 9;
      T acc = T(0), accp = T(0);
#pragma unroll
      for (int c = 0; c < SH_NV; c++) {
        acc += L.B[q][kk][c] * L.u[c];
        if (MODE == SH_ADJ_PRODUCT) accp += L.B[q][kk][c] * L.psi[c];
      }
      // the function evaluations use e[8] = 0 (TACSShellElement::evalPointQuantity)
      bool drop_drill = (MODE != SH_RESIDUAL && MODE != SH_ADJ_PRODUCT && MODE != SH_XPT_ADJ);
      L.e[q][kk] = (drop_drill && kk == 8) ? T(0) : acc;
      if (MODE == SH_ADJ_PRODUCT) L.ep[q][kk] = accp;
    }
  }
  __syncthreads();

  if (MODE == SH_XPT_ADJ) {
    // psi . r with r_c = sum_q detw_q B_q[:,c] . s_q, summed in lane order
    if (active && lane < SH_NV) {
      T acc = T(0);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        T s[9];
        sh_stress<T>(M, t, L.e[q], s);
        T dot = T(0);
#pragma unroll
        for (int kk = 0; kk < 9; kk++) dot += L.B[q][kk][lane] * s[kk];
        acc += L.detw[q] * dot;
      }
      L.red[lane] = acc * L.psi[lane];
    }
    __syncthreads();
    if (active && lane == 0) {
      T sum = T(0);
      for (int c = 0; c < SH_NV; c++) sum += L.red[c];
      es[(size_t)el * 12 + dir] = P.scale * sh_dpart(sum);
    }
    return;
  }

  if (MODE == SH_RESIDUAL || MODE == SH_COMPLIANCE_SV) {
    // r_c = sum_q detw_q B_q[:,c] . s_q  (s = Cs e, doubled for dC/du)
    if (active && lane < SH_NV) {
      T acc = T(0);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        T s[9];
        sh_stress<T>(M, t, L.e[q], s);
        T dot = T(0);
#pragma unroll
        for (int kk = 0; kk < 9; kk++) dot += L.B[q][kk][lane] * s[kk];
        acc += L.detw[q] * dot;
      }
      if (MODE == SH_COMPLIANCE_SV) acc *= T(2) * T(P.scale);
      ev[(size_t)el * SH_NV + lane] = double(acc);
    }
    return;
  }

  if (MODE == SH_KS_SV) {
    if (active && lane < 4) {
      int q = lane;
      T de[9], dt;
      T f = T(P.safety) * sh_failure<T>(M, t, L.e[q], de, dt);
      T coef = exp(T(P.ks_weight) * (f - T(P.ks_max))) / T(P.ks_sum) * L.detw[q] *
               T(P.safety) * T(P.scale);
#pragma unroll
      for (int kk = 0; kk < 9; kk++) L.de[q][kk] = coef * de[kk];
    }
    __syncthreads();
    if (active && lane < SH_NV) {
      T acc = T(0);
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int kk = 0; kk < 9; kk++) acc += L.B[q][kk][lane] * L.de[q][kk];
      ev[(size_t)el * SH_NV + lane] = double(acc);
    }
    return;
  }

  // ---- per-element scalars: lanes 0..3 evaluate one quadrature point each
  if (active && lane < 4) {
    int q = lane;
    const T* e = L.e[q];
    T v = T(0);
    if (MODE == SH_COMPLIANCE || MODE == SH_XPT_COMPLIANCE) {
      T s[9];
      sh_stress<T>(M, t, e, s);
      T dot = T(0);
#pragma unroll
      for (int kk = 0; kk < 9; kk++) dot += e[kk] * s[kk];
      v = L.detw[q] * dot;
    } else if (MODE == SH_KS_MAX || MODE == SH_KS_SUM || MODE == SH_KS_DV || MODE == SH_XPT_KS) {
      T de[9], dt;
      T f = T(P.safety) * sh_failure<T>(M, t, e, de, dt);
      if (MODE == SH_KS_MAX) v = f;
      else if (MODE == SH_KS_SUM || MODE == SH_XPT_KS)
        v = L.detw[q] * exp(T(P.ks_weight) * (f - T(P.ks_max)));
      else
        v = T(P.scale) * exp(T(P.ks_weight) * (f - T(P.ks_max))) / T(P.ks_sum) * L.detw[q] *
            T(P.safety) * dt;
    } else if (MODE == SH_ADJ_PRODUCT || MODE == SH_COMPLIANCE_DV) {
      // TACSIsoShellConstitutive::addStressDVSens (:229-251): d/dt of A, B, D
      const T* x = (MODE == SH_ADJ_PRODUCT) ? L.ep[q] : e;
      T a0 = M.Q0 * x[0] + M.Q1 * x[1], a1 = M.Q1 * x[0] + M.Q0 * x[1], a2 = M.Q5 * x[2];
      T b0 = M.Q0 * x[3] + M.Q1 * x[4], b1 = M.Q1 * x[3] + M.Q0 * x[4], b2 = M.Q5 * x[5];
      T dI = (T(3) * M.toff * M.toff + T(0.25)) * t * t;
      v = a0 * e[0] + a1 * e[1] + a2 * e[2] + dI * (b0 * e[3] + b1 * e[4] + b2 * e[5]) +
          M.kcorr * M.Q5 * (x[6] * e[6] + x[7] * e[7] + M.drill_reg * x[8] * e[8]);
      if (M.toff != T(0))
        v += T(-2) * M.toff * t *
             (a0 * e[3] + a1 * e[4] + a2 * e[5] + b0 * e[0] + b1 * e[1] + b2 * e[2]);
      v *= T(P.scale) * L.detw[q];
    }
    L.val[q] = v;
  }
  __syncthreads();
  if (XPT) {
    if (active && lane == 0) {
      const T sum = ((L.val[0] + L.val[1]) + L.val[2]) + L.val[3];
      double dv = sh_dpart(sum);
      if (MODE == SH_XPT_KS) dv /= P.ks_weight * P.ks_sum;
      es[(size_t)el * 12 + dir] = P.scale * dv;
    }
    return;
  }
  if (active && lane == 0) {
    double r;
    if (MODE == SH_KS_MAX)
      r = fmax(fmax(double(L.val[0]), double(L.val[1])), fmax(double(L.val[2]), double(L.val[3])));
    else
      r = ((double(L.val[0]) + double(L.val[1])) + double(L.val[2])) + double(L.val[3]);
    es[el] = r;
  }
}

// Sum element matrices into BCSR rows: one wavefront per block row, values
// of the row spread over the lanes; contributions added in ascending element
// order (n2e entries are (element << 5 | local node)).  Dirichlet rows are
// fused: constrained rows become zero with a unit diagonal, columns are kept
// (TACSParallelMat::applyBCs, BCSRMat::zeroRow src/bpmat/BCSRMat.cpp:2027).
template <int B, int NPE, typename T>
__global__ __launch_bounds__(256) void gather_elem_matrices(
    int nrows, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
    const int* __restrict__ conn, const T* __restrict__ ke, const int* __restrict__ rowp,
    const int* __restrict__ cols, const int* __restrict__ bcvars, double* __restrict__ A,
    float* __restrict__ A32) {
  constexpr int NV = B * NPE, BB = B * B;
  const int lane = threadIdx.x % TACS_WAVE;
  const int wpb = blockDim.x / TACS_WAVE;
  for (int row = blockIdx.x * wpb + threadIdx.x / TACS_WAVE; row < nrows; row += gridDim.x * wpb) {
    const int r0 = rowp[row], nvals = (rowp[row + 1] - r0) * BB;
    const int p0 = n2e_ptr[row], p1 = n2e_ptr[row + 1];
    const int mask = bcvars ? bcvars[row] : 0;
    for (int v = lane; v < nvals; v += TACS_WAVE) {
      const int j = v / BB, rc = v % BB, i = rc / B, jj = rc % B;
      const int col = cols[r0 + j];
      double acc = 0.0;
      for (int p = p0; p < p1; p++) {
        const int packed = n2e[p];
        const int e = packed >> 5, la = packed & 31;
        const int* ce = &conn[(size_t)e * NPE];
#pragma unroll
        for (int lb = 0; lb < NPE; lb++)
          if (ce[lb] == col) acc += double(ke[(size_t)e * NV * NV + (size_t)(B * la + i) * NV + B * lb + jj]);
      }
      if (mask & (1 << i)) acc = (col == row && i == jj) ? 1.0 : 0.0;
      A[(size_t)r0 * BB + v] = acc;
      if (A32) A32[(size_t)r0 * BB + v] = (float)acc;
    }
  }
}

// Node gather of element vectors with B values per node, ascending element order.
template <int B, int NV>
__global__ void gather_elem_vectors_b(int nnodes, const int* __restrict__ n2e_ptr,
                                      const int* __restrict__ n2e, const double* __restrict__ ev,
                                      int accumulate, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  double s[B];
#pragma unroll
  for (int k = 0; k < B; k++) s[k] = accumulate ? out[(size_t)B * i + k] : 0.0;
  for (int p = n2e_ptr[i]; p < n2e_ptr[i + 1]; p++) {
    int packed = n2e[p];
    const double* v = &ev[(size_t)(packed >> 5) * NV + B * (packed & 31)];
#pragma unroll
    for (int k = 0; k < B; k++) s[k] += v[k];
  }
#pragma unroll
  for (int k = 0; k < B; k++) out[(size_t)B * i + k] = s[k];
}

// Uniform pressure on shell elements (TACSShellPressure::addResidual,
// TACSShellPressure.h:51-91): ev += N^T f, f = -p detXd n at each quadrature
// point, n the interpolated node normal.  One thread per loaded element.
template <typename T>
__device__ __forceinline__ void sh_pressure_elem(const T X[12], T p, T r[24]) {
  T fn[12];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    T g[6];
    sh_grad3<T>(X, (i & 1) ? T(1) : T(-1), (i & 2) ? T(1) : T(-1), g);
    T a[3] = {g[0], g[2], g[4]}, b[3] = {g[1], g[3], g[5]}, f[3];
    sh_cross<T>(a, b, f);
    T nrm = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (nrm != T(0)) {
      T rn = T(1) / nrm;
      f[0] *= rn;
      f[1] *= rn;
      f[2] *= rn;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) fn[3 * i + k] = f[k];
  }
#pragma unroll
  for (int k = 0; k < 24; k++) r[k] = T(0);
#pragma unroll
  for (int qp = 0; qp < 4; qp++) {
    T x = T(sh_gauss(qp & 1)), y = T(sh_gauss(qp >> 1));
    T g[6], n[3], Xd[9], Xinv[9];
    sh_grad3<T>(X, x, y, g);
    sh_interp3<T>(fn, x, y, n);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      Xd[3 * k] = g[2 * k];
      Xd[3 * k + 1] = g[2 * k + 1];
      Xd[3 * k + 2] = n[k];
    }
    T det = sh_inv3<T>(Xd, Xinv);
    T f[3] = {p * -det * n[0], p * -det * n[1], p * -det * n[2]};
#pragma unroll
    for (int a = 0; a < 4; a++) {
      T N, Nx, Ny;
      sh_shape<T>(a, x, y, N, Nx, Ny);
#pragma unroll
      for (int k = 0; k < 3; k++) r[6 * a + k] += N * f[k];
    }
  }
}

__global__ void shell_pressure_vectors(int naux, const int* __restrict__ aux_elem,
                                       const double* __restrict__ pres, const int* __restrict__ conn,
                                       const double* __restrict__ Xg, double lambda,
                                       double* __restrict__ ev) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= naux) return;
  const int* c = &conn[(size_t)aux_elem[q] * 4];
  double X[12], r[24];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int k = 0; k < 3; k++) X[3 * a + k] = Xg[3 * (size_t)c[a] + k];
  sh_pressure_elem<double>(X, lambda * pres[q], r);
#pragma unroll
  for (int k = 0; k < 24; k++) ev[(size_t)q * 24 + k] = r[k];
}

// Node-coordinate derivative of the pressure load: one thread per (loaded
// element, coordinate k), out[12 q + k] = scale d/dX_k (psi . r).
__global__ void shell_pressure_xpt(int naux, const int* __restrict__ aux_elem,
                                   const double* __restrict__ pres, const int* __restrict__ conn,
                                   const double* __restrict__ Xg, const double* __restrict__ psi,
                                   double scale, double* __restrict__ out) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = gt / 12, dir = gt % 12;
  if (q >= naux) return;
  const int* c = &conn[(size_t)aux_elem[q] * 4];
  ShDual X[12], r[24];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int k = 0; k < 3; k++) X[3 * a + k] = sh_seed<ShDual>(Xg[3 * (size_t)c[a] + k], 3 * a + k == dir);
  sh_pressure_elem<ShDual>(X, ShDual(pres[q]), r);
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int k = 0; k < 3; k++) acc += psi[6 * (size_t)c[a] + k] * r[6 * a + k].d;
  out[(size_t)q * 12 + dir] = scale * acc;
}

extern "C" {

static ShellParams shell_params(const ShellMaterial* m, double scale, double ks_weight,
                                double ks_max, double ks_sum, double safety) {
  ShellParams P;
  P.Q0 = m->Q[0];
  P.Q1 = m->Q[1];
  P.Q5 = m->Q[5];
  P.kcorr = m->kcorr;
  P.drill_reg = m->drill_reg;
  P.ys = m->ys;
  P.scale = scale;
  P.ks_weight = ks_weight;
  P.ks_max = ks_max;
  P.ks_sum = ks_sum;
  P.safety = safety;
  P.mass = 0.0;
  P.rho = m->rho;
  P.ref_axis = m->ref_axis;
  P.toff = m->toff;
  P.mtab = m->table;
  P.emat = m->elem_mat;
  for (int k = 0; k < 3; k++) P.axis[k] = m->axis[k];
  return P;
}

#define SH_LAUNCH(T, MODE)                                                                        \
  hipLaunchKernelGGL((shell_elem_kernel<T, MODE>), dim3(grid), dim3(SH_TEAM * SH_EPB), 0, stream, \
                     nelem, conn, X, u, psi, telem, P, (T*)ke, ev, es)

int tacs_hip_shell_elem_kernel(int prec, int mode, int nelem, const int* conn, const double* X,
                               const double* u, const double* psi, const double* telem,
                               const ShellMaterial* mat, double scale, double ks_weight,
                               double ks_max, double ks_sum, double safety, void* ke, double* ev,
                               double* es, hipStream_t stream) {
  if (nelem <= 0) return 0;
  if (!mat || (prec != 64 && prec != 32)) return TACS_HIP_ERR_ARG;
  ShellParams P = shell_params(mat, scale, ks_weight, ks_max, ks_sum, safety);
  int grid = (nelem + SH_EPB - 1) / SH_EPB;
  if (prec == 32) {
    if (mode != SH_MATRIX) return TACS_HIP_ERR_ARG;
    SH_LAUNCH(float, SH_MATRIX);
    return (int)hipGetLastError();
  }
  switch (mode) {
    case SH_MATRIX: SH_LAUNCH(double, SH_MATRIX); break;
    case SH_RESIDUAL: SH_LAUNCH(double, SH_RESIDUAL); break;
    case SH_COMPLIANCE: SH_LAUNCH(double, SH_COMPLIANCE); break;
    case SH_KS_MAX: SH_LAUNCH(double, SH_KS_MAX); break;
    case SH_KS_SUM: SH_LAUNCH(double, SH_KS_SUM); break;
    case SH_KS_SV: SH_LAUNCH(double, SH_KS_SV); break;
    case SH_KS_DV: SH_LAUNCH(double, SH_KS_DV); break;
    case SH_ADJ_PRODUCT: SH_LAUNCH(double, SH_ADJ_PRODUCT); break;
    case SH_COMPLIANCE_SV: SH_LAUNCH(double, SH_COMPLIANCE_SV); break;
    case SH_COMPLIANCE_DV: SH_LAUNCH(double, SH_COMPLIANCE_DV); break;
    default: return TACS_HIP_ERR_ARG;
  }
  return (int)hipGetLastError();
}

int tacs_hip_shell_xpt_kernel(int mode, int nelem, const int* conn, const double* X,
                              const double* u, const double* psi, const double* telem,
                              const ShellMaterial* mat, double scale, double ks_weight,
                              double ks_max, double ks_sum, double safety, double* es,
                              hipStream_t stream) {
  if (nelem <= 0) return 0;
  if (!mat || !es || mode < 0 || mode > 2 || (mode == 0 && !psi)) return TACS_HIP_ERR_ARG;
  ShellParams P = shell_params(mat, scale, ks_weight, ks_max, ks_sum, safety);
  const int grid = (int)(((long)nelem * 12 + SH_EPB - 1) / SH_EPB);
  ShDual* ke = NULL;
  double* ev = NULL;
  if (mode == 0)
    hipLaunchKernelGGL((shell_elem_kernel<ShDual, SH_XPT_ADJ>), dim3(grid), dim3(SH_TEAM * SH_EPB),
                       0, stream, nelem, conn, X, u, psi, telem, P, ke, ev, es);
  else if (mode == 1)
    hipLaunchKernelGGL((shell_elem_kernel<ShDual, SH_XPT_COMPLIANCE>), dim3(grid),
                       dim3(SH_TEAM * SH_EPB), 0, stream, nelem, conn, X, u, psi, telem, P, ke, ev,
                       es);
  else
    hipLaunchKernelGGL((shell_elem_kernel<ShDual, SH_XPT_KS>), dim3(grid), dim3(SH_TEAM * SH_EPB),
                       0, stream, nelem, conn, X, u, psi, telem, P, ke, ev, es);
  return (int)hipGetLastError();
}

int tacs_hip_shell_pressure_xpt(int naux, const int* aux_elem, const double* pressure,
                                const int* conn, const double* X, const double* psi, double scale,
                                double* out, hipStream_t stream) {
  if (naux <= 0) return 0;
  const long nt = (long)naux * 12;
  hipLaunchKernelGGL(shell_pressure_xpt, dim3((unsigned)((nt + 127) / 128)), dim3(128), 0, stream,
                     naux, aux_elem, pressure, conn, X, psi, scale, out);
  return (int)hipGetLastError();
}

int tacs_hip_shell_mat_type(int prec, int nelem, const int* conn, const double* X,
                            const double* u, const double* telem, const ShellMaterial* mat,
                            double kscale, double mscale, double gscale, void* ke, double* ev,
                            hipStream_t stream) {
  if (gscale != 0.0 && (prec != 64 || !u)) return TACS_HIP_ERR_ARG;
  if (nelem <= 0) return 0;
  if (!mat || (prec != 64 && prec != 32) || (prec == 32 && ev)) return TACS_HIP_ERR_ARG;
  ShellParams P = shell_params(mat, kscale, 0.0, 0.0, 1.0, 1.0);
  P.mass = mscale;
  int grid = (nelem + SH_EPB - 1) / SH_EPB;
  const double* psi = NULL;
  double* es = NULL;
  if (prec == 32)
    SH_LAUNCH(float, SH_MATRIX);
  else
    SH_LAUNCH(double, SH_MATRIX);
  if (gscale != 0.0) {
    // + gscale G: the geometric stiffness along the state u, added to ke
    P.scale = gscale;
    ev = NULL;
    SH_LAUNCH(double, SH_GEO);
  }
  return (int)hipGetLastError();
}

int tacs_hip_gather_elem_matrices(int b, int npe, int prec, int nrows, const int* n2e_ptr,
                                  const int* n2e, const int* conn, const void* ke, const int* rowp,
                                  const int* cols, const int* bcvars, double* A, float* A32,
                                  hipStream_t stream) {
  if (nrows <= 0) return 0;
  int grid = tacs_grid_for(nrows, 4, 1 << 20);
  if (b == 6 && npe == 4 && prec == 64)
    hipLaunchKernelGGL((gather_elem_matrices<6, 4, double>), dim3(grid), dim3(256), 0, stream,
                       nrows, n2e_ptr, n2e, conn, (const double*)ke, rowp, cols, bcvars, A, A32);
  else if (b == 6 && npe == 4 && prec == 32)
    hipLaunchKernelGGL((gather_elem_matrices<6, 4, float>), dim3(grid), dim3(256), 0, stream,
                       nrows, n2e_ptr, n2e, conn, (const float*)ke, rowp, cols, bcvars, A, A32);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_gather_elem_vectors_b(int b, int nv, int nnodes, const int* n2e_ptr, const int* n2e,
                                   const double* ev, int accumulate, double* out,
                                   hipStream_t stream) {
  if (nnodes <= 0) return 0;
  int grid = (nnodes + 255) / 256;
  if (b == 6 && nv == 24)
    hipLaunchKernelGGL((gather_elem_vectors_b<6, 24>), dim3(grid), dim3(256), 0, stream, nnodes,
                       n2e_ptr, n2e, ev, accumulate, out);
  else if (b == 3 && nv == 24)
    hipLaunchKernelGGL((gather_elem_vectors_b<3, 24>), dim3(grid), dim3(256), 0, stream, nnodes,
                       n2e_ptr, n2e, ev, accumulate, out);
  else if (b == 3 && nv == 12)
    hipLaunchKernelGGL((gather_elem_vectors_b<3, 12>), dim3(grid), dim3(256), 0, stream, nnodes,
                       n2e_ptr, n2e, ev, accumulate, out);
  else if (b == 3 && nv == 81)
    hipLaunchKernelGGL((gather_elem_vectors_b<3, 81>), dim3(grid), dim3(256), 0, stream, nnodes,
                       n2e_ptr, n2e, ev, accumulate, out);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_shell_pressure_vectors(int naux, const int* aux_elem, const double* pressure,
                                    const int* conn, const double* X, double lambda, double* ev,
                                    hipStream_t stream) {
  if (naux <= 0) return 0;
  hipLaunchKernelGGL(shell_pressure_vectors, dim3((naux + 127) / 128), dim3(128), 0, stream, naux,
                     aux_elem, pressure, conn, X, lambda, ev);
  return (int)hipGetLastError();
}

}  // extern "C"
