// Kernels of the batched (multi right-hand side) GMRES: R independent GMRES
// recurrences that share every pass over the matrix.  A block vector stores
// the R right-hand sides interleaved, V[i*R + j] (dof i, right-hand side j),
// the layout of the block-CSR SpMM (bsr3_spmm), so one SpMM reads the matrix
// once for all R systems.  Per right-hand side j: Hessenberg H_j (column-major,
// (m+1) x m), rotations and residual vector are contiguous slices of device
// arrays with stride sH / sR; the Givens update and back-substitution run one
// thread per right-hand side.  R <= 64.
#include "tacs_hip_common.h"

#define BRED_BLOCKS 1024
#define BRED_THREADS 256
#define BRED_KMAX 16

static inline int bred_blocks(size_t rows) {
  size_t b = rows / 256;
  if (b < 1) b = 1;
  if (b > BRED_BLOCKS) b = BRED_BLOCKS;
  return (int)b;
}

// Round R up to a power of two (threads per row group).
__host__ __device__ inline int pow2_ceil(int r) {
  int p = 1;
  while (p < r) p <<= 1;
  return p;
}

// Partial multi-dots: part[((v * R + j) * NB) + b] = block b's share of
// sum_i W_v(i, j) q[i*R + j], v < k <= BRED_KMAX, where W_v(i, j) is the block
// vector W[v*ldw + i*R + j] or, SHARED, the single vector W[v*ldw + i] common
// to every right-hand side (a recycled space).  Thread t handles right-hand
// side j = t % Rp of rows t / Rp, t / Rp + 256/Rp, ...
template <bool SHARED>
__global__ __launch_bounds__(BRED_THREADS) void bmdot_partial(size_t n, int R, int k,
                                                              const double* __restrict__ W,
                                                              size_t ldw,
                                                              const double* __restrict__ q,
                                                              double* __restrict__ part) {
  __shared__ double sh[BRED_KMAX][BRED_THREADS];
  const int Rp = pow2_ceil(R);
  const int j = threadIdx.x % Rp, r0 = threadIdx.x / Rp, rstride = BRED_THREADS / Rp;
  double acc[BRED_KMAX];
#pragma unroll
  for (int v = 0; v < BRED_KMAX; v++) acc[v] = 0.0;
  if (j < R) {
    for (size_t i = (size_t)blockIdx.x * rstride + r0; i < n; i += (size_t)gridDim.x * rstride) {
      double qi = q[i * R + j];
      const size_t w0 = SHARED ? i : i * R + j;
#pragma unroll
      for (int v = 0; v < BRED_KMAX; v++)
        if (v < k) acc[v] += W[(size_t)v * ldw + w0] * qi;
    }
  }
#pragma unroll
  for (int v = 0; v < BRED_KMAX; v++) sh[v][threadIdx.x] = acc[v];
  __syncthreads();
  for (int t = threadIdx.x; t < k * R; t += BRED_THREADS) {
    int v = t / R, jj = t % R;
    double s = 0.0;
    for (int rr = 0; rr < rstride; rr++) s += sh[v][rr * Rp + jj];
    part[((size_t)v * R + jj) * gridDim.x + blockIdx.x] = s;
  }
}

// out[j * sR + v] (+)= sum_b part[(v*R + j)*nb + b]; one wave per (v, j)
__global__ void bmdot_finalize(int nb, int k, int R, const double* __restrict__ part,
                               double* __restrict__ out, size_t sR, int accumulate) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= k * R) return;
  const int v = w / R, j = w % R;
  const double* p = &part[(size_t)w * nb];
  double t = 0.0;
  for (int b = lane; b < nb; b += 64) t += p[b];
  t = wave_sum(t);
  if (lane == 0) {
    double* o = &out[(size_t)j * sR + v];
    *o = accumulate ? *o + t : t;
  }
}

// q[i*R + j] -= sum_{v<k} W_v(i, j) c[j*sR + v]  (W_v(i, j) as in bmdot_partial)
template <bool SHARED>
__global__ void bcgs_update(size_t n, int R, int k, const double* __restrict__ W, size_t ldw,
                            const double* __restrict__ c, size_t sR, double* __restrict__ q) {
  const size_t N = n * R;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (size_t)gridDim.x * blockDim.x) {
    int j = (int)(t % R);
    const size_t w0 = SHARED ? t / R : t;
    double s = q[t];
    for (int v = 0; v < k; v++) s -= W[(size_t)v * ldw + w0] * c[(size_t)j * sR + v];
    q[t] = s;
  }
}

// y[i*R + j] (+)= sum_{v < min(cnt[j], kmax)} W_v(i, j) c[j*sR + v] (cnt NULL: all kmax)
template <bool SHARED>
__global__ void blincomb(size_t n, int R, int kmax, const double* __restrict__ W, size_t ldw,
                         const double* __restrict__ c, size_t sR, const int* __restrict__ cnt,
                         double* __restrict__ y, int add) {
  const size_t N = n * R;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (size_t)gridDim.x * blockDim.x) {
    int j = (int)(t % R);
    const size_t w0 = SHARED ? t / R : t;
    int kj = (cnt && cnt[j] < kmax) ? cnt[j] : kmax;
    double s = add ? y[t] : 0.0;
    for (int v = 0; v < kj; v++) s += W[(size_t)v * ldw + w0] * c[(size_t)j * sR + v];
    y[t] = s;
  }
}

// z_j[v] = sum_{i < cnt[j]} B_j[v + i*ldb] y_j[i], v < k: the recycled-space
// coefficients of a cycle's correction, one thread per (j, v)
__global__ void bsmall_gemv(int R, int k, const int* __restrict__ cnt,
                            const double* __restrict__ B, size_t sB, int ldb,
                            const double* __restrict__ y, size_t sy, double* __restrict__ z,
                            size_t sz) {
  int t = threadIdx.x + blockIdx.x * blockDim.x;
  if (t >= R * k) return;
  int j = t / k, v = t % k;
  const double* b = &B[(size_t)j * sB];
  const double* yj = &y[(size_t)j * sy];
  double s = 0.0;
  for (int i = 0; i < cnt[j]; i++) s += b[v + (size_t)i * ldb] * yj[i];
  z[(size_t)j * sz + v] = s;
}

// ---------------------------------------------------------------------------
// Delayed classical Gram-Schmidt (DCGS2: D. Bielich, J. Langou, S. Thomas,
// K. Swirydowicz, I. Yamazaki, E. G. Boman, "Low-synch Gram-Schmidt with
// delayed reorthogonalization for Krylov solvers", Parallel Computing 112
// (2022)) needs, per Arnoldi step, one sweep of dot products of the basis
// against two vectors and one sweep updating two vectors.
// ---------------------------------------------------------------------------
#define BDQ_KMAX 32
#define BDQ_C 4

// Partial dots of the concatenated basis [W1 (k1 vectors), W2 (k2 vectors)]
// (block vectors, stride ldw; entries v0 .. v0 + kk of the list, kk <=
// BDQ_KMAX) against NQ vectors q_0, q_1: part[((v*NQ + qi)*R + j)*nb + blk].
// A block is 4 wavefronts over the same rows; wave w takes the basis vectors
// v0 + 8w .. v0 + 8w + 7 (so each thread keeps 8 NQ sums and the q entries,
// read by all four waves, come from cache).  Within a wave, lanes are (row
// group g, right-hand side j); the G row groups are reduced by lane shuffles.
template <int NQ>
__global__ __launch_bounds__(256) void bdotq_partial(size_t n, int R, int v0, int kk, int k1,
                                                     const double* __restrict__ W1,
                                                     const double* __restrict__ W2, size_t ldw,
                                                     const double* __restrict__ q0,
                                                     const double* __restrict__ q1,
                                                     double* __restrict__ part) {
  const int Rp = pow2_ceil(R);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int j = lane % Rp, g = lane / Rp, G = 64 / Rp;
  const int vb = wv * 8;  // first vector (relative to v0) of this wave
  const int nv = min(8, kk - vb);
  if (nv <= 0) return;
  const double* Wv[8];
#pragma unroll
  for (int v = 0; v < 8; v++) {
    const int vv = v0 + vb + (v < nv ? v : 0);
    Wv[v] = vv < k1 ? W1 + (size_t)vv * ldw : W2 + (size_t)(vv - k1) * ldw;
  }
  double acc[NQ][8];
#pragma unroll
  for (int qi = 0; qi < NQ; qi++)
#pragma unroll
    for (int v = 0; v < 8; v++) acc[qi][v] = 0.0;
  // each pass takes BDQ_C consecutive groups of G rows, so a wave reads
  // BDQ_C * 512 contiguous bytes of every vector (with 512-byte pieces 2 MB
  // apart the sweep reached 4.6 TB/s at 128^3, with 2 KB pieces 5.1 TB/s)
  if (j < R) {
    for (size_t i0 = (size_t)blockIdx.x * G * BDQ_C; i0 < n; i0 += (size_t)gridDim.x * G * BDQ_C) {
      double a0[BDQ_C], a1[BDQ_C];
      bool in[BDQ_C];
#pragma unroll
      for (int c = 0; c < BDQ_C; c++) {
        const size_t i = i0 + (size_t)G * c + g;
        in[c] = i < n;
        const size_t t = in[c] ? i * R + j : 0;
        a0[c] = in[c] ? q0[t] : 0.0;
        a1[c] = (NQ > 1 && in[c]) ? q1[t] : 0.0;
      }
#pragma unroll
      for (int v = 0; v < 8; v++) {
        if (v < nv) {
          double w[BDQ_C];
#pragma unroll
          for (int c = 0; c < BDQ_C; c++)
            w[c] = in[c] ? Wv[v][(i0 + (size_t)G * c + g) * R + j] : 0.0;
#pragma unroll
          for (int c = 0; c < BDQ_C; c++) {
            acc[0][v] += w[c] * a0[c];
            if (NQ > 1) acc[NQ - 1][v] += w[c] * a1[c];
          }
        }
      }
    }
  }
  for (int off = Rp; off < 64; off <<= 1) {
#pragma unroll
    for (int qi = 0; qi < NQ; qi++)
#pragma unroll
      for (int v = 0; v < 8; v++) acc[qi][v] += __shfl_xor(acc[qi][v], off, 64);
  }
  if (g == 0 && j < R) {
#pragma unroll
    for (int qi = 0; qi < NQ; qi++)
#pragma unroll
      for (int v = 0; v < 8; v++)
        if (v < nv)
          part[((size_t)((vb + v) * NQ + qi) * R + j) * gridDim.x + blockIdx.x] = acc[qi][v];
  }
}

// out_qi[j*sR + v0 + v] = sum_b part[((v*NQ + qi)*R + j)*nb + b]; one wave each
__global__ void bdotq_finalize(int nb, int kk, int NQ, int R, const double* __restrict__ part,
                               double* __restrict__ out0, double* __restrict__ out1, size_t sR,
                               int v0) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= kk * NQ * R) return;
  const int j = w % R, vq = w / R, qi = vq % NQ, v = vq / NQ;
  const double* p = &part[(size_t)w * nb];
  double s = 0.0;
  for (int b = lane; b < nb; b += 64) s += p[b];
  s = wave_sum(s);
  if (lane == 0) (qi == 0 ? out0 : out1)[(size_t)j * sR + v0 + v] = s;
}

// The update sweep of a DCGS2 step, per right-hand side j with coefficients
// cf_j = coef + j*sC = [s (k) | t (k) | 1/omega | tau], k = k1 + k2, over the
// basis [W1, W2]:
//   u <- (u - W s) / omega                      (delayed reorthogonalisation)
//   z <- (z - W t - u_new tau) / omega          (first pass of the new vector)
// (z NULL: the u update alone).  A thread takes BDU_E entries 256 apart,
// which belong to the same right-hand side when R divides 256, so each
// coefficient it loads serves BDU_E entries.
#define BDU_E 4
__global__ __launch_bounds__(256) void bdcgs_update(size_t n, int R, int k1,
                                                    const double* __restrict__ W1, int k2,
                                                    const double* __restrict__ W2, size_t ldw,
                                                    const double* __restrict__ coef, size_t sC,
                                                    double* __restrict__ u, double* __restrict__ z) {
  const size_t N = n * R;
  const int k = k1 + k2;
  for (size_t base = (size_t)blockIdx.x * (256 * BDU_E); base < N;
       base += (size_t)gridDim.x * (256 * BDU_E)) {
    const size_t t0 = base + threadIdx.x;
    const double* cf = &coef[(size_t)(t0 % R) * sC];
    double as[BDU_E], at[BDU_E];
    bool ok[BDU_E];
    size_t te[BDU_E];
#pragma unroll
    for (int e = 0; e < BDU_E; e++) {
      as[e] = 0.0;
      at[e] = 0.0;
      ok[e] = t0 + (size_t)e * 256 < N;
      // entries past the end re-read the last one and are not written: the
      // loads stay unconditional, so the unrolled passes issue them together
      te[e] = ok[e] ? t0 + (size_t)e * 256 : N - 1;
    }
    // four basis vectors per pass keep 16 loads per thread in flight (small
    // systems have too few waves to cover the memory latency otherwise)
#pragma unroll 4
    for (int v = 0; v < k; v++) {
      const double* w = v < k1 ? W1 + (size_t)v * ldw : W2 + (size_t)(v - k1) * ldw;
      const double cs = cf[v], ct = cf[k + v];
#pragma unroll
      for (int e = 0; e < BDU_E; e++) {
        const double x = w[te[e]];
        as[e] += x * cs;
        at[e] += x * ct;
      }
    }
    const double io = cf[2 * k], tau = cf[2 * k + 1];
#pragma unroll
    for (int e = 0; e < BDU_E; e++)
      if (ok[e]) {
        const size_t t = t0 + (size_t)e * 256;
        const double un = (u[t] - as[e]) * io;
        u[t] = un;
        if (z) z[t] = (z[t] - at[e] - un * tau) * io;
      }
  }
}

// General R (not dividing 256): one entry per thread
__global__ __launch_bounds__(256) void bdcgs_update1(size_t n, int R, int k1,
                                                     const double* __restrict__ W1, int k2,
                                                     const double* __restrict__ W2, size_t ldw,
                                                     const double* __restrict__ coef, size_t sC,
                                                     double* __restrict__ u,
                                                     double* __restrict__ z) {
  const size_t N = n * R;
  const int k = k1 + k2;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (size_t)gridDim.x * blockDim.x) {
    const double* cf = &coef[(size_t)(t % R) * sC];
    double as = 0.0, at = 0.0;
    for (int v = 0; v < k; v++) {
      const double w = v < k1 ? W1[(size_t)v * ldw + t] : W2[(size_t)(v - k1) * ldw + t];
      as += w * cf[v];
      at += w * cf[k + v];
    }
    const double io = cf[2 * k], tau = cf[2 * k + 1];
    const double un = (u[t] - as) * io;
    u[t] = un;
    if (z) z[t] = (z[t] - at - un * tau) * io;
  }
}

// ---------------------------------------------------------------------------
// The small per-step computation of a DCGS2 Arnoldi step of the recycling
// (GCRO-DR) cycle, on the device so that the host does not wait for each
// step.  One wavefront per right-hand side j.  Inputs: the dot lists of the
// step, a = Q^T u_i (p = Kd + i entries, then |u_i|^2) and b = Q^T z (then
// u_i^T z), Q = [C (Kd vectors), V_0..V_{i-1}].  State of j (ls + j*sLS,
// ld = rows of the least-squares matrix, nc = ld - 1 columns):
//   Hb, Hr [ld x nc column-major]: the least-squares matrix, plain / rotated
//   g, c [ld]: its right-hand side, rotated / plain
//   rc, rs [ld]: rotation q acts on rows (q, q + 1)
//   pend [ld]: first-pass coefficients of the column being finished
// and ist[4j..4j+3] = converged, failed, finished columns (jend), columns in
// the least-squares matrix.  Per step:
//   omega = sqrt(|u_i|^2 - |s|^2);
//   i > 0: column kd + i - 1 = pend + s (least-squares row order: the
//          kd_j recycled rows, then V), subdiagonal omega; Givens update;
//          converged when |g| <= tgt_j or omega = 0, failed when not finite;
//   tau = (u_i^T z - s . t) / omega;
//   pend = ([t; tau] - sum_{a < i} Hb(:, kd + a) s_V[a]) / omega;
//   cf = [s | t | 1/omega | tau] for the update sweep (zeros once j stopped).
// flags (host-visible, may be NULL) receives ist after the step.
// ---------------------------------------------------------------------------
#define BDC_LDMAX 512
__global__ __launch_bounds__(64) void bdcgs_coeffs(int i, int Kd, int ld, const int* __restrict__ kdv,
                                                   const double* __restrict__ tgtv,
                                                   const double* __restrict__ dA,
                                                   const double* __restrict__ dB, size_t S,
                                                   double* __restrict__ ls, size_t sLS,
                                                   int* __restrict__ ist, double* __restrict__ cf,
                                                   size_t sC, int* __restrict__ flags) {
  __shared__ double col[BDC_LDMAX];
  __shared__ int stop;
  const int j = blockIdx.x, lane = threadIdx.x;
  const int p = Kd + i, nc = ld - 1;
  const double* a = dA + (size_t)j * S;
  const double* b = dB + (size_t)j * S;
  double* cj = cf + (size_t)j * sC;
  int* st = ist + 4 * j;
  double* Hb = ls + (size_t)j * sLS;
  double* Hr = Hb + (size_t)ld * nc;
  double* g = Hr + (size_t)ld * nc;
  double* cc = g + ld;
  double* rc = cc + ld;
  double* rs = rc + ld;
  double* pend = rs + ld;
  bool done = st[0] != 0;
  if (!done) {
    double ss = 0.0, sb = 0.0;
    for (int r = lane; r < p; r += 64) {
      ss += a[r] * a[r];
      sb += a[r] * b[r];
    }
    ss = wave_sum(ss);
    sb = wave_sum(sb);
    const double w2 = a[p] - ss;
    const double om = (w2 > 0.0 && isfinite(w2)) ? sqrt(w2) : 0.0;
    const int kd = kdv[j];
    if (i > 0) {
      const int jj = kd + i - 1;  // the column finished at this step
      for (int r = lane; r < ld; r += 64) {
        double v = 0.0;
        if (r < kd + i)
          v = pend[r] + (r < kd ? a[r] : a[Kd + (r - kd)]);
        else if (r == kd + i)
          v = om;
        Hb[(size_t)jj * ld + r] = v;
        col[r] = v;
      }
      __syncthreads();
      if (lane == 0) {
        for (int q = 0; q < jj; q++) {
          const double x = col[q], y = col[q + 1];
          col[q] = rc[q] * x + rs[q] * y;
          col[q + 1] = -rs[q] * x + rc[q] * y;
        }
        const double x = col[jj], y = col[jj + 1];
        const double h = hypot(x, y);
        const double cs = h == 0.0 ? 1.0 : x / h, sn = h == 0.0 ? 0.0 : y / h;
        col[jj] = cs * x + sn * y;
        col[jj + 1] = -sn * x + cs * y;
        const double g0 = g[jj], g1 = g[jj + 1];
        g[jj] = cs * g0 + sn * g1;
        g[jj + 1] = -sn * g0 + cs * g1;
        rc[jj] = cs;
        rs[jj] = sn;
        st[3] = jj + 1;
        const double est = fabs(g[jj + 1]);
        stop = 0;
        if (!isfinite(est)) {
          st[0] = 1;
          st[1] = 1;
          st[2] = i - 1;
          stop = 1;
        } else if (est <= tgtv[j] || om == 0.0) {
          st[0] = 1;
          st[2] = i;
        }
      }
      __syncthreads();
      for (int r = lane; r < ld; r += 64) Hr[(size_t)jj * ld + r] = col[r];
      done = stop != 0;
    } else if (lane == 0) {
      cc[kd] = om;  // u_0 = C s_C + omega v_0
      g[kd] = om;
    }
    if (!done) {
      const double io = om > 0.0 ? 1.0 / om : 0.0;
      const double tau = (b[p] - sb) * io;
      for (int r = lane; r <= kd + i; r += 64) {
        double v = r < kd + i ? (r < kd ? b[r] : b[Kd + (r - kd)]) : tau;
        for (int q = 0; q < i; q++)
          if (r <= kd + q + 1) v -= Hb[(size_t)(kd + q) * ld + r] * a[Kd + q];
        pend[r] = v * io;
      }
      for (int r = lane; r < p; r += 64) {
        cj[r] = a[r];
        cj[p + r] = b[r];
      }
      if (lane == 0) {
        cj[2 * p] = io;
        cj[2 * p + 1] = tau;
      }
    }
  }
  if (done)
    for (int r = lane; r < 2 * p + 2; r += 64) cj[r] = 0.0;
  __syncthreads();
  if (flags && lane < 4) flags[4 * j + lane] = st[lane];
}

// y[j*sy + v] += x[j*sx + v], v < k (the second Gram-Schmidt pass into H)
__global__ void badd(int R, int k, const double* __restrict__ x, size_t sx, double* __restrict__ y,
                     size_t sy) {
  int t = threadIdx.x + blockIdx.x * blockDim.x;
  if (t >= R * k) return;
  int j = t / k, v = t % k;
  y[(size_t)j * sy + v] += x[(size_t)j * sx + v];
}

// x[i*R + j] *= s[j]
__global__ void bscale(size_t n, int R, const double* __restrict__ s, double* __restrict__ x) {
  const size_t N = n * R;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (size_t)gridDim.x * blockDim.x)
    x[t] *= s[t % R];
}

// nrm[j*sN] = sqrt(sq[j*sS]), inv[j] = 1 / nrm (0 for a zero norm)
__global__ void bsqrt_inv(int R, const double* __restrict__ sq, size_t sS, double* __restrict__ nrm,
                          size_t sN, double* __restrict__ inv) {
  int j = threadIdx.x + blockIdx.x * blockDim.x;
  if (j >= R) return;
  double v = sqrt(sq[(size_t)j * sS]);
  if (nrm) nrm[(size_t)j * sN] = v;
  inv[j] = v > 0.0 ? 1.0 / v : 0.0;
}

// Givens update of column i of every H_j (the recurrence of GMRES::solve,
// src/bpmat/KSM.cpp:862-895), one thread per right-hand side.
__global__ void bgivens(int R, int i, int m, double* __restrict__ H, size_t sH,
                        double* __restrict__ Qc, double* __restrict__ Qs, size_t sQ,
                        double* __restrict__ res, size_t sRes, double* __restrict__ st) {
  int j = threadIdx.x + blockIdx.x * blockDim.x;
  if (j >= R) return;
  double* h = &H[(size_t)j * sH + (size_t)i * (m + 1)];
  double* qc = &Qc[(size_t)j * sQ];
  double* qs = &Qs[(size_t)j * sQ];
  double* rs = &res[(size_t)j * sRes];
  for (int k = 0; k < i; k++) {
    double h1 = h[k], h2 = h[k + 1];
    h[k] = h1 * qc[k] + h2 * qs[k];
    h[k + 1] = -h1 * qs[k] + h2 * qc[k];
  }
  double h1 = h[i], h2 = h[i + 1];
  double sq = sqrt(h1 * h1 + h2 * h2);
  double c = sq != 0.0 ? h1 / sq : 1.0, s = sq != 0.0 ? h2 / sq : 0.0;
  qc[i] = c;
  qs[i] = s;
  h[i] = h1 * c + h2 * s;
  h[i + 1] = 0.0;
  rs[i + 1] = -rs[i] * s;
  rs[i] = rs[i] * c;
  st[j] = fabs(rs[i + 1]);
}

// y_j = R_j^{-1} res_j over the first cnt[j] columns (in place in res)
__global__ void bbacksolve(int R, int m, const int* __restrict__ cnt, const double* __restrict__ H,
                           size_t sH, double* __restrict__ res, size_t sRes) {
  int j = threadIdx.x + blockIdx.x * blockDim.x;
  if (j >= R) return;
  const double* h = &H[(size_t)j * sH];
  double* rs = &res[(size_t)j * sRes];
  int nit = cnt[j];
  for (int i = nit - 1; i >= 0; i--) {
    double s = rs[i];
    for (int k = i + 1; k < nit; k++) s -= h[(size_t)k * (m + 1) + i] * rs[k];
    rs[i] = s / h[(size_t)i * (m + 1) + i];
  }
}

// y = D^{-1} x per b x b block (b = 3 or 6) for interleaved block vectors
// (block-Jacobi)
template <int B, typename TD>
__global__ void bdiag_apply(int nrows, int R, const TD* __restrict__ Dinv,
                            const double* __restrict__ x, double* __restrict__ y) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nrows * R) return;
  int node = (int)(t / R), j = (int)(t % R);
  const TD* d = &Dinv[(size_t)B * B * node];
  double xc[B];
#pragma unroll
  for (int c = 0; c < B; c++) xc[c] = x[((size_t)B * node + c) * R + j];
#pragma unroll
  for (int r = 0; r < B; r++) {
    double v = (double)d[B * r] * xc[0];
#pragma unroll
    for (int c = 1; c < B; c++) v += (double)d[B * r + c] * xc[c];
    y[((size_t)B * node + r) * R + j] = v;
  }
}

// One step of the Chebyshev iteration for A z = x preconditioned by the
// block-Jacobi inverse (Saad, Iterative Methods for Sparse Linear Systems,
// 2nd ed., Alg. 12.1), over interleaved block vectors: with w = D^{-1}
// (x - A z) (first step: w = D^{-1} x, z = 0), d = c1 d + c2 w and z += d.
template <int B, typename TD>
__global__ void bcheb_step(int nrows, int R, const TD* __restrict__ Dinv,
                           const double* __restrict__ x, const double* __restrict__ Az,
                           double* __restrict__ d, double* __restrict__ z, double c1, double c2,
                           int first, float* __restrict__ z32) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nrows * R) return;
  const int node = (int)(t / R), j = (int)(t % R);
  const TD* D = &Dinv[(size_t)B * B * node];
  double rc[B];
#pragma unroll
  for (int c = 0; c < B; c++) {
    const size_t o = ((size_t)B * node + c) * R + j;
    rc[c] = first ? x[o] : x[o] - Az[o];
  }
#pragma unroll
  for (int r = 0; r < B; r++) {
    double w = (double)D[B * r] * rc[0];
#pragma unroll
    for (int c = 1; c < B; c++) w += (double)D[B * r + c] * rc[c];
    const size_t o = ((size_t)B * node + r) * R + j;
    const double dn = first ? c2 * w : c1 * d[o] + c2 * w;
    d[o] = dn;
    const double zn = first ? dn : z[o] + dn;
    z[o] = zn;
    if (z32) z32[o] = (float)zn;
  }
}

// Multi right-hand-side 3x3 SpMM, Y = A X, over interleaved block vectors
// (X[(3*node + c)*R + j]), FP64 (TA = double) or FP32 (TA = float) values,
// FP64 vectors and accumulation.  One wavefront per block row; the 64 lanes
// are G = 64 / RP slots of RP >= R lanes: lane (slot g, right-hand side j)
// takes blocks g, g + G, ... of the row, so each step of the wave reads G
// consecutive blocks (coalesced) and, per block, the 3R contiguous entries of
// X at its column node; the three row sums are reduced over the slots.
template <typename TA, int RP>
__global__ __launch_bounds__(256) void bsr3_spmm_slots(int nrows, int R,
                                                       const int* __restrict__ rowp,
                                                       const int* __restrict__ cols,
                                                       const TA* __restrict__ A,
                                                       const double* __restrict__ X,
                                                       double* __restrict__ Y) {
  constexpr int G = 64 / RP;
  const int lane = threadIdx.x & 63;
  const int j = lane % RP, g = lane / RP;
  const bool okj = j < R;
  // XCD-aware rows: workgroups go round-robin to the 8 XCDs, so XCD x
  // (blockIdx % 8) takes the contiguous eighth x of the rows and its L2 holds
  // the X entries its neighbouring rows share
  const int wid = threadIdx.x >> 6;
  const int xcd = blockIdx.x & 7, nbx = gridDim.x >> 3;
  const int r0 = (int)((long)nrows * xcd / 8), r1 = (int)((long)nrows * (xcd + 1) / 8);
  for (int row = r0 + (blockIdx.x >> 3) * 4 + wid; row < r1; row += nbx * 4) {
    const int rs = rowp[row], re = rowp[row + 1];
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    if (okj) {
      for (int s = rs + g; s < re; s += G) {
        const int col = cols[s];
        const TA* a = &A[(size_t)9 * s];
        const double* xx = &X[(size_t)3 * col * R + j];
        const double x0 = xx[0], x1 = xx[R], x2 = xx[2 * R];
        y0 += (double)a[0] * x0 + (double)a[1] * x1 + (double)a[2] * x2;
        y1 += (double)a[3] * x0 + (double)a[4] * x1 + (double)a[5] * x2;
        y2 += (double)a[6] * x0 + (double)a[7] * x1 + (double)a[8] * x2;
      }
    }
#pragma unroll
    for (int off = RP; off < 64; off <<= 1) {
      y0 += __shfl_xor(y0, off, 64);
      y1 += __shfl_xor(y1, off, 64);
      y2 += __shfl_xor(y2, off, 64);
    }
    if (g == 0 && okj) {
      double* y = &Y[(size_t)3 * row * R + j];
      y[0] = y0;
      y[R] = y1;
      y[2 * R] = y2;
    }
  }
}

// The same product for R divisible by 4: each lane takes four consecutive
// right-hand sides (16-byte loads of X, 12 sums), so NL = R/4 lanes cover a
// block and a wave works on G = 64/NL blocks at once.
typedef double spmm_dbl2 __attribute__((ext_vector_type(2)));
typedef float spmm_flt4 __attribute__((ext_vector_type(4)));
template <typename TA, int NL, int SPU = 1, int NTA = 1>
__global__ __launch_bounds__(256) void bsr3_spmm_quad(int nrows, int R,
                                                      const int* __restrict__ rowp,
                                                      const int* __restrict__ cols,
                                                      const TA* __restrict__ A,
                                                      const double* __restrict__ X,
                                                      double* __restrict__ Y) {
  constexpr int G = 64 / NL;
  const int lane = threadIdx.x & 63;
  const int jq = lane % NL, g = lane / NL;  // right-hand sides 4 jq .. 4 jq + 3
  const bool okq = 4 * jq < R;               // NL is R/4 rounded up to a power of two
  const int wid = threadIdx.x >> 6;
  const int xcd = blockIdx.x & 7, nbx = gridDim.x >> 3;
  const int r0 = (int)((long)nrows * xcd / 8), r1 = (int)((long)nrows * (xcd + 1) / 8);
  for (int row = r0 + (blockIdx.x >> 3) * 4 + wid; row < r1; row += nbx * 4) {
    const int rs = rowp[row], re = rowp[row + 1];
    double y[3][4];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int q = 0; q < 4; q++) y[r][q] = 0.0;
    // SPU blocks per lane per trip, all loads issued before the arithmetic
    // (a block past the row end re-reads a valid block and is masked)
    for (int s0 = rs + g; okq && s0 < re; s0 += SPU * G) {
      int sb[SPU];
      bool use[SPU];
      int col[SPU];
#pragma unroll
      for (int u = 0; u < SPU; u++) {
        use[u] = s0 + u * G < re;
        sb[u] = use[u] ? s0 + u * G : s0;
        col[u] = cols[sb[u]];
      }
      TA av[SPU][9];
#pragma unroll
      for (int u = 0; u < SPU; u++)
#pragma unroll
        for (int e = 0; e < 9; e++)
          av[u][e] = NTA ? __builtin_nontemporal_load(&A[(size_t)9 * sb[u] + e])
                         : A[(size_t)9 * sb[u] + e];
      double xv[SPU][3][4];
#pragma unroll
      for (int u = 0; u < SPU; u++) {
        const double* xb = &X[(size_t)3 * col[u] * R + 4 * jq];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const spmm_dbl2 lo = *(const spmm_dbl2*)(xb + (size_t)c * R);
          const spmm_dbl2 hi = *(const spmm_dbl2*)(xb + (size_t)c * R + 2);
          xv[u][c][0] = lo.x;
          xv[u][c][1] = lo.y;
          xv[u][c][2] = hi.x;
          xv[u][c][3] = hi.y;
        }
      }
#pragma unroll
      for (int u = 0; u < SPU; u++) {
        if (!use[u]) continue;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int q = 0; q < 4; q++)
            y[r][q] += (double)av[u][3 * r] * xv[u][0][q] + (double)av[u][3 * r + 1] * xv[u][1][q] +
                       (double)av[u][3 * r + 2] * xv[u][2][q];
      }
    }
#pragma unroll
    for (int off = NL; off < 64; off <<= 1)
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) y[r][q] += __shfl_xor(y[r][q], off, 64);
    if (g == 0 && okq) {
#pragma unroll
      for (int r = 0; r < 3; r++) {
        double* yo = &Y[((size_t)3 * row + r) * R + 4 * jq];
        *(spmm_dbl2*)yo = spmm_dbl2{y[r][0], y[r][1]};
        *(spmm_dbl2*)(yo + 2) = spmm_dbl2{y[r][2], y[r][3]};
      }
    }
  }
}

// Forms and launch variants of the R right-hand-side product.  R even and
// at most 12: two right-hand sides per lane (bsr3_spmm_pair; hex27 40^3,
// FP32 values: 4 RHS 0.46 ms against 0.77 ms four per lane, 8 RHS 0.62
// against 0.80), variants 0 (1 block per lane per trip) and 1 (2).  R a
// multiple of 4 from 16: four per lane (bsr3_spmm_quad), variants by blocks
// per lane per trip SPU and non-temporal value loads NTA: 0 (1, on), 1 (2,
// on), 2 (1, off), 3 (2, off), 4 (4, on).  Other R: one lane per right-hand
// side (bsr3_spmm_slots), one variant.  The variants of a form give the same
// sums (each lane still adds its blocks g, g + G, ... in order).
static constexpr int kSpmmQuadVariants = 5, kSpmmPairVariants = 2;
static inline bool spmm_pair_form(int R) { return R % 2 == 0 && R <= 12; }
static inline bool spmm_quad_form(int R) { return R % 4 == 0 && R >= 16; }
static inline int spmm_nvariants(int R) {
  return spmm_pair_form(R) ? kSpmmPairVariants : spmm_quad_form(R) ? kSpmmQuadVariants : 1;
}
template <typename TA>
static void spmm_pair_launch(int spu, int grid, int nrows, int R, const int* rowp, const int* cols,
                             const TA* A, const double* X, double* Y, hipStream_t stream);

// The same product with two right-hand sides per lane and NL = R / 2 lanes
// per block (R even): each component's R values of X at a column node are
// one R x 8-byte run read by the NL lanes of a block with 16-byte loads (a
// 128-byte line for R = 16), and a block's 9 values come in with three
// vector loads (two of four values, one of one) that its NL lanes share, so
// a wave step of G = 64 / NL blocks touches about 4 cache lines per block
// where the four-per-lane form touches about 9.  SPU blocks per lane per
// trip (loads first; a block past the row end is masked).
typedef float spmm_flt4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double spmm_dbl2u __attribute__((ext_vector_type(2), aligned(8)));
template <typename TA>
__device__ __forceinline__ void spmm_load_block(const TA* __restrict__ a, TA v[9]) {
  if constexpr (sizeof(TA) == 4) {
    const spmm_flt4u p0 = *(const spmm_flt4u*)a;
    const spmm_flt4u p1 = *(const spmm_flt4u*)(a + 4);
    v[0] = p0.x, v[1] = p0.y, v[2] = p0.z, v[3] = p0.w;
    v[4] = p1.x, v[5] = p1.y, v[6] = p1.z, v[7] = p1.w;
    v[8] = a[8];
  } else {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const spmm_dbl2u p = *(const spmm_dbl2u*)(a + 2 * q);
      v[2 * q] = p.x;
      v[2 * q + 1] = p.y;
    }
    v[8] = a[8];
  }
}
template <typename TA, int NL, int SPU>
__global__ __launch_bounds__(256) void bsr3_spmm_pair(int nrows, int R,
                                                      const int* __restrict__ rowp,
                                                      const int* __restrict__ cols,
                                                      const TA* __restrict__ A,
                                                      const double* __restrict__ X,
                                                      double* __restrict__ Y) {
  constexpr int G = 64 / NL;
  const int lane = threadIdx.x & 63;
  const int jp = lane % NL, g = lane / NL;  // right-hand sides 2 jp, 2 jp + 1
  const bool okp = 2 * jp < R;
  const int wid = threadIdx.x >> 6;
  const int xcd = blockIdx.x & 7, nbx = gridDim.x >> 3;
  const int r0 = (int)((long)nrows * xcd / 8), r1 = (int)((long)nrows * (xcd + 1) / 8);
  for (int row = r0 + (blockIdx.x >> 3) * 4 + wid; row < r1; row += nbx * 4) {
    const int rs = rowp[row], re = rowp[row + 1];
    double y[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int s0 = rs + g; okp && s0 < re; s0 += SPU * G) {
      int sb[SPU];
      bool use[SPU];
#pragma unroll
      for (int u = 0; u < SPU; u++) {
        use[u] = s0 + u * G < re;
        sb[u] = use[u] ? s0 + u * G : s0;
      }
      TA av[SPU][9];
      spmm_dbl2 xv[SPU][3];
#pragma unroll
      for (int u = 0; u < SPU; u++) {
        const int col = cols[sb[u]];
        spmm_load_block<TA>(&A[(size_t)9 * sb[u]], av[u]);
        const double* xb = &X[(size_t)3 * col * R + 2 * jp];
#pragma unroll
        for (int c = 0; c < 3; c++) xv[u][c] = *(const spmm_dbl2*)(xb + (size_t)c * R);
      }
#pragma unroll
      for (int u = 0; u < SPU; u++) {
        if (!use[u]) continue;
#pragma unroll
        for (int r = 0; r < 3; r++) {
          y[r][0] += (double)av[u][3 * r] * xv[u][0].x + (double)av[u][3 * r + 1] * xv[u][1].x +
                     (double)av[u][3 * r + 2] * xv[u][2].x;
          y[r][1] += (double)av[u][3 * r] * xv[u][0].y + (double)av[u][3 * r + 1] * xv[u][1].y +
                     (double)av[u][3 * r + 2] * xv[u][2].y;
        }
      }
    }
#pragma unroll
    for (int off = NL; off < 64; off <<= 1)
#pragma unroll
      for (int r = 0; r < 3; r++) {
        y[r][0] += __shfl_xor(y[r][0], off, 64);
        y[r][1] += __shfl_xor(y[r][1], off, 64);
      }
    if (g == 0 && okp) {
#pragma unroll
      for (int r = 0; r < 3; r++)
        *(spmm_dbl2*)&Y[((size_t)3 * row + r) * R + 2 * jp] = spmm_dbl2{y[r][0], y[r][1]};
    }
  }
}

// One Chebyshev step of the batched preconditioner fused into the product:
// the quad-form product a = A z32 (FP32 z), then per row and right-hand side
// w = D^{-1} (x - a), d = c1 d + c2 w, z = z + d (FP64) and z32out = (float) z.
// z32 is a copy of z that the products of the polynomial read (half the
// gather traffic of FP64 z); z32out must differ from z32in (other rows still
// read it).  R = 4 NL, 16 <= R; Dinv in FP32 (the values' precision).
// mode 1 (three-term form): d holds the iterate before z and receives
// (1 + c1) z - c1 d + c2 w, z is only read; mode 2 (the step into that
// form): d holds the last d and receives z + c1 d + c2 w.  Both read z and
// d and write d: one FP64 access per entry fewer than mode 0.
template <typename TA, int NL, int SPU = 1, int NTA = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void bsr3_spmm_quad_cheb(
    int nrows, int R, const int* __restrict__ rowp, const int* __restrict__ cols,
    const TA* __restrict__ A, const float* __restrict__ X, const double* __restrict__ xr,
    const float* __restrict__ Dinv, double* __restrict__ d, double* __restrict__ z,
    float* __restrict__ Xout, double c1, double c2, int mode) {
  constexpr int G = 64 / NL;
  const int lane = threadIdx.x & 63;
  const int jq = lane % NL, g = lane / NL;
  const bool okq = 4 * jq < R;
  const int wid = threadIdx.x >> 6;
  const int xcd = blockIdx.x & 7, nbx = gridDim.x >> 3;
  const int r0 = (int)((long)nrows * xcd / 8), r1 = (int)((long)nrows * (xcd + 1) / 8);
  for (int row = r0 + (blockIdx.x >> 3) * 4 + wid; row < r1; row += nbx * 4) {
    const int rs = rowp[row], re = rowp[row + 1];
    double y[3][4];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int q = 0; q < 4; q++) y[r][q] = 0.0;
    // software pipeline over the wave's steps of G blocks: the next step's
    // column and values are loaded after this step's X gathers, so a step
    // waits on one memory latency (the gathers), not on the column load
    // and then the gathers
    if (okq && rs + g < re) {
      int colN = cols[rs + g];
      TA avN[9];
#pragma unroll
      for (int e = 0; e < 9; e++)
        avN[e] = NTA ? __builtin_nontemporal_load(&A[(size_t)9 * (rs + g) + e])
                     : A[(size_t)9 * (rs + g) + e];
      for (int s0 = rs + g; s0 < re; s0 += G) {
        const int col = colN;
        TA av[9];
#pragma unroll
        for (int e = 0; e < 9; e++) av[e] = avN[e];
        double xv[3][4];
        const float* xb = &X[(size_t)3 * col * R + 4 * jq];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const spmm_flt4 v = *(const spmm_flt4*)(xb + (size_t)c * R);
          xv[c][0] = v.x;
          xv[c][1] = v.y;
          xv[c][2] = v.z;
          xv[c][3] = v.w;
        }
        const int s1 = s0 + G;
        if (s1 < re) {
          colN = cols[s1];
#pragma unroll
          for (int e = 0; e < 9; e++)
            avN[e] = NTA ? __builtin_nontemporal_load(&A[(size_t)9 * s1 + e])
                         : A[(size_t)9 * s1 + e];
        }
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int q = 0; q < 4; q++)
            y[r][q] += (double)av[3 * r] * xv[0][q] + (double)av[3 * r + 1] * xv[1][q] +
                       (double)av[3 * r + 2] * xv[2][q];
      }
    }
    for (int off = NL; off < 64; off <<= 1)
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) y[r][q] += __shfl_xor(y[r][q], off, 64);
    // every lane holds the row sums of its right-hand sides; the G lanes of
    // each jq share its 12 (block row r, right-hand side q) outputs
    for (int k = g; okq && k < 12; k += G) {
      const int r = k >> 2, q = k & 3;
      double yq[3];
#pragma unroll
      for (int c = 0; c < 3; c++)
        yq[c] = q == 0 ? y[c][0] : (q == 1 ? y[c][1] : (q == 2 ? y[c][2] : y[c][3]));
      const size_t base = (size_t)3 * row * R + 4 * jq + q;
      const double rc0 = xr[base] - yq[0], rc1 = xr[base + R] - yq[1],
                   rc2 = xr[base + 2 * (size_t)R] - yq[2];
      const float* D = &Dinv[(size_t)9 * row + 3 * r];
      double w = (double)D[0] * rc0;
      w += (double)D[1] * rc1;
      w += (double)D[2] * rc2;
      const size_t o = base + (size_t)r * R;
      double zn;
      if (mode == 0) {
        const double dn = c1 * d[o] + c2 * w;
        d[o] = dn;
        zn = z[o] + dn;
        z[o] = zn;
      } else {
        const double zo = z[o], dp = d[o];
        zn = mode == 1 ? (1.0 + c1) * zo - c1 * dp + c2 * w : zo + (c1 * dp + c2 * w);
        d[o] = zn;
      }
      Xout[o] = (float)zn;
    }
  }
}

template <typename TA, int NLV>
static void spmm_quad_launch(int variant, int grid, int nrows, int R, const int* rowp,
                             const int* cols, const TA* A, const double* X, double* Y,
                             hipStream_t stream) {
#define SPMM_QUAD(SPU, NTA)                                                                      \
  hipLaunchKernelGGL((bsr3_spmm_quad<TA, NLV, SPU, NTA>), dim3(grid), dim3(256), 0, stream, nrows, \
                     R, rowp, cols, A, X, Y)
  switch (variant) {
    case 1: SPMM_QUAD(2, 1); break;
    case 2: SPMM_QUAD(1, 0); break;
    case 3: SPMM_QUAD(2, 0); break;
    case 4: SPMM_QUAD(4, 1); break;
    default: SPMM_QUAD(1, 1); break;
  }
#undef SPMM_QUAD
}

template <typename TA>
static void spmm_pair_launch(int spu, int grid, int nrows, int R, const int* rowp, const int* cols,
                             const TA* A, const double* X, double* Y, hipStream_t stream) {
#define SPMM_PAIR(NLV)                                                                          \
  do {                                                                                          \
    if (spu == 1)                                                                               \
      hipLaunchKernelGGL((bsr3_spmm_pair<TA, NLV, 1>), dim3(grid), dim3(256), 0, stream, nrows, \
                         R, rowp, cols, A, X, Y);                                               \
    else                                                                                        \
      hipLaunchKernelGGL((bsr3_spmm_pair<TA, NLV, 2>), dim3(grid), dim3(256), 0, stream, nrows, \
                         R, rowp, cols, A, X, Y);                                               \
  } while (0)
  switch (pow2_ceil(R / 2)) {
    case 1: SPMM_PAIR(1); break;
    case 2: SPMM_PAIR(2); break;
    case 4: SPMM_PAIR(4); break;
    default: SPMM_PAIR(8); break;
  }
#undef SPMM_PAIR
}

template <typename TA>
static int spmm_slots_launch(int variant, int nrows, int R, const int* rowp, const int* cols,
                             const TA* A, const double* X, double* Y, hipStream_t stream) {
  if (R > 64) return TACS_HIP_ERR_ARG;
  if (nrows <= 0 || R <= 0) return 0;
  if (variant < 0 || variant >= spmm_nvariants(R)) return TACS_HIP_ERR_ARG;
  // a multiple of 8 workgroups (one share per XCD)
  const int grid = 8 * tacs_grid_for((nrows + 7) / 8, 4, 8192);
  if (spmm_pair_form(R)) {
    spmm_pair_launch<TA>(variant + 1, grid, nrows, R, rowp, cols, A, X, Y, stream);
    return (int)hipGetLastError();
  }
  if (spmm_quad_form(R)) {
    // 16-byte aligned X rows: four right-hand sides per lane
    switch (pow2_ceil(R / 4)) {
      case 1: spmm_quad_launch<TA, 1>(variant, grid, nrows, R, rowp, cols, A, X, Y, stream); break;
      case 2: spmm_quad_launch<TA, 2>(variant, grid, nrows, R, rowp, cols, A, X, Y, stream); break;
      case 4: spmm_quad_launch<TA, 4>(variant, grid, nrows, R, rowp, cols, A, X, Y, stream); break;
      case 8: spmm_quad_launch<TA, 8>(variant, grid, nrows, R, rowp, cols, A, X, Y, stream); break;
      default: spmm_quad_launch<TA, 16>(variant, grid, nrows, R, rowp, cols, A, X, Y, stream); break;
    }
    return (int)hipGetLastError();
  }
#define SPMM_CASE(RPV)                                                                          \
  hipLaunchKernelGGL((bsr3_spmm_slots<TA, RPV>), dim3(grid), dim3(256), 0, stream, nrows, R, rowp, \
                     cols, A, X, Y)
  switch (pow2_ceil(R)) {
    case 1: SPMM_CASE(1); break;
    case 2: SPMM_CASE(2); break;
    case 4: SPMM_CASE(4); break;
    case 8: SPMM_CASE(8); break;
    case 16: SPMM_CASE(16); break;
    case 32: SPMM_CASE(32); break;
    default: SPMM_CASE(64); break;
  }
#undef SPMM_CASE
  return (int)hipGetLastError();
}

// Per right-hand side basis change: out_c(:, j) (+)= sum_{v<nin} W_v(:, j)
// P_j[v + c*ldp], c < nout <= BBC_MAXOUT, with P_j = P + j*sP (the recycled
// space update of the batched GCRO-DR).  A block of 1024 threads owns 1024
// consecutive entries; the coefficients of vch (<= 8, at most 64 KB) inputs
// for every right-hand side are staged in LDS (R x vch x nout), so each
// multiply reads its coefficient from LDS instead of a scattered global load.
#define BBC_MAXOUT 32
#define BBC_THREADS 1024
__global__ __launch_bounds__(BBC_THREADS) void bbasis_combine(size_t n, int R, int nin,
                                                              const double* __restrict__ W,
                                                              size_t ldw,
                                                              const double* __restrict__ P,
                                                              size_t sP, int ldp, int nout,
                                                              double* __restrict__ out, size_t ldo,
                                                              int add, int vch) {
  extern __shared__ double Ps[];  // [R][vch][nout], right-hand side stride js (odd: no bank conflicts)
  const int js = vch * nout + 1;
  const size_t N = n * R;
  for (size_t base = (size_t)blockIdx.x * BBC_THREADS; base < N;
       base += (size_t)gridDim.x * BBC_THREADS) {
    const size_t t = base + threadIdx.x;
    const bool ok = t < N;
    const int j = ok ? (int)(t % R) : 0;
    double acc[BBC_MAXOUT];
#pragma unroll
    for (int c = 0; c < BBC_MAXOUT; c++) acc[c] = 0.0;
    for (int v0 = 0; v0 < nin; v0 += vch) {
      const int vc = min(vch, nin - v0);
      __syncthreads();
      for (int e = threadIdx.x; e < R * vc * nout; e += BBC_THREADS) {
        const int jj = e / (vc * nout), r = e - jj * (vc * nout);
        const int v = r / nout, c = r - v * nout;
        Ps[(size_t)jj * js + v * nout + c] = P[(size_t)jj * sP + v0 + v + (size_t)c * ldp];
      }
      __syncthreads();
      if (ok) {
        for (int v = 0; v < vc; v++) {
          const double w = W[(size_t)(v0 + v) * ldw + t];
          const double* pp = &Ps[(size_t)j * js + v * nout];
#pragma unroll
          for (int c = 0; c < BBC_MAXOUT; c++)
            if (c < nout) acc[c] += w * pp[c];
        }
      }
    }
    if (ok) {
#pragma unroll
      for (int c = 0; c < BBC_MAXOUT; c++)
        if (c < nout) {
          double* o = &out[(size_t)c * ldo + t];
          *o = add ? *o + acc[c] : acc[c];
        }
    }
  }
}

// Per right-hand side tall-skinny products out_j = X_j^T Y_j (X_j: pc block
// vectors, Y_j: qc block vectors, interleaved), the chunk (v0.., b0..) of a
// p x q product.  A block stages BTS_TR rows of all pc + qc vectors for every
// right-hand side in LDS ([vector][row][j]); thread tiles are 4 x 4 entries
// (v, b) of one right-hand side, at most BTS_MAXT tiles per thread, so each
// LDS read serves 4 multiply-adds.  part[(blk * R + j) * pc * qc + e] holds a
// block's sums (e = b * pc + v).
#define BTS_TR 4
#define BTS_MAXT 2
#define BTS_BLOCKS 512
__global__ __launch_bounds__(256) void btsgemm_partial(size_t n, int R, int pc, const double* __restrict__ X,
                                                       int qc, const double* __restrict__ Y,
                                                       size_t ldw, double* __restrict__ part) {
  extern __shared__ double tile[];  // [(pc + qc)][BTS_TR][R]
  const int P4 = (pc + 3) / 4, Q4 = (qc + 3) / 4;
  const int ntiles = R * P4 * Q4;
  double acc[BTS_MAXT][16];
#pragma unroll
  for (int t = 0; t < BTS_MAXT; t++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[t][e] = 0.0;
  const int nv = pc + qc;
  for (size_t r0 = (size_t)blockIdx.x * BTS_TR; r0 < n; r0 += (size_t)gridDim.x * BTS_TR) {
    const int nr = (int)((n - r0) < BTS_TR ? (n - r0) : BTS_TR);
    __syncthreads();
    for (int t = threadIdx.x; t < nv * BTS_TR * R; t += 256) {
      const int v = t / (BTS_TR * R), rr = (t / R) % BTS_TR, j = t % R;
      double val = 0.0;
      if (rr < nr) {
        const size_t o = (r0 + rr) * R + j;
        val = v < pc ? X[(size_t)v * ldw + o] : Y[(size_t)(v - pc) * ldw + o];
      }
      tile[t] = val;
    }
    __syncthreads();
#pragma unroll
    for (int tt = 0; tt < BTS_MAXT; tt++) {
      const int e = threadIdx.x + 256 * tt;
      if (e < ntiles) {
        const int j = e % R, v4 = (e / R) % P4, b4 = e / (R * P4);
        for (int rr = 0; rr < BTS_TR; rr++) {
          double xv[4], yv[4];
#pragma unroll
          for (int a = 0; a < 4; a++) {
            const int v = min(4 * v4 + a, pc - 1), b = min(4 * b4 + a, qc - 1);
            xv[a] = tile[((size_t)v * BTS_TR + rr) * R + j];
            yv[a] = tile[((size_t)(pc + b) * BTS_TR + rr) * R + j];
          }
#pragma unroll
          for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[tt][a * 4 + c] += xv[a] * yv[c];
        }
      }
    }
  }
#pragma unroll
  for (int tt = 0; tt < BTS_MAXT; tt++) {
    const int e = threadIdx.x + 256 * tt;
    if (e < ntiles) {
      const int j = e % R, v4 = (e / R) % P4, b4 = e / (R * P4);
      double* o = &part[((size_t)blockIdx.x * R + j) * pc * qc];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const int v = 4 * v4 + a, b = 4 * b4 + c;
          if (v < pc && b < qc) o[(size_t)b * pc + v] = acc[tt][a * 4 + c];
        }
    }
  }
}

// out[j*sR + (b0 + b)*ldo + v0 + v] = sum over blocks
__global__ void btsgemm_finalize(int nb, int R, int pc, int qc, const double* __restrict__ part,
                                 double* __restrict__ out, size_t sR, int ldo, int v0, int b0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = pc * qc;
  if (t >= R * per) return;
  const int j = t / per, e = t % per, v = e % pc, b = e / pc;
  double s = 0.0;
  for (int blk = 0; blk < nb; blk++) s += part[((size_t)blk * R + j) * per + e];
  out[(size_t)j * sR + (size_t)(b0 + b) * ldo + v0 + v] = s;
}

// X[i*R + j] = x[i] for every j
__global__ void bbroadcast(size_t n, int R, const double* __restrict__ x, double* __restrict__ X) {
  const size_t N = n * R;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (size_t)gridDim.x * blockDim.x)
    X[t] = x[t / R];
}

// X[i*R + j] = x[i] (pack) or x[i] = X[i*R + j] (unpack)
__global__ void bpack(size_t n, int R, int j, const double* __restrict__ x,
                      double* __restrict__ X, int unpack) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    if (unpack)
      X[i] = x[i * R + j];
    else
      X[i * R + j] = x[i];
  }
}

static inline int bgrid(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  if (g < 1) g = 1;
  return (int)g;
}

// ---------------------------------------------------------------------------
// Segmented systems: the R diagonal blocks of a block-diagonal 3x3-block
// matrix, block j being the node range [off[j], off[j+1]) of one BCSR matrix
// (the meshes of a multi-mesh assembler, no coupling between them), solved as
// R independent systems by the batched solver.  A segmented block vector
// keeps local node i of system j at X[(3 i + c) R + j], i < maxn = the
// largest system; entries of local nodes beyond a system's size are zero.
// Every column of a row of block j must lie in block j (the host checks).
// ---------------------------------------------------------------------------
__global__ void seg_pack(int maxn, int R, const int* __restrict__ off, const double* __restrict__ x,
                         double* __restrict__ X) {
  const size_t total = (size_t)3 * maxn * R;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(t % R);
    const size_t ic = t / R;
    const int i = (int)(ic / 3), c = (int)(ic - 3 * (size_t)i);
    const int nj = off[j + 1] - off[j];
    X[t] = i < nj ? x[3 * ((size_t)off[j] + i) + c] : 0.0;
  }
}

__global__ void seg_unpack(int nnodes, int R, const int* __restrict__ off,
                           const double* __restrict__ X, double* __restrict__ x) {
  const size_t total = (size_t)3 * nnodes;
  for (size_t d = (size_t)blockIdx.x * blockDim.x + threadIdx.x; d < total;
       d += (size_t)gridDim.x * blockDim.x) {
    const int node = (int)(d / 3), c = (int)(d - 3 * (size_t)node);
    int lo = 0, hi = R - 1;  // the system whose range holds node
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= node) lo = mid;
      else hi = mid - 1;
    }
    x[d] = X[((size_t)3 * (node - off[lo]) + c) * R + lo];
  }
}

// Segmented block vector <-> whole-matrix vector through an LDS tile of SEG_T
// local nodes x R systems: the segmented side is read / written as one
// contiguous run of 3 SEG_T R values, the whole-matrix side as R runs of
// 3 SEG_T values (one per system).  to_union: x[3 (off[j] + i) + c] =
// X[(3 i + c) R + j]; from_union: X = the segmented form of x (zeros beyond
// a system's size).
#define SEG_T 32
template <bool TO_UNION>
__global__ __launch_bounds__(256) void seg_transpose(int maxn, int R, const int* __restrict__ off,
                                                      const double* __restrict__ src,
                                                      double* __restrict__ dst) {
  __shared__ double tile[3 * SEG_T * 64];
  const int i0 = blockIdx.x * SEG_T;
  const int ni = min(SEG_T, maxn - i0);
  const int nv = 3 * ni * R;  // values of the tile
  const size_t base = (size_t)3 * i0 * R;
  if (TO_UNION) {
    for (int t = threadIdx.x; t < nv; t += blockDim.x) tile[t] = src[base + t];
    __syncthreads();
    // system j's run: local values 3 i + c, i < min(ni, n_j - i0)
    for (int t = threadIdx.x; t < nv; t += blockDim.x) {
      const int j = t / (3 * ni), q = t - j * 3 * ni;  // q = 3 (i - i0) + c
      const int nj = off[j + 1] - off[j];
      if (i0 + q / 3 < nj) dst[3 * ((size_t)off[j] + i0) + q] = tile[q * R + j];
    }
  } else {
    for (int t = threadIdx.x; t < nv; t += blockDim.x) {
      const int j = t / (3 * ni), q = t - j * 3 * ni;
      const int nj = off[j + 1] - off[j];
      tile[q * R + j] = i0 + q / 3 < nj ? src[3 * ((size_t)off[j] + i0) + q] : 0.0;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nv; t += blockDim.x) dst[base + t] = tile[t];
  }
}

// y = D^{-1} x per local node (i, j): the inverse diagonal block of union
// node off[j] + i; zero beyond a system's size.
template <typename TD>
__global__ void seg_bdiag_apply(int maxn, int R, const int* __restrict__ off,
                                const TD* __restrict__ Dinv, const double* __restrict__ x,
                                double* __restrict__ y) {
  const size_t total = (size_t)maxn * R;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(t % R), i = (int)(t / R);
    const int nj = off[j + 1] - off[j];
    double* yo = &y[(size_t)3 * i * R + j];
    if (i >= nj) {
      yo[0] = 0.0;
      yo[R] = 0.0;
      yo[2 * R] = 0.0;
      continue;
    }
    const TD* d = &Dinv[9 * ((size_t)off[j] + i)];
    const double* xi = &x[(size_t)3 * i * R + j];
    const double x0 = xi[0], x1 = xi[R], x2 = xi[2 * R];
#pragma unroll
    for (int r = 0; r < 3; r++)
      yo[(size_t)r * R] = (double)d[3 * r] * x0 + (double)d[3 * r + 1] * x1 + (double)d[3 * r + 2] * x2;
  }
}

// Y = A X per local row (i, j): the blocks of union row off[j] + i in slot
// order (FP64 sums; TA double or float values), columns taken relative to
// off[j]; zero beyond a system's size.  Threads take j fastest, so the
// writes of the R systems of a local node are contiguous.
template <typename TA>
__global__ void seg_spmm3(int maxn, int R, const int* __restrict__ off,
                          const int* __restrict__ rowp, const int* __restrict__ cols,
                          const TA* __restrict__ A, const double* __restrict__ X,
                          double* __restrict__ Y) {
  const size_t total = (size_t)maxn * R;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(t % R), i = (int)(t / R);
    const int o = off[j], nj = off[j + 1] - o;
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    if (i < nj) {
      const int row = o + i;
      for (int s = rowp[row]; s < rowp[row + 1]; s++) {
        const int cl = cols[s] - o;
        const TA* a = &A[(size_t)9 * s];
        const double* xc = &X[(size_t)3 * cl * R + j];
        const double x0 = xc[0], x1 = xc[R], x2 = xc[2 * R];
        y0 += (double)a[0] * x0 + (double)a[1] * x1 + (double)a[2] * x2;
        y1 += (double)a[3] * x0 + (double)a[4] * x1 + (double)a[5] * x2;
        y2 += (double)a[6] * x0 + (double)a[7] * x1 + (double)a[8] * x2;
      }
    }
    double* yo = &Y[(size_t)3 * i * R + j];
    yo[0] = y0;
    yo[R] = y1;
    yo[2 * R] = y2;
  }
}

extern "C" {

size_t tacs_hip_batch_scratch_bytes(int R) {
  return (size_t)BRED_BLOCKS * BRED_KMAX * (R > 0 ? R : 1) * sizeof(double);
}

static int bmdot_launch(bool shared, size_t n, int R, int k, const double* W, size_t ldw,
                        const double* q, double* scratch, double* out, size_t sR, int accumulate,
                        hipStream_t stream) {
  if (R < 1 || R > 64 || k < 1) return TACS_HIP_ERR_ARG;
  for (int v0 = 0; v0 < k; v0 += BRED_KMAX) {
    int kk = k - v0 < BRED_KMAX ? k - v0 : BRED_KMAX;
    int nb = bred_blocks(n);
    if (shared)
      hipLaunchKernelGGL(bmdot_partial<true>, dim3(nb), dim3(BRED_THREADS), 0, stream, n, R, kk,
                         W + (size_t)v0 * ldw, ldw, q, scratch);
    else
      hipLaunchKernelGGL(bmdot_partial<false>, dim3(nb), dim3(BRED_THREADS), 0, stream, n, R, kk,
                         W + (size_t)v0 * ldw, ldw, q, scratch);
    int waves = kk * R;
    hipLaunchKernelGGL(bmdot_finalize, dim3((waves + 3) / 4), dim3(256), 0, stream, nb, kk, R,
                       scratch, out + v0, sR, accumulate);
  }
  return (int)hipGetLastError();
}

int tacs_hip_bmdot(size_t n, int R, int k, const double* W, size_t ldw, const double* q,
                   double* scratch, double* out, size_t sR, int accumulate, hipStream_t stream) {
  return bmdot_launch(false, n, R, k, W, ldw, q, scratch, out, sR, accumulate, stream);
}

int tacs_hip_bmdot_shared(size_t n, int R, int k, const double* W, size_t ldw, const double* q,
                          double* scratch, double* out, size_t sR, int accumulate,
                          hipStream_t stream) {
  return bmdot_launch(true, n, R, k, W, ldw, q, scratch, out, sR, accumulate, stream);
}

int tacs_hip_bcgs_update(size_t n, int R, int k, const double* W, size_t ldw, const double* c,
                         size_t sR, double* q, hipStream_t stream) {
  if (k <= 0) return 0;
  hipLaunchKernelGGL(bcgs_update<false>, dim3(bgrid(n * R)), dim3(256), 0, stream, n, R, k, W, ldw,
                     c, sR, q);
  return (int)hipGetLastError();
}

int tacs_hip_bcgs_update_shared(size_t n, int R, int k, const double* W, size_t ldw,
                                const double* c, size_t sR, double* q, hipStream_t stream) {
  if (k <= 0) return 0;
  hipLaunchKernelGGL(bcgs_update<true>, dim3(bgrid(n * R)), dim3(256), 0, stream, n, R, k, W, ldw,
                     c, sR, q);
  return (int)hipGetLastError();
}

int tacs_hip_blincomb(size_t n, int R, int kmax, const double* W, size_t ldw, const double* c,
                      size_t sR, const int* cnt, double* y, int add, hipStream_t stream) {
  hipLaunchKernelGGL(blincomb<false>, dim3(bgrid(n * R)), dim3(256), 0, stream, n, R, kmax, W, ldw,
                     c, sR, cnt, y, add);
  return (int)hipGetLastError();
}

int tacs_hip_blincomb_shared(size_t n, int R, int k, const double* W, size_t ldw, const double* c,
                             size_t sR, double* y, int add, hipStream_t stream) {
  hipLaunchKernelGGL(blincomb<true>, dim3(bgrid(n * R)), dim3(256), 0, stream, n, R, k, W, ldw, c,
                     sR, (const int*)NULL, y, add);
  return (int)hipGetLastError();
}

int tacs_hip_bsmall_gemv(int R, int k, const int* cnt, const double* B, size_t sB, int ldb,
                         const double* y, size_t sy, double* z, size_t sz, hipStream_t stream) {
  if (R * k <= 0) return 0;
  hipLaunchKernelGGL(bsmall_gemv, dim3((R * k + 255) / 256), dim3(256), 0, stream, R, k, cnt, B,
                     sB, ldb, y, sy, z, sz);
  return (int)hipGetLastError();
}

size_t tacs_hip_bdotq_scratch_bytes(int R) {
  return (size_t)4096 * 2 * BDQ_KMAX * (R > 0 ? R : 1) * sizeof(double);
}

int tacs_hip_bdotq(size_t n, int R, int nq, int k1, const double* W1, int k2, const double* W2,
                   size_t ldw, const double* q0, const double* q1, double* scratch, double* out0,
                   double* out1, size_t sR, hipStream_t stream) {
  if (R < 1 || R > 64 || nq < 1 || nq > 2 || k1 < 0 || k2 < 0) return TACS_HIP_ERR_ARG;
  const int k = k1 + k2;
  const int G = 64 / pow2_ceil(R);
  size_t nbs = (n + (size_t)G * 4 * BDQ_C - 1) / ((size_t)G * 4 * BDQ_C);  // >= 4 passes per wave
  if (nbs > 4096) nbs = 4096;
  if (nbs < 1) nbs = 1;
  const int nb = (int)nbs;
  for (int v0 = 0; v0 < k; v0 += BDQ_KMAX) {
    const int kk = k - v0 < BDQ_KMAX ? k - v0 : BDQ_KMAX;
    if (nq == 1)
      hipLaunchKernelGGL(bdotq_partial<1>, dim3(nb), dim3(256), 0, stream, n, R, v0, kk, k1, W1,
                         W2, ldw, q0, q0, scratch);
    else
      hipLaunchKernelGGL(bdotq_partial<2>, dim3(nb), dim3(256), 0, stream, n, R, v0, kk, k1, W1,
                         W2, ldw, q0, q1, scratch);
    const int waves = kk * nq * R;
    hipLaunchKernelGGL(bdotq_finalize, dim3((waves + 3) / 4), dim3(256), 0, stream, nb, kk, nq, R,
                       scratch, out0, out1, sR, v0);
  }
  return (int)hipGetLastError();
}

int tacs_hip_bdcgs_update(size_t n, int R, int k1, const double* W1, int k2, const double* W2,
                          size_t ldw, const double* coef, size_t sC, double* u, double* z,
                          hipStream_t stream) {
  if (256 % R == 0) {
    size_t g = (n * R + 256 * BDU_E - 1) / (256 * BDU_E);
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL(bdcgs_update, dim3((int)g), dim3(256), 0, stream, n, R, k1, W1, k2, W2, ldw,
                       coef, sC, u, z);
  } else {
    hipLaunchKernelGGL(bdcgs_update1, dim3(bgrid(n * R)), dim3(256), 0, stream, n, R, k1, W1, k2,
                       W2, ldw, coef, sC, u, z);
  }
  return (int)hipGetLastError();
}

size_t tacs_hip_bdcgs_state_stride(int ld) { return (size_t)2 * ld * (ld - 1) + 5 * (size_t)ld; }

int tacs_hip_bdcgs_coeffs(int R, int i, int Kd, int ld, const int* kd, const double* tgt,
                          const double* dA, const double* dB, size_t S, double* ls, size_t sLS,
                          int* ist, double* cf, size_t sC, int* flags, hipStream_t stream) {
  if (R < 1 || R > 64 || i < 0 || Kd < 0 || ld > BDC_LDMAX || Kd + i + 2 > ld ||
      sLS < tacs_hip_bdcgs_state_stride(ld))
    return TACS_HIP_ERR_ARG;
  hipLaunchKernelGGL(bdcgs_coeffs, dim3(R), dim3(64), 0, stream, i, Kd, ld, kd, tgt, dA, dB, S, ls,
                     sLS, ist, cf, sC, flags);
  return (int)hipGetLastError();
}

int tacs_hip_badd(int R, int k, const double* x, size_t sx, double* y, size_t sy,
                  hipStream_t stream) {
  if (R * k <= 0) return 0;
  hipLaunchKernelGGL(badd, dim3((R * k + 255) / 256), dim3(256), 0, stream, R, k, x, sx, y, sy);
  return (int)hipGetLastError();
}

int tacs_hip_bscale(size_t n, int R, const double* s, double* x, hipStream_t stream) {
  hipLaunchKernelGGL(bscale, dim3(bgrid(n * R)), dim3(256), 0, stream, n, R, s, x);
  return (int)hipGetLastError();
}

int tacs_hip_bsqrt_inv(int R, const double* sq, size_t sS, double* nrm, size_t sN, double* inv,
                       hipStream_t stream) {
  hipLaunchKernelGGL(bsqrt_inv, dim3(1), dim3(64), 0, stream, R, sq, sS, nrm, sN, inv);
  return (int)hipGetLastError();
}

int tacs_hip_bgivens(int R, int i, int m, double* H, size_t sH, double* Qc, double* Qs, size_t sQ,
                     double* res, size_t sRes, double* st, hipStream_t stream) {
  hipLaunchKernelGGL(bgivens, dim3(1), dim3(64), 0, stream, R, i, m, H, sH, Qc, Qs, sQ, res, sRes,
                     st);
  return (int)hipGetLastError();
}

int tacs_hip_bbacksolve(int R, int m, const int* cnt, const double* H, size_t sH, double* res,
                        size_t sRes, hipStream_t stream) {
  hipLaunchKernelGGL(bbacksolve, dim3(1), dim3(64), 0, stream, R, m, cnt, H, sH, res, sRes);
  return (int)hipGetLastError();
}

int tacs_hip_bdiag_apply(int b, int prec, int nrows, int R, const void* Dinv, const double* x,
                         double* y, hipStream_t stream) {
  if ((b != 3 && b != 6) || (prec != 32 && prec != 64)) return TACS_HIP_ERR_ARG;
  size_t N = (size_t)nrows * R;
  if (N == 0) return 0;
  dim3 g((unsigned)((N + 255) / 256)), t(256);
  if (b == 3 && prec == 64)
    hipLaunchKernelGGL((bdiag_apply<3, double>), g, t, 0, stream, nrows, R, (const double*)Dinv, x, y);
  else if (b == 3)
    hipLaunchKernelGGL((bdiag_apply<3, float>), g, t, 0, stream, nrows, R, (const float*)Dinv, x, y);
  else if (prec == 64)
    hipLaunchKernelGGL((bdiag_apply<6, double>), g, t, 0, stream, nrows, R, (const double*)Dinv, x, y);
  else
    hipLaunchKernelGGL((bdiag_apply<6, float>), g, t, 0, stream, nrows, R, (const float*)Dinv, x, y);
  return (int)hipGetLastError();
}

int tacs_hip_bsr3_spmm_mixed(int nrows, int nrhs, const int* rowp, const int* cols, const float* A,
                             const double* X, double* Y, hipStream_t stream) {
  return spmm_slots_launch<float>(tacs_tune_variant(TACS_TUNE_SPMM, 100 + nrhs, nrows), nrows,
                                  nrhs, rowp, cols, A, X, Y, stream);
}

int tacs_hip_bsr3_spmm(int nrows, int nrhs, const int* rowp, const int* cols, const double* A,
                       const double* X, double* Y, hipStream_t stream) {
  return spmm_slots_launch<double>(tacs_tune_variant(TACS_TUNE_SPMM, nrhs, nrows), nrows, nrhs,
                                   rowp, cols, A, X, Y, stream);
}

int tacs_hip_bsr3_spmm_nvariants(int nrhs) { return spmm_nvariants(nrhs); }

int tacs_hip_bsr3_spmm_cheb(int nrows, int R, const int* rowp, const int* cols, const float* A,
                            const float* X, const double* x, const float* Dinv, double* d,
                            double* z, float* Xout, double c1, double c2, hipStream_t stream) {
  return tacs_hip_bsr3_spmm_cheb_mode(0, nrows, R, rowp, cols, A, X, x, Dinv, d, z, Xout, c1, c2,
                                      stream);
}

int tacs_hip_bsr3_spmm_cheb_mode(int mode, int nrows, int R, const int* rowp, const int* cols,
                                 const float* A, const float* X, const double* x,
                                 const float* Dinv, double* d, double* z, float* Xout, double c1,
                                 double c2, hipStream_t stream) {
  if (!spmm_quad_form(R) || R > 64 || X == Xout || mode < 0 || mode > 2 || d == z)
    return TACS_HIP_ERR_ARG;
  if (nrows <= 0) return 0;
  const int grid = 8 * tacs_grid_for((nrows + 7) / 8, 4, 8192);
#define SPMM_CHEB(NLV)                                                                            \
  hipLaunchKernelGGL((bsr3_spmm_quad_cheb<float, NLV, 1, 0>), dim3(grid), dim3(256), 0, stream,  \
                     nrows, R, rowp, cols, A, X, x, Dinv, d, z, Xout, c1, c2, mode)
  switch (pow2_ceil(R / 4)) {
    case 4: SPMM_CHEB(4); break;
    case 8: SPMM_CHEB(8); break;
    default: SPMM_CHEB(16); break;
  }
#undef SPMM_CHEB
  return (int)hipGetLastError();
}

int tacs_hip_bsr3_spmm_variant(int variant, int prec_vals, int nrows, int nrhs, const int* rowp,
                               const int* cols, const void* A, const double* X, double* Y,
                               hipStream_t stream) {
  if (prec_vals == 64)
    return spmm_slots_launch<double>(variant, nrows, nrhs, rowp, cols, (const double*)A, X, Y,
                                     stream);
  if (prec_vals == 32)
    return spmm_slots_launch<float>(variant, nrows, nrhs, rowp, cols, (const float*)A, X, Y,
                                    stream);
  return TACS_HIP_ERR_ARG;
}

int tacs_hip_bbasis_combine(size_t n, int R, int nin, const double* W, size_t ldw, const double* P,
                            size_t sP, int ldp, int nout, double* out, size_t ldo, int add,
                            hipStream_t stream) {
  if (nout < 1 || nout > BBC_MAXOUT || nin < 0 || R < 1 || R > 64) return TACS_HIP_ERR_ARG;
  int vch = (int)((65536 / ((size_t)R * sizeof(double)) - 1) / nout);  // >= 3
  if (vch > 8) vch = 8;
  const size_t lds = (size_t)R * (vch * nout + 1) * sizeof(double);
  size_t g = (n * R + BBC_THREADS - 1) / BBC_THREADS;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(bbasis_combine, dim3((int)g), dim3(BBC_THREADS), lds, stream, n, R, nin, W,
                     ldw, P, sP, ldp, nout, out, ldo, add, vch);
  return (int)hipGetLastError();
}

int tacs_hip_bbroadcast(size_t n, int R, const double* x, double* X, hipStream_t stream) {
  hipLaunchKernelGGL(bbroadcast, dim3(bgrid(n * R)), dim3(256), 0, stream, n, R, x, X);
  return (int)hipGetLastError();
}

size_t tacs_hip_btsgemm_scratch_bytes(int R) {
  (void)R;
  return (size_t)BTS_BLOCKS * 256 * BTS_MAXT * 16 * sizeof(double);
}

int tacs_hip_btsgemm(size_t n, int R, int p, const double* X, int q, const double* Y, size_t ldw,
                     double* scratch, double* out, size_t sR, int ldo, hipStream_t stream) {
  if (R < 1 || R > 64 || p < 0 || q < 0) return TACS_HIP_ERR_ARG;
  if (p == 0 || q == 0) return 0;
  // chunks of pc x qc with R * ceil(pc/4) * ceil(qc/4) <= 256 * BTS_MAXT tiles
  // and (pc + qc) BTS_TR R doubles of LDS <= 64 KB
  const int maxt = 256 * BTS_MAXT;
  int Q4 = (q + 3) / 4, P4 = (p + 3) / 4;
  int q4c = Q4, p4c = P4;
  while ((long)R * p4c * q4c > maxt || (long)(4 * (p4c + q4c)) * BTS_TR * R * 8 > 65536) {
    if (p4c >= q4c && p4c > 1)
      p4c = (p4c + 1) / 2;
    else if (q4c > 1)
      q4c = (q4c + 1) / 2;
    else
      return TACS_HIP_ERR_ARG;
  }
  size_t nb = (n + BTS_TR - 1) / BTS_TR;
  if (nb > BTS_BLOCKS) nb = BTS_BLOCKS;
  for (int b0 = 0; b0 < q; b0 += 4 * q4c) {
    const int qc = q - b0 < 4 * q4c ? q - b0 : 4 * q4c;
    for (int v0 = 0; v0 < p; v0 += 4 * p4c) {
      const int pc = p - v0 < 4 * p4c ? p - v0 : 4 * p4c;
      const size_t lds = (size_t)(pc + qc) * BTS_TR * R * sizeof(double);
      hipLaunchKernelGGL(btsgemm_partial, dim3((int)nb), dim3(256), lds, stream, n, R, pc,
                         X + (size_t)v0 * ldw, qc, Y + (size_t)b0 * ldw, ldw, scratch);
      const int tot = R * pc * qc;
      hipLaunchKernelGGL(btsgemm_finalize, dim3((tot + 255) / 256), dim3(256), 0, stream, (int)nb,
                         R, pc, qc, scratch, out, sR, ldo, v0, b0);
    }
  }
  return (int)hipGetLastError();
}

int tacs_hip_bpack(size_t n, int R, int j, const double* x, double* X, hipStream_t stream) {
  hipLaunchKernelGGL(bpack, dim3(bgrid(n)), dim3(256), 0, stream, n, R, j, x, X, 0);
  return (int)hipGetLastError();
}

int tacs_hip_bunpack(size_t n, int R, int j, const double* X, double* x, hipStream_t stream) {
  hipLaunchKernelGGL(bpack, dim3(bgrid(n)), dim3(256), 0, stream, n, R, j, X, x, 1);
  return (int)hipGetLastError();
}

int tacs_hip_bcheb_step(int b, int prec, int nrows, int R, const void* Dinv, const double* x,
                        const double* Az, double* d, double* z, double c1, double c2, int first,
                        float* z32, hipStream_t stream) {
  if ((b != 3 && b != 6) || (prec != 32 && prec != 64) || R < 1) return TACS_HIP_ERR_ARG;
  const size_t N = (size_t)nrows * R;
  if (N == 0) return 0;
  dim3 g((unsigned)((N + 255) / 256)), t(256);
#define CHEB(BB, TD)                                                                             \
  hipLaunchKernelGGL((bcheb_step<BB, TD>), g, t, 0, stream, nrows, R, (const TD*)Dinv, x, Az, d, \
                     z, c1, c2, first, z32)
  if (b == 3 && prec == 64)
    CHEB(3, double);
  else if (b == 3)
    CHEB(3, float);
  else if (prec == 64)
    CHEB(6, double);
  else
    CHEB(6, float);
#undef CHEB
  return (int)hipGetLastError();
}

int tacs_hip_seg_pack(int maxn, int R, const int* off, const double* x, double* X,
                      hipStream_t stream) {
  if (R < 1 || R > 64 || maxn < 0) return TACS_HIP_ERR_ARG;
  const size_t total = (size_t)3 * maxn * R;
  if (total == 0) return 0;
  hipLaunchKernelGGL(seg_pack, dim3(bgrid(total)), dim3(256), 0, stream, maxn, R, off, x, X);
  return (int)hipGetLastError();
}

int tacs_hip_seg_unpack(int nnodes, int R, const int* off, const double* X, double* x,
                        hipStream_t stream) {
  if (R < 1 || R > 64 || nnodes < 0) return TACS_HIP_ERR_ARG;
  if (nnodes == 0) return 0;
  hipLaunchKernelGGL(seg_unpack, dim3(bgrid((size_t)3 * nnodes)), dim3(256), 0, stream, nnodes, R,
                     off, X, x);
  return (int)hipGetLastError();
}

int tacs_hip_seg_to_union(int maxn, int R, const int* off, const double* X, double* x,
                          hipStream_t stream) {
  if (R < 1 || R > 64 || maxn < 0) return TACS_HIP_ERR_ARG;
  if (maxn == 0) return 0;
  hipLaunchKernelGGL(seg_transpose<true>, dim3((maxn + SEG_T - 1) / SEG_T), dim3(256), 0, stream,
                     maxn, R, off, X, x);
  return (int)hipGetLastError();
}

int tacs_hip_seg_from_union(int maxn, int R, const int* off, const double* x, double* X,
                            hipStream_t stream) {
  if (R < 1 || R > 64 || maxn < 0) return TACS_HIP_ERR_ARG;
  if (maxn == 0) return 0;
  hipLaunchKernelGGL(seg_transpose<false>, dim3((maxn + SEG_T - 1) / SEG_T), dim3(256), 0, stream,
                     maxn, R, off, x, X);
  return (int)hipGetLastError();
}

int tacs_hip_seg_bdiag_apply(int prec, int maxn, int R, const int* off, const void* Dinv,
                             const double* x, double* y, hipStream_t stream) {
  if (R < 1 || R > 64 || maxn < 0 || (prec != 64 && prec != 32)) return TACS_HIP_ERR_ARG;
  const size_t total = (size_t)maxn * R;
  if (total == 0) return 0;
  if (prec == 64)
    hipLaunchKernelGGL(seg_bdiag_apply<double>, dim3(bgrid(total)), dim3(256), 0, stream, maxn, R,
                       off, (const double*)Dinv, x, y);
  else
    hipLaunchKernelGGL(seg_bdiag_apply<float>, dim3(bgrid(total)), dim3(256), 0, stream, maxn, R,
                       off, (const float*)Dinv, x, y);
  return (int)hipGetLastError();
}

int tacs_hip_seg_spmm3(int prec, int maxn, int R, const int* off, const int* rowp,
                       const int* cols, const void* A, const double* X, double* Y,
                       hipStream_t stream) {
  if (R < 1 || R > 64 || maxn < 0 || (prec != 64 && prec != 32)) return TACS_HIP_ERR_ARG;
  const size_t total = (size_t)maxn * R;
  if (total == 0) return 0;
  if (prec == 64)
    hipLaunchKernelGGL(seg_spmm3<double>, dim3(bgrid(total)), dim3(256), 0, stream, maxn, R, off,
                       rowp, cols, (const double*)A, X, Y);
  else
    hipLaunchKernelGGL(seg_spmm3<float>, dim3(bgrid(total)), dim3(256), 0, stream, maxn, R, off,
                       rowp, cols, (const float*)A, X, Y);
  return (int)hipGetLastError();
}

}  // extern "C"
