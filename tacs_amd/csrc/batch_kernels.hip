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

// y = D^{-1} x per 3x3 block for interleaved block vectors (block-Jacobi)
template <typename TD>
__global__ void bdiag_apply3(int nrows, int R, const TD* __restrict__ Dinv,
                             const double* __restrict__ x, double* __restrict__ y) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nrows * R) return;
  int node = (int)(t / R), j = (int)(t % R);
  const TD* d = &Dinv[(size_t)9 * node];
  double x0 = x[((size_t)3 * node) * R + j], x1 = x[((size_t)3 * node + 1) * R + j],
         x2 = x[((size_t)3 * node + 2) * R + j];
#pragma unroll
  for (int r = 0; r < 3; r++)
    y[((size_t)3 * node + r) * R + j] = (double)d[3 * r] * x0 + (double)d[3 * r + 1] * x1 +
                                        (double)d[3 * r + 2] * x2;
}

// Multi right-hand-side 3x3 SpMM with FP32 values and FP64 block vectors:
// one wavefront per block row, lane = (right-hand side, row of the block).
__global__ __launch_bounds__(256) void bsr3_spmm_mixed(int nrows, int nrhs,
                                                       const int* __restrict__ rowp,
                                                       const int* __restrict__ cols,
                                                       const float* __restrict__ A,
                                                       const double* __restrict__ X,
                                                       double* __restrict__ Y) {
  __shared__ float blk[4][9 * 8];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = gridDim.x * 4;
  for (int row = blockIdx.x * 4 + wid; row < nrows; row += nw) {
    const int rs = rowp[row];
    const int nb = rowp[row + 1] - rs;
    for (int j0 = 0; j0 < nrhs; j0 += 21) {
      const int jj = lane / 3, r = lane - 3 * jj;
      const int j = j0 + jj;
      const bool ok = (lane < 63) && (j < nrhs);
      double acc = 0.0;
      for (int s0 = 0; s0 < nb; s0 += 8) {
        const int nbs = min(8, nb - s0);
        for (int k = lane; k < nbs * 9; k += 64) blk[wid][k] = A[(size_t)(rs + s0) * 9 + k];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (ok) {
          for (int s = 0; s < nbs; s++) {
            int col = cols[rs + s0 + s];
            const float* a = &blk[wid][9 * s + 3 * r];
            const double* xx = &X[(size_t)3 * col * nrhs + j];
            acc += (double)a[0] * xx[0] + (double)a[1] * xx[nrhs] + (double)a[2] * xx[2 * nrhs];
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      if (ok) Y[(size_t)(3 * row + r) * nrhs + j] = acc;
    }
  }
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

int tacs_hip_bdiag_apply3(int prec, int nrows, int R, const void* Dinv, const double* x,
                          double* y, hipStream_t stream) {
  size_t N = (size_t)nrows * R;
  if (N == 0) return 0;
  if (prec == 64)
    hipLaunchKernelGGL((bdiag_apply3<double>), dim3((N + 255) / 256), dim3(256), 0, stream, nrows,
                       R, (const double*)Dinv, x, y);
  else
    hipLaunchKernelGGL((bdiag_apply3<float>), dim3((N + 255) / 256), dim3(256), 0, stream, nrows,
                       R, (const float*)Dinv, x, y);
  return (int)hipGetLastError();
}

int tacs_hip_bsr3_spmm_mixed(int nrows, int nrhs, const int* rowp, const int* cols, const float* A,
                             const double* X, double* Y, hipStream_t stream) {
  if (nrows <= 0 || nrhs <= 0) return 0;
  int grid = tacs_grid_for(nrows, 4, 65536);
  hipLaunchKernelGGL(bsr3_spmm_mixed, dim3(grid), dim3(256), 0, stream, nrows, nrhs, rowp, cols,
                     A, X, Y);
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

}  // extern "C"
