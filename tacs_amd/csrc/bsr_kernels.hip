// gfx950 kernels for TACS's block-CSR matrix (BCSRMat, row-major b x b
// blocks, src/bpmat/BCSRMat.h:34): SpMV / multi-RHS SpMM, the block-Jacobi
// preconditioner (inverse diagonal blocks, as BCSRMat::factorDiag
// src/bpmat/BCSRMat.cpp:1527 computes them), Dirichlet rows
// (BCSRMat::zeroRow :2027) and precision conversion.
#include <algorithm>

#include "tacs_hip_common.h"

#define WAVE_SYNC()                                        \
  do {                                                     \
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); \
    __builtin_amdgcn_wave_barrier();                       \
  } while (0)

// ---------------------------------------------------------------------------
// 3x3 SpMV: one wavefront per block row.  Lane l = 9*bo + 3*r + c (bo < 7)
// owns entry (r, c) of blocks bo, bo+7, bo+14, ... of the row, so every load
// instruction of the wave reads 63 consecutive values of A.  Rows of up to
// 7*NT blocks (28 for NT = 4: every row of a structured 8-node hex mesh) are
// handled with all loads of the row issued before the arithmetic; the
// values are streamed with non-temporal loads so that x stays in cache.
// Partial sums are reduced over c and bo with cross-lane shuffles.
// y = A x (+ y if add).  TA: value type of A; T: vectors (FP32 values with
// FP64 vectors is the fixed FP32-rounded operator of the deflated solver).
// ---------------------------------------------------------------------------
template <typename TA, typename T, typename ACC, int NT, int R>
__global__ __launch_bounds__(256) void bsr3_spmv(int nrows, const int* __restrict__ rowp,
                                                 const int* __restrict__ cols,
                                                 const TA* __restrict__ A,
                                                 const T* __restrict__ x, T* __restrict__ y,
                                                 int add) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int bo = lane / 9;
  const int rc = lane - 9 * bo;
  const int c = rc % 3;
  const bool ok = lane < 63;
  // XCD-aware rows: workgroups go round-robin to the 8 XCDs, so XCD x
  // (blockIdx % 8) takes the contiguous eighth x of the rows, and the x
  // entries that neighbouring rows share stay in that XCD's L2
  const int xcd = blockIdx.x & 7, nbx = gridDim.x >> 3;
  const int rbeg = (int)((long)nrows * xcd / 8), rend = (int)((long)nrows * (xcd + 1) / 8);
  const int nw = nbx * 4;
  // each wave handles R consecutive rows per pass, with every load of the R
  // rows in flight together; the row pointers of the next pass are loaded
  // during this one, so a pass waits on one memory latency, not two
  int row0 = rbeg + ((blockIdx.x >> 3) * 4 + wid) * R;
  int pre[R + 1];
#pragma unroll
  for (int r = 0; r <= R; r++) pre[r] = row0 + r <= rend ? rowp[row0 + r] : 0;
  for (; row0 < rend; row0 += nw * R) {
    int rs[R], nb[R];
    bool fast = true;
#pragma unroll
    for (int r = 0; r < R; r++) {
      int row = row0 + r;
      rs[r] = row < rend ? pre[r] : 0;
      nb[r] = row < rend ? pre[r + 1] - pre[r] : 0;
      fast = fast && nb[r] <= 7 * NT;
    }
    const int nxt = row0 + nw * R;
#pragma unroll
    for (int r = 0; r <= R; r++) pre[r] = nxt + r <= rend ? rowp[nxt + r] : 0;
    ACC acc[R];
    if (fast && sizeof(TA) == 8) {
      int col[R][NT];
      TA av[R][NT];
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int t = 0; t < NT; t++) {
          int s = bo + 7 * t;
          col[r][t] = (ok && s < nb[r]) ? cols[rs[r] + s] : -1;
        }
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int t = 0; t < NT; t++) {
          int s = bo + 7 * t;
          av[r][t] = (ok && s < nb[r])
                         ? __builtin_nontemporal_load(&A[((size_t)rs[r] + s) * 9 + rc])
                         : (TA)0;
        }
#pragma unroll
      for (int r = 0; r < R; r++) {
        acc[r] = 0;
#pragma unroll
        for (int t = 0; t < NT; t++)
          if (col[r][t] >= 0) acc[r] += (ACC)av[r][t] * (ACC)x[3 * col[r][t] + c];
      }
    } else if (fast) {
      // FP32 values: every load is unconditional (an absent block reads
      // entry 0 and is masked in the sum), so the compiler issues all of them
      // before the first wait instead of one branch-and-wait per block (with
      // the conditional form each FP32 load was followed by its wait and
      // conversion)
      int col[R][NT];
      TA av[R][NT];
      T xv[R][NT];
      bool use[R][NT];
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const int s = bo + 7 * t;
          use[r][t] = ok && s < nb[r];
          const int e = use[r][t] ? rs[r] + s : 0;
          col[r][t] = cols[e];
          av[r][t] = __builtin_nontemporal_load(&A[(size_t)e * 9 + (use[r][t] ? rc : 0)]);
        }
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int t = 0; t < NT; t++) xv[r][t] = x[3 * col[r][t] + c];
#pragma unroll
      for (int r = 0; r < R; r++) {
        acc[r] = 0;
#pragma unroll
        for (int t = 0; t < NT; t++)
          acc[r] += use[r][t] ? (ACC)av[r][t] * (ACC)xv[r][t] : (ACC)0;
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) {
        acc[r] = 0;
        if (ok)
          for (int s = bo; s < nb[r]; s += 7)
            acc[r] += (ACC)__builtin_nontemporal_load(&A[((size_t)rs[r] + s) * 9 + rc]) *
                      (ACC)x[3 * cols[rs[r] + s] + c];
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      // reduce over c (lanes l, l+1, l+2 with c = 0), then over bo: lane q
      // (< 3) gathers lanes 9*b + 3*q
      ACC a1 = __shfl(acc[r], (lane + 1) & 63, 64);
      ACC a2 = __shfl(acc[r], (lane + 2) & 63, 64);
      ACC s3 = acc[r] + a1 + a2;
      ACC tot = 0;
#pragma unroll
      for (int b = 0; b < 7; b++) tot += __shfl(s3, (9 * b + 3 * lane) & 63, 64);
      if (lane < 3 && row0 + r < rend) {
        size_t o = 3 * (size_t)(row0 + r) + lane;
        y[o] = add ? (T)((ACC)y[o] + tot) : (T)tot;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Generic b x b SpMV (b = 6 for shells): one wavefront per block row, lanes
// stride over the row's contiguous values.
// ---------------------------------------------------------------------------
template <int B, typename TA, typename T, typename ACC>
__global__ __launch_bounds__(256) void bsr_spmv_generic(int nrows, const int* __restrict__ rowp,
                                                        const int* __restrict__ cols,
                                                        const TA* __restrict__ A,
                                                        const T* __restrict__ x,
                                                        T* __restrict__ y, int add) {
  constexpr int B2 = B * B;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = gridDim.x * 4;
  for (int row = blockIdx.x * 4 + wid; row < nrows; row += nw) {
    const int rs = rowp[row];
    const int nval = (rowp[row + 1] - rs) * B2;
    ACC acc[B];
#pragma unroll
    for (int r = 0; r < B; r++) acc[r] = 0;
    const TA* Ar = A + (size_t)rs * B2;
    for (int k = lane; k < nval; k += 64) {
      int s = k / B2, rc = k - s * B2, r = rc / B, c = rc - r * B;
      ACC v = (ACC)Ar[k] * (ACC)x[(size_t)cols[rs + s] * B + c];
#pragma unroll
      for (int rr = 0; rr < B; rr++) acc[rr] += (rr == r) ? v : (ACC)0;
    }
#pragma unroll
    for (int r = 0; r < B; r++) {
      ACC t = acc[r];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
      acc[r] = t;
    }
    if (lane < B) {
      ACC t = 0;
#pragma unroll
      for (int r = 0; r < B; r++) t = (r == lane) ? acc[r] : t;
      size_t o = (size_t)B * row + lane;
      y[o] = add ? (T)((ACC)y[o] + t) : (T)t;
    }
  }
}

// ---------------------------------------------------------------------------
// Diagonal-block positions, block-Jacobi inverse and application.
// ---------------------------------------------------------------------------
__global__ void bsr_find_diag(int nrows, const int* __restrict__ rowp,
                              const int* __restrict__ cols, int* __restrict__ diag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  int lo = rowp[i], hi = rowp[i + 1] - 1, d = -1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int v = cols[mid];
    if (v == i) {
      d = mid;
      break;
    }
    if (v < i) lo = mid + 1;
    else hi = mid - 1;
  }
  diag[i] = d;
}

template <int B>
__global__ void bsr_diag_inverse(int nrows, const int* __restrict__ diag,
                                 const double* __restrict__ A, double* __restrict__ Dinv,
                                 float* __restrict__ Dinv32, int* __restrict__ status) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  double m[B * B], r[B * B];
  int d = diag[i];
  if (d < 0) {
    atomicExch(status, TACS_HIP_ERR_MISSING_ENTRY);
    return;
  }
#pragma unroll
  for (int k = 0; k < B * B; k++) {
    m[k] = A[(size_t)d * B * B + k];
    r[k] = (k % (B + 1) == 0) ? 1.0 : 0.0;
  }
  // Gauss-Jordan with partial pivoting
  for (int c = 0; c < B; c++) {
    int p = c;
    for (int ii = c + 1; ii < B; ii++)
      if (fabs(m[B * ii + c]) > fabs(m[B * p + c])) p = ii;
    if (m[B * p + c] == 0.0) {
      atomicExch(status, TACS_HIP_ERR_SINGULAR);
      return;
    }
    if (p != c) {
      for (int k = 0; k < B; k++) {
        double t = m[B * c + k];
        m[B * c + k] = m[B * p + k];
        m[B * p + k] = t;
        t = r[B * c + k];
        r[B * c + k] = r[B * p + k];
        r[B * p + k] = t;
      }
    }
    double dd = 1.0 / m[B * c + c];
    for (int k = 0; k < B; k++) {
      m[B * c + k] *= dd;
      r[B * c + k] *= dd;
    }
    for (int ii = 0; ii < B; ii++) {
      if (ii != c) {
        double f = m[B * ii + c];
        for (int k = 0; k < B; k++) {
          m[B * ii + k] -= f * m[B * c + k];
          r[B * ii + k] -= f * r[B * c + k];
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < B * B; k++) {
    if (Dinv) Dinv[(size_t)i * B * B + k] = r[k];
    if (Dinv32) Dinv32[(size_t)i * B * B + k] = (float)r[k];
  }
}

// y = D^{-1} x per block; one thread per block row
template <int B, typename TD, typename T>
__global__ void bsr_diag_apply(int nrows, const TD* __restrict__ Dinv, const T* __restrict__ x,
                               T* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  T xx[B];
#pragma unroll
  for (int c = 0; c < B; c++) xx[c] = x[(size_t)B * i + c];
  const TD* d = Dinv + (size_t)B * B * i;
#pragma unroll
  for (int r = 0; r < B; r++) {
    T v = 0;
#pragma unroll
    for (int c = 0; c < B; c++) v += (T)d[B * r + c] * xx[c];
    y[(size_t)B * i + r] = v;
  }
}

// Dirichlet rows on an assembled matrix: one wavefront per BC node.
__global__ void bsr_zero_rows(int b, int nbc, const int* __restrict__ nodes,
                              const int* __restrict__ vars, const int* __restrict__ rowp,
                              const int* __restrict__ cols, double* __restrict__ A,
                              int ident) {
  int q = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (q >= nbc) return;
  int row = nodes[q], mask = vars[q];
  int rs = rowp[row], b2 = b * b;
  int nval = (rowp[row + 1] - rs) * b2;
  for (int k = lane; k < nval; k += 64) {
    int s = k / b2, rc = k - s * b2, r = rc / b, c = rc - r * b;
    if (mask & (1 << r)) {
      A[(size_t)rs * b2 + k] = (ident && cols[rs + s] == row && r == c) ? 1.0 : 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Sliced ELL copy of a b x b BCSR matrix (b = 3 or 6) for the product (no
// counterpart in the reference, whose BCSRMat is the only layout; the BCSR
// stays the matrix of record).  Rows are cut into slices of 64 (one
// wavefront); slice s has width w_s = the longest row of the slice in blocks;
// slot t of row 64 s + lane has its column at cs[coff[s] + 64 t + lane] and
// value q (row-major in the block) at vs[b^2 coff[s] + 64 (b^2 t + q) + lane].
// Padding slots point at the row itself with zero values.  With lane = row
// every load of a wavefront is 64 consecutive entries, no cross-lane
// reduction is needed, and each row adds its blocks in column order with the
// reference's in-block order (BCSRMatVecMult3 src/bpmat/BCSRMatMult3.cpp:28).
// ---------------------------------------------------------------------------
// Matrix values of the sliced ELL copies: double, float, or bfloat16 kept
// as its 16 bits (unsigned short; the upper half of a float).
__device__ __forceinline__ double sell_val(double a) { return a; }
__device__ __forceinline__ double sell_val(float a) { return (double)a; }
__device__ __forceinline__ double sell_val(unsigned short h) {
  return (double)__uint_as_float((unsigned int)h << 16);
}
// float -> bfloat16, rounded to nearest even (NaN kept a NaN)
__device__ __forceinline__ unsigned short f32_to_bf16(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
template <typename TA>
struct SellDinv {  // type of the inverse diagonal blocks that go with TA values
  typedef TA type;
};
template <>
struct SellDinv<unsigned short> {
  typedef float type;
};

template <int B, typename TA, typename T, bool NT = true>
__device__ inline void sell_slot(const TA* __restrict__ vs, long v, const T* __restrict__ x, int col,
                                 double* yr) {
  TA a[B * B];
#pragma unroll
  for (int q = 0; q < B * B; q++)
    a[q] = NT ? __builtin_nontemporal_load(&vs[v + 64 * q]) : vs[v + 64 * q];
  double xc[B];
#pragma unroll
  for (int c = 0; c < B; c++) xc[c] = (double)x[(size_t)B * col + c];
#pragma unroll
  for (int r = 0; r < B; r++) {
    double t = sell_val(a[B * r]) * xc[0];
#pragma unroll
    for (int c = 1; c < B; c++) t += sell_val(a[B * r + c]) * xc[c];
    yr[r] += t;
  }
}

// U slots per iteration: every load of the U slots is issued before the
// first block is multiplied; the blocks are still added in slot order, so
// every U gives the same sums.  NT: the values are read with non-temporal
// loads (streamed past the caches, which keep x).  The launch takes U and
// NT from the tuning table (sell_variant).
// Chebyshev epilogue of the product (CHEB): with the product a = A z of
// the row, w = D^{-1} (rhs - a), d = c1 d + c2 w and zout = z + d for the
// row (one step of the Chebyshev iteration for A z = rhs, Saad, Alg. 12.1),
// instead of writing a.  zout must not be z (other rows still read z).
// Three-term form (zprev not NULL; d unused): with d = z - zprev the same
// step is zout = (1 + c1) z - c1 zprev + c2 w, which reads zprev instead of
// reading and writing d; zout may be zprev (each row reads its own entries
// before writing them).
struct SellCheb {
  const void* Dinv;  // b x b inverse diagonal blocks, the value type of the matrix
  const double* rhs;
  double* d;
  double* zout;
  double c1, c2;
  const double* zprev;
};
template <int B, int U, typename TA, typename T, bool NT = true, bool CHEB = false>
__global__ __launch_bounds__(256) void sell_spmv(int nslices, int nrows,
                                                 const int* __restrict__ width,
                                                 const long* __restrict__ coff,
                                                 const int* __restrict__ cs,
                                                 const TA* __restrict__ vs,
                                                 const T* __restrict__ x, T* __restrict__ y,
                                                 SellCheb ch = SellCheb{}) {
  constexpr int B2 = B * B;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int s = blockIdx.x * 4 + wid; s < nslices; s += gridDim.x * 4) {
    const int w = width[s];
    const long co = coff[s] + lane, vo = B2 * coff[s] + lane;
    double yr[B];
#pragma unroll
    for (int r = 0; r < B; r++) yr[r] = 0.0;
    int t = 0;
    if (U > 1) {
      for (; t + U <= w; t += U) {
        int cu[U];
        TA a[U][B2];
#pragma unroll
        for (int u = 0; u < U; u++) cu[u] = cs[co + 64 * (t + u)];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
          for (int q = 0; q < B2; q++)
            a[u][q] = NT ? __builtin_nontemporal_load(&vs[vo + 64 * (B2 * (t + u) + q)])
                         : vs[vo + 64 * (B2 * (t + u) + q)];
#pragma unroll
        for (int u = 0; u < U; u++) {
          double xc[B];
#pragma unroll
          for (int c = 0; c < B; c++) xc[c] = (double)x[(size_t)B * cu[u] + c];
#pragma unroll
          for (int r = 0; r < B; r++) {
            double v = sell_val(a[u][B * r]) * xc[0];
#pragma unroll
            for (int c = 1; c < B; c++) v += sell_val(a[u][B * r + c]) * xc[c];
            yr[r] += v;
          }
        }
      }
    }
    for (; t < w; t++) sell_slot<B, TA, T, NT>(vs, vo + 64 * B2 * t, x, cs[co + 64 * t], yr);
    const int row = 64 * s + lane;
    if (row < nrows) {
      if constexpr (CHEB) {
        typedef typename SellDinv<TA>::type TD;
        const TD* D = (const TD*)ch.Dinv + (size_t)B2 * row;
        double rr[B];
#pragma unroll
        for (int r = 0; r < B; r++) rr[r] = ch.rhs[(size_t)B * row + r] - yr[r];
#pragma unroll
        for (int r = 0; r < B; r++) {
          double wv = (double)D[B * r] * rr[0];
#pragma unroll
          for (int c = 1; c < B; c++) wv += (double)D[B * r + c] * rr[c];
          const size_t o = (size_t)B * row + r;
          if (ch.zprev) {
            const double zr = (double)x[o];
            ch.zout[o] = (1.0 + ch.c1) * zr - ch.c1 * ch.zprev[o] + ch.c2 * wv;
          } else {
            const double dn = ch.c1 * ch.d[o] + ch.c2 * wv;
            ch.d[o] = dn;
            ch.zout[o] = (double)x[o] + dn;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < B; r++) y[(size_t)B * row + r] = (T)yr[r];
      }
    }
  }
}

// Sliced ELL copy of the block-CSR values (and, with cs, the columns): one
// lane per row of a 64-row slice, T slots per pass.  A lane issues the loads
// of its row's T consecutive blocks (one contiguous run of T b^2 values) back
// to back, so every cache line of the run is fetched once and consumed by the
// burst, then writes slot t, value q: the wave's 64 lanes store 64
// consecutive entries.  Padding slots (t beyond a row's length) are zero
// blocks in the row's own column.  v64, v32 and v16 receive the FP64, FP32
// and bfloat16 copies in the same pass.
template <int B>
struct SellFillCfg {
  static constexpr int T = B == 3 ? 4 : 1;  // slots per pass (8: 1.3x fetch, but occupancy 2 and slower)
};
template <int B, int TW>
__device__ __forceinline__ void sell_fill_pass(int lane, int t0, int rs, int nb, int rowc,
                                               long vo, long co, const int* __restrict__ cols,
                                               const double* __restrict__ A,
                                               double* __restrict__ v64, float* __restrict__ v32,
                                               unsigned short* __restrict__ v16,
                                               int* __restrict__ cs) {
  constexpr int B2 = B * B;
  const double* a = A + (size_t)B2 * (rs + t0);
  double v[TW * B2];
#pragma unroll
  for (int t = 0; t < TW; t++) {
    const bool use = t0 + t < nb;
#pragma unroll
    for (int q = 0; q < B2; q++) v[t * B2 + q] = use ? a[t * B2 + q] : 0.0;
  }
#pragma unroll
  for (int t = 0; t < TW; t++) {
#pragma unroll
    for (int q = 0; q < B2; q++) {
      const size_t o = vo + 64 * ((size_t)B2 * (t0 + t) + q) + lane;
      if (v64) v64[o] = v[t * B2 + q];
      if (v32) v32[o] = (float)v[t * B2 + q];
      if (v16) v16[o] = f32_to_bf16((float)v[t * B2 + q]);
    }
    if (cs) cs[co + 64 * (t0 + t) + lane] = t0 + t < nb ? cols[rs + t0 + t] : rowc;
  }
}
template <int B>
__global__ __launch_bounds__(256) void sell_fill(int nslices, int nrows, const int* __restrict__ rowp,
                                                 const int* __restrict__ cols,
                                                 const int* __restrict__ width,
                                                 const long* __restrict__ coff,
                                                 const double* __restrict__ A,
                                                 double* __restrict__ v64, float* __restrict__ v32,
                                                 unsigned short* __restrict__ v16,
                                                 int* __restrict__ cs) {
  constexpr int B2 = B * B, T = SellFillCfg<B>::T;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int s = blockIdx.x * 4 + wid; s < nslices; s += gridDim.x * 4) {
    const int w = width[s];
    const long co = coff[s], vo = B2 * co;
    const int row = 64 * s + lane;
    const int rs = row < nrows ? rowp[row] : 0;
    const int nb = row < nrows ? rowp[row + 1] - rs : 0;
    const int rowc = row < nrows ? row : 0;
    int t0 = 0;
    for (; t0 + T <= w; t0 += T)
      sell_fill_pass<B, T>(lane, t0, rs, nb, rowc, vo, co, cols, A, v64, v32, v16, cs);
    for (; t0 < w; t0++)
      sell_fill_pass<B, 1>(lane, t0, rs, nb, rowc, vo, co, cols, A, v64, v32, v16, cs);
  }
}

// The same copies (3x3 blocks) staged through LDS, one wave per slice and
// T slots per chunk.  Read: row after row of the slice, lane k < 9 T takes
// value k of the row's run of T blocks (a run is 9 T contiguous values of
// the BCSR array, so each load instruction reads one contiguous piece; the
// row's start and length are wave-uniform), RB rows in flight per batch;
// the columns one lane per row.  Store, after a barrier: lane l writes row
// l's slots from LDS, each store instruction 64 consecutive entries of a
// copy.  NTL: the BCSR values are read with non-temporal loads; NTS: the
// copies are written with non-temporal stores.
template <int T, int RB, bool NTL, bool NTS>
__global__ __launch_bounds__(64) void sell_fill3_lds(int nslices, int nrows,
                                                     const int* __restrict__ rowp,
                                                     const int* __restrict__ cols,
                                                     const int* __restrict__ width,
                                                     const long* __restrict__ coff,
                                                     const double* __restrict__ A,
                                                     double* __restrict__ v64,
                                                     float* __restrict__ v32,
                                                     unsigned short* __restrict__ v16,
                                                     int* __restrict__ cs) {
  constexpr int B2 = 9, RUN = T * B2, LD = RUN % 2 ? RUN : RUN + 1;
  static_assert(RUN <= 64, "a run is at most one wave wide");
  __shared__ double buf[64 * LD];
  const int lane = threadIdx.x;
  for (int s = blockIdx.x; s < nslices; s += gridDim.x) {
    const int w = width[s];
    const long co = coff[s], vo = (long)B2 * co;
    const int row = 64 * s + lane;
    const int rs = row < nrows ? rowp[row] : 0;
    const int nb = row < nrows ? rowp[row + 1] - rs : 0;
    const int rowc = row < nrows ? row : 0;
    for (int t0 = 0; t0 < w; t0 += T) {
      int c[T];
#pragma unroll
      for (int t = 0; t < T; t++) c[t] = t0 + t < nb ? cols[rs + t0 + t] : rowc;
      __syncthreads();  // the previous chunk's stores have read buf
#pragma unroll 1
      for (int r0 = 0; r0 < 64; r0 += RB) {
        double v[RB];
#pragma unroll
        for (int j = 0; j < RB; j++) {
          const int rsr = __builtin_amdgcn_readlane(rs, r0 + j);
          const int nbr = __builtin_amdgcn_readlane(nb, r0 + j);
          const int len = B2 * min(max(nbr - t0, 0), T);
          const double* a = &A[(size_t)B2 * (rsr + t0) + lane];
          v[j] = lane < len ? (NTL ? __builtin_nontemporal_load(a) : *a) : 0.0;
        }
#pragma unroll
        for (int j = 0; j < RB; j++)
          if (lane < RUN) buf[(r0 + j) * LD + lane] = v[j];
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < T; t++) {
        if (t0 + t >= w) break;
#pragma unroll
        for (int q = 0; q < B2; q++) {
          const double x = buf[lane * LD + B2 * t + q];
          const size_t o = vo + 64 * ((size_t)B2 * (t0 + t) + q) + lane;
          if (NTS) {
            if (v64) __builtin_nontemporal_store(x, &v64[o]);
            if (v32) __builtin_nontemporal_store((float)x, &v32[o]);
            if (v16) __builtin_nontemporal_store(f32_to_bf16((float)x), &v16[o]);
          } else {
            if (v64) v64[o] = x;
            if (v32) v32[o] = (float)x;
            if (v16) v16[o] = f32_to_bf16((float)x);
          }
        }
        if (cs) cs[co + 64 * (t0 + t) + lane] = c[t];
      }
    }
  }
}

__global__ void convert_f64_to_f32(size_t n, const double* __restrict__ x, float* __restrict__ y) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = (float)x[i];
}

// x[b i + r] = 0 where row r of block row i is an identity row (Dirichlet
// row: diagonal value 1, every other value of the row 0).
__global__ void bsr_zero_identity_rows(int b, int nrows, const int* __restrict__ rowp,
                                       const int* __restrict__ cols,
                                       const double* __restrict__ A, double* __restrict__ x) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nrows * b) return;
  const int node = (int)(t / b), r = (int)(t % b);
  bool ident = true;
  for (int s = rowp[node]; s < rowp[node + 1] && ident; s++) {
    const double* a = &A[((size_t)s * b + r) * b];
    for (int c = 0; c < b; c++) {
      const double want = (cols[s] == node && c == r) ? 1.0 : 0.0;
      if (a[c] != want) ident = false;
    }
  }
  if (ident) x[t] = 0.0;
}

extern "C" {

int tacs_hip_bsr_spmv(int b, int nrows, const int* rowp, const int* cols, const double* A,
                      const double* x, double* y, int add, hipStream_t stream) {
  if (nrows <= 0) return 0;
  int grid = 8 * tacs_grid_for((nrows + 7) / 8, 8, 8192);  // a multiple of 8 (XCD shares)
  if (b == 3)
    hipLaunchKernelGGL((bsr3_spmv<double, double, double, 4, 2>), dim3(grid), dim3(256), 0, stream, nrows, rowp,
                       cols, A, x, y, add);
  else if (b == 6)
    hipLaunchKernelGGL((bsr_spmv_generic<6, double, double, double>), dim3(grid), dim3(256), 0, stream,
                       nrows, rowp, cols, A, x, y, add);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_bsr_spmv_f32(int b, int nrows, const int* rowp, const int* cols, const float* A,
                          const float* x, float* y, int add, hipStream_t stream) {
  if (nrows <= 0) return 0;
  int grid = 8 * tacs_grid_for((nrows + 7) / 8, 8, 8192);  // a multiple of 8 (XCD shares)
  if (b == 3)
    hipLaunchKernelGGL((bsr3_spmv<float, float, float, 4, 2>), dim3(grid), dim3(256), 0, stream, nrows, rowp,
                       cols, A, x, y, add);
  else if (b == 6)
    hipLaunchKernelGGL((bsr_spmv_generic<6, float, float, float>), dim3(grid), dim3(256), 0, stream,
                       nrows, rowp, cols, A, x, y, add);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_bsr_spmv_mixed(int b, int nrows, const int* rowp, const int* cols, const float* A,
                            const double* x, double* y, int add, hipStream_t stream) {
  if (nrows <= 0) return 0;
  int grid = 8 * tacs_grid_for((nrows + 7) / 8, 8, 8192);  // a multiple of 8 (XCD shares)
  if (b == 3)
    hipLaunchKernelGGL((bsr3_spmv<float, double, double, 4, 2>), dim3(grid), dim3(256), 0, stream,
                       nrows, rowp, cols, A, x, y, add);
  else if (b == 6)
    hipLaunchKernelGGL((bsr_spmv_generic<6, float, double, double>), dim3(grid), dim3(256), 0,
                       stream, nrows, rowp, cols, A, x, y, add);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_bsr_diag_apply_mixed(int b, int nrows, const float* Dinv, const double* x, double* y,
                                  hipStream_t stream) {
  if (nrows <= 0) return 0;
  dim3 g((nrows + 255) / 256), t(256);
  if (b == 3)
    hipLaunchKernelGGL((bsr_diag_apply<3, float, double>), g, t, 0, stream, nrows, Dinv, x, y);
  else if (b == 6)
    hipLaunchKernelGGL((bsr_diag_apply<6, float, double>), g, t, 0, stream, nrows, Dinv, x, y);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

size_t tacs_hip_sell_slots(int nrows, const int* rowp_host, int* width, long* coff) {
  const int ns = (nrows + 63) / 64;
  long off = 0;
  for (int s = 0; s < ns; s++) {
    int w = 0;
    for (int r = 64 * s; r < 64 * s + 64 && r < nrows; r++) w = std::max(w, rowp_host[r + 1] - rowp_host[r]);
    if (width) width[s] = w;
    if (coff) coff[s] = off;
    off += 64L * w;
  }
  if (coff) coff[ns] = off;
  return (size_t)off;
}

// Launch variants of the fill (all write the same arrays): 3x3 blocks 0 the
// lane-per-row kernel, 1-3 the LDS-staged kernel with (T slots per chunk,
// RB rows per batch of loads, non-temporal loads, non-temporal stores):
// (4, 16, yes, no), (7, 32, yes, yes), (7, 32, no, yes); 6x6 blocks only
// the lane-per-row kernel.  At hex8 128^3 (57.1 M blocks, all three copies)
// they take 4.5, 2.7, 2.4 and 2.3 ms: temporal loads keep the cache lines a
// run shares with the row's next run.
int tacs_hip_sell_fill_nvariants(int b) { return b == 3 ? 4 : b == 6 ? 1 : 0; }
static int sell_fill_default_variant(int b) { return b == 3 ? 3 : 0; }

int tacs_hip_sell_fill_variant(int variant, int b, int nrows, const int* rowp, const int* cols,
                               const int* width, const long* coff, const double* A, double* v64,
                               float* v32, unsigned short* v16, int* cs, hipStream_t stream) {
  const int ns = (nrows + 63) / 64;
  if ((b != 3 && b != 6) || (!v64 && !v32 && !v16)) return TACS_HIP_ERR_ARG;
  if (variant < 0 || variant >= tacs_hip_sell_fill_nvariants(b)) return TACS_HIP_ERR_ARG;
  if (ns <= 0) return 0;
  if (b == 3 && variant > 0) {
    const int grid = std::min(ns, 65536);
#define FILL_LDS(TT, RB, NTL, NTS)                                                         \
  hipLaunchKernelGGL((sell_fill3_lds<TT, RB, NTL, NTS>), dim3(grid), dim3(64), 0, stream, ns,   \
                     nrows, rowp, cols, width, coff, A, v64, v32, v16, cs)
    switch (variant) {
      case 1: FILL_LDS(4, 16, true, false); break;
      case 2: FILL_LDS(7, 32, true, true); break;
      default: FILL_LDS(7, 32, false, true); break;
    }
#undef FILL_LDS
    return (int)hipGetLastError();
  }
  const int grid = tacs_grid_for(ns, 4, 65536);
  if (b == 3)
    hipLaunchKernelGGL(sell_fill<3>, dim3(grid), dim3(256), 0, stream, ns, nrows, rowp, cols,
                       width, coff, A, v64, v32, v16, cs);
  else
    hipLaunchKernelGGL(sell_fill<6>, dim3(grid), dim3(256), 0, stream, ns, nrows, rowp, cols,
                       width, coff, A, v64, v32, v16, cs);
  return (int)hipGetLastError();
}

int tacs_hip_sell_fill3(int b, int nrows, const int* rowp, const int* cols, const int* width,
                        const long* coff, const double* A, double* v64, float* v32,
                        unsigned short* v16, int* cs, hipStream_t stream) {
  return tacs_hip_sell_fill_variant(sell_fill_default_variant(b), b, nrows, rowp, cols, width,
                                    coff, A, v64, v32, v16, cs, stream);
}

int tacs_hip_sell_fill2(int b, int nrows, const int* rowp, const int* cols, const int* width,
                        const long* coff, const double* A, double* v64, float* v32, int* cs,
                        hipStream_t stream) {
  if (!v64 && !v32) return TACS_HIP_ERR_ARG;
  return tacs_hip_sell_fill3(b, nrows, rowp, cols, width, coff, A, v64, v32, NULL, cs, stream);
}

int tacs_hip_sell_fill(int b, int prec, int nrows, const int* rowp, const int* cols,
                       const int* width, const long* coff, const double* A, void* vs, int* cs,
                       hipStream_t stream) {
  if (prec != 64 && prec != 32) return TACS_HIP_ERR_ARG;
  return tacs_hip_sell_fill2(b, nrows, rowp, cols, width, coff, A,
                             prec == 64 ? (double*)vs : NULL, prec == 32 ? (float*)vs : NULL, cs,
                             stream);
}

// Variants of the sliced ELL product (all give the same sums): 3x3 blocks
// U in {1, 2, 4, 8} with (variants 0-3) or without (4-7) non-temporal value
// loads; 6x6 blocks U in {1, 2}, with (0, 1) or without (2, 3).
int tacs_hip_sell_spmv_nvariants(int b) { return b == 3 ? 8 : b == 6 ? 4 : 0; }

int tacs_hip_sell_spmv_variant(int variant, int b, int prec_vals, int prec_vecs, int nrows,
                               const int* width, const long* coff, const int* cs, const void* vs,
                               const void* x, void* y, hipStream_t stream) {
  const int ns = (nrows + 63) / 64;
  const int kind = (prec_vals == 64 && prec_vecs == 64)   ? 0
                   : (prec_vals == 32 && prec_vecs == 64) ? 1
                   : (prec_vals == 32 && prec_vecs == 32) ? 2
                                                          : -1;
  if ((b != 3 && b != 6) || kind < 0) return TACS_HIP_ERR_ARG;
  if (variant < 0 || variant >= tacs_hip_sell_spmv_nvariants(b)) return TACS_HIP_ERR_ARG;
  if (ns <= 0) return 0;
  const int grid = tacs_grid_for(ns, 4, 65536);
#define SELL_SPMV(BB, U, NT)                                                                    \
  do {                                                                                          \
    if (kind == 0)                                                                              \
      hipLaunchKernelGGL((sell_spmv<BB, U, double, double, NT>), dim3(grid), dim3(256), 0, stream, \
                         ns, nrows, width, coff, cs, (const double*)vs, (const double*)x,      \
                         (double*)y);                                                           \
    else if (kind == 1)                                                                         \
      hipLaunchKernelGGL((sell_spmv<BB, U, float, double, NT>), dim3(grid), dim3(256), 0, stream, \
                         ns, nrows, width, coff, cs, (const float*)vs, (const double*)x,       \
                         (double*)y);                                                           \
    else                                                                                        \
      hipLaunchKernelGGL((sell_spmv<BB, U, float, float, NT>), dim3(grid), dim3(256), 0, stream, \
                         ns, nrows, width, coff, cs, (const float*)vs, (const float*)x,        \
                         (float*)y);                                                            \
  } while (0)
  if (b == 3) {
    switch (variant) {
      case 0: SELL_SPMV(3, 1, true); break;
      case 1: SELL_SPMV(3, 2, true); break;
      case 2: SELL_SPMV(3, 4, true); break;
      case 3: SELL_SPMV(3, 8, true); break;
      case 4: SELL_SPMV(3, 1, false); break;
      case 5: SELL_SPMV(3, 2, false); break;
      case 6: SELL_SPMV(3, 4, false); break;
      default: SELL_SPMV(3, 8, false); break;
    }
  } else {
    switch (variant) {
      case 0: SELL_SPMV(6, 1, true); break;
      case 1: SELL_SPMV(6, 2, true); break;
      case 2: SELL_SPMV(6, 1, false); break;
      default: SELL_SPMV(6, 2, false); break;
    }
  }
#undef SELL_SPMV
  return (int)hipGetLastError();
}

static int sell_cheb_launch(int variant, int b, int prec_vals, int nrows, const int* width,
                            const long* coff, const int* cs, const void* vs, const double* z,
                            const SellCheb ch, hipStream_t stream);

int tacs_hip_sell_cheb_step(int variant, int b, int prec_vals, int nrows, const int* width,
                            const long* coff, const int* cs, const void* vs, const double* z,
                            double* zout, const double* rhs, const void* Dinv, double* d,
                            double c1, double c2, hipStream_t stream) {
  if (z == zout) return TACS_HIP_ERR_ARG;
  return sell_cheb_launch(variant, b, prec_vals, nrows, width, coff, cs, vs, z,
                          SellCheb{Dinv, rhs, d, zout, c1, c2, NULL}, stream);
}

int tacs_hip_sell_cheb_step3(int variant, int b, int prec_vals, int nrows, const int* width,
                             const long* coff, const int* cs, const void* vs, const double* z,
                             double* zprev, const double* rhs, const void* Dinv, double c1,
                             double c2, hipStream_t stream) {
  if (!zprev || z == zprev) return TACS_HIP_ERR_ARG;
  return sell_cheb_launch(variant, b, prec_vals, nrows, width, coff, cs, vs, z,
                          SellCheb{Dinv, rhs, NULL, zprev, c1, c2, zprev}, stream);
}

static int sell_cheb_launch(int variant, int b, int prec_vals, int nrows, const int* width,
                            const long* coff, const int* cs, const void* vs, const double* z,
                            const SellCheb ch, hipStream_t stream) {
  const int ns = (nrows + 63) / 64;
  if ((b != 3 && b != 6) || (prec_vals != 64 && prec_vals != 32 && prec_vals != 16))
    return TACS_HIP_ERR_ARG;
  if (variant < 0 || variant >= tacs_hip_sell_spmv_nvariants(b)) return TACS_HIP_ERR_ARG;
  if (ns <= 0) return 0;
  const int grid = tacs_grid_for(ns, 4, 65536);
#define SELL_CHEB(BB, U, NT)                                                                     \
  do {                                                                                           \
    if (prec_vals == 64)                                                                         \
      hipLaunchKernelGGL((sell_spmv<BB, U, double, double, NT, true>), dim3(grid), dim3(256), 0, \
                         stream, ns, nrows, width, coff, cs, (const double*)vs, z,              \
                         (double*)NULL, ch);                                                     \
    else if (prec_vals == 32)                                                                    \
      hipLaunchKernelGGL((sell_spmv<BB, U, float, double, NT, true>), dim3(grid), dim3(256), 0,  \
                         stream, ns, nrows, width, coff, cs, (const float*)vs, z,               \
                         (double*)NULL, ch);                                                     \
    else                                                                                         \
      hipLaunchKernelGGL((sell_spmv<BB, U, unsigned short, double, NT, true>), dim3(grid),      \
                         dim3(256), 0, stream, ns, nrows, width, coff, cs,                      \
                         (const unsigned short*)vs, z, (double*)NULL, ch);                      \
  } while (0)
  if (b == 3) {
    switch (variant) {
      case 0: SELL_CHEB(3, 1, true); break;
      case 1: SELL_CHEB(3, 2, true); break;
      case 2: SELL_CHEB(3, 4, true); break;
      case 3: SELL_CHEB(3, 8, true); break;
      case 4: SELL_CHEB(3, 1, false); break;
      case 5: SELL_CHEB(3, 2, false); break;
      case 6: SELL_CHEB(3, 4, false); break;
      default: SELL_CHEB(3, 8, false); break;
    }
  } else {
    switch (variant) {
      case 0: SELL_CHEB(6, 1, true); break;
      case 1: SELL_CHEB(6, 2, true); break;
      case 2: SELL_CHEB(6, 1, false); break;
      default: SELL_CHEB(6, 2, false); break;
    }
  }
#undef SELL_CHEB
  return (int)hipGetLastError();
}

int tacs_hip_sell_spmv(int b, int prec_vals, int prec_vecs, int nrows, const int* width,
                       const long* coff, const int* cs, const void* vs, const void* x, void* y,
                       hipStream_t stream) {
  const int kind = (prec_vals == 64 && prec_vecs == 64) ? 0 : prec_vecs == 64 ? 1 : 2;
  return tacs_hip_sell_spmv_variant(tacs_tune_variant(TACS_TUNE_SELL_SPMV, 10 * b + kind, nrows),
                                    b, prec_vals, prec_vecs, nrows, width, coff, cs, vs, x, y,
                                    stream);
}

int tacs_hip_bsr_find_diag(int nrows, const int* rowp, const int* cols, int* diag,
                           hipStream_t stream) {
  if (nrows <= 0) return 0;
  hipLaunchKernelGGL(bsr_find_diag, dim3((nrows + 255) / 256), dim3(256), 0, stream, nrows, rowp,
                     cols, diag);
  return (int)hipGetLastError();
}

int tacs_hip_bsr_diag_inverse(int b, int nrows, const int* diag, const double* A, double* Dinv,
                              float* Dinv32, int* status, hipStream_t stream) {
  if (nrows <= 0) return 0;
  dim3 g((nrows + 127) / 128), t(128);
  if (b == 3)
    hipLaunchKernelGGL(bsr_diag_inverse<3>, g, t, 0, stream, nrows, diag, A, Dinv, Dinv32, status);
  else if (b == 6)
    hipLaunchKernelGGL(bsr_diag_inverse<6>, g, t, 0, stream, nrows, diag, A, Dinv, Dinv32, status);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_bsr_diag_apply(int b, int nrows, const double* Dinv, const double* x, double* y,
                            hipStream_t stream) {
  if (nrows <= 0) return 0;
  dim3 g((nrows + 255) / 256), t(256);
  if (b == 3)
    hipLaunchKernelGGL((bsr_diag_apply<3, double, double>), g, t, 0, stream, nrows, Dinv, x, y);
  else if (b == 6)
    hipLaunchKernelGGL((bsr_diag_apply<6, double, double>), g, t, 0, stream, nrows, Dinv, x, y);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_bsr_diag_apply_f32(int b, int nrows, const float* Dinv, const float* x, float* y,
                                hipStream_t stream) {
  if (nrows <= 0) return 0;
  dim3 g((nrows + 255) / 256), t(256);
  if (b == 3)
    hipLaunchKernelGGL((bsr_diag_apply<3, float, float>), g, t, 0, stream, nrows, Dinv, x, y);
  else if (b == 6)
    hipLaunchKernelGGL((bsr_diag_apply<6, float, float>), g, t, 0, stream, nrows, Dinv, x, y);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_bsr_zero_rows(int b, int nbc, const int* nodes, const int* vars, const int* rowp,
                           const int* cols, double* A, int ident, hipStream_t stream) {
  if (nbc <= 0) return 0;
  hipLaunchKernelGGL(bsr_zero_rows, dim3((nbc + 3) / 4), dim3(256), 0, stream, b, nbc, nodes, vars,
                     rowp, cols, A, ident);
  return (int)hipGetLastError();
}

int tacs_hip_convert_f64_f32(size_t n, const double* x, float* y, hipStream_t stream) {
  if (n == 0) return 0;
  int grid = (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
  hipLaunchKernelGGL(convert_f64_to_f32, dim3(grid), dim3(256), 0, stream, n, x, y);
  return (int)hipGetLastError();
}

int tacs_hip_bsr_zero_identity_rows(int b, int nrows, const int* rowp, const int* cols,
                                    const double* A, double* x, hipStream_t stream) {
  if (b < 1 || b > 8) return TACS_HIP_ERR_ARG;
  const size_t n = (size_t)nrows * b;
  if (n == 0) return 0;
  hipLaunchKernelGGL(bsr_zero_identity_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     stream, b, nrows, rowp, cols, A, x);
  return (int)hipGetLastError();
}

}  // extern "C"
