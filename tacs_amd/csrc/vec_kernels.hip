// gfx950 vector kernels for TACSBVec and the Krylov solvers: axpy/scale,
// deterministic two-launch reductions whose results stay in device memory
// (so Arnoldi steps never wait on the host), the fused modified-Gram-Schmidt step, the
// classical-Gram-Schmidt multi-dot / update pair, and the Givens update of
// the Hessenberg column (GMRES::solve, src/bpmat/KSM.cpp:785-958).
#include <mutex>

#include "tacs_hip_common.h"

#define RED_BLOCKS 2048
#define RED_THREADS 256
#define RED_MAXK 32

// Blocks of a reduction over n entries: about 8 entries per thread, at most
// RED_BLOCKS (a function of n only, so results are reproducible).
static inline int red_blocks(size_t n) {
  size_t b = n / 2048;
  if (b < 1) b = 1;
  if (b > RED_BLOCKS) b = RED_BLOCKS;
  return (int)b;
}

// Reduction scratch: per-block partial sums.  Every reduction is two
// launches: the first writes one partial per block with ordinary stores, the
// second (one block) combines the partials in block order.  The kernel
// boundary makes the partials visible; no atomics and no device-scope fences
// (a device-scope release writes back the L2 of the XCD) are needed, and the
// result is deterministic.
struct RedScratch {
  double partial[RED_BLOCKS * RED_MAXK];
};

__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; w++) t += sh[w];
  return t;
}

// partial[b] = sum over the block's share of w.q; optionally q -= (*hprev)
// wprev first (fused MGS step: ModifiedGramSchmidt KSM.cpp:526).
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void dot_partial(size_t n, const T* __restrict__ wprev,
                                                           const double* __restrict__ hprev,
                                                           T* q, const T* w, RedScratch* rs) {
  __shared__ double sh[RED_THREADS / 64];
  double acc = 0.0, acc2 = 0.0;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  if (wprev) {
    const double h = *hprev;
    if (w == q) {  // q -= h*wprev, then |q|^2
      for (; i < n; i += stride) {
        T qi = (T)((double)q[i] - h * (double)wprev[i]);
        q[i] = qi;
        acc += (double)qi * (double)qi;
      }
    } else {
      for (; i < n; i += stride) {
        T qi = (T)((double)q[i] - h * (double)wprev[i]);
        q[i] = qi;
        acc += (double)w[i] * (double)qi;
      }
    }
  } else {
    for (; i + stride < n; i += 2 * stride) {
      acc += (double)w[i] * (double)q[i];
      acc2 += (double)w[i + stride] * (double)q[i + stride];
    }
    if (i < n) acc += (double)w[i] * (double)q[i];
  }
  double bs = block_sum(acc + acc2, sh);
  if (threadIdx.x == 0) rs->partial[blockIdx.x] = bs;
}

// out[0] = sum of nb partials (sqrt if take_sqrt; out_inv = 1/out)
__global__ __launch_bounds__(RED_THREADS) void dot_finalize(int nb, const RedScratch* rs,
                                                            double* __restrict__ out,
                                                            int take_sqrt,
                                                            double* __restrict__ out_inv) {
  __shared__ double sh[RED_THREADS / 64];
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) t += rs->partial[b];
  t = block_sum(t, sh);
  if (threadIdx.x == 0) {
    double r = take_sqrt ? sqrt(t) : t;
    out[0] = r;
    if (out_inv) out_inv[0] = 1.0 / r;
  }
}

// Multi-dot partials: partial[j][b] = block b's share of W_j . q (classical
// GS).  Each thread keeps k partial sums; one block reduction phase.
template <typename T, int KMAX>
__global__ __launch_bounds__(RED_THREADS) void mdot_partial(size_t n, int k, const T* __restrict__ W,
                                                            size_t ldw, const T* __restrict__ q,
                                                            RedScratch* rs) {
  __shared__ double sh[RED_THREADS / 64][KMAX];
  double acc[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; j++) acc[j] = 0.0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double qi = (double)q[i];
#pragma unroll
    for (int j = 0; j < KMAX; j++)
      if (j < k) acc[j] += (double)W[(size_t)j * ldw + i] * qi;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < KMAX; j++) {
    if (j < k) {
      double v = wave_sum(acc[j]);
      if (lane == 0) sh[wid][j] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < k) {
    double t = 0.0;
    for (int w = 0; w < RED_THREADS / 64; w++) t += sh[w][threadIdx.x];
    rs->partial[(size_t)threadIdx.x * RED_BLOCKS + blockIdx.x] = t;
  }
}

// h[j] (+)= sum_b partial[j][b]; wave w handles vectors w, w + 4, ...
__global__ __launch_bounds__(RED_THREADS) void mdot_finalize(int nb, int k, const RedScratch* rs,
                                                             double* __restrict__ h,
                                                             int accumulate) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int j = wid; j < k; j += RED_THREADS / 64) {
    const double* pj = &rs->partial[(size_t)j * RED_BLOCKS];
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    int bb = lane;
    for (; bb + 192 < nb; bb += 256) {
      t0 += pj[bb];
      t1 += pj[bb + 64];
      t2 += pj[bb + 128];
      t3 += pj[bb + 192];
    }
    for (; bb < nb; bb += 64) t0 += pj[bb];
    double t = wave_sum((t0 + t1) + (t2 + t3));
    if (lane == 0) h[j] = accumulate ? h[j] + t : t;
  }
}

// q -= sum_j c[j] W_j  (c in device memory, k <= RED_MAXK)
template <typename T>
__global__ void cgs_update(size_t n, int k, const T* __restrict__ W, size_t ldw,
                           const double* __restrict__ c, const double* __restrict__ c2,
                           T* __restrict__ q) {
  __shared__ double cs[RED_MAXK];
  if (threadIdx.x < k) cs[threadIdx.x] = c[threadIdx.x] - (c2 ? c2[threadIdx.x] : 0.0);
  __syncthreads();
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double s = (double)q[i];
    for (int j = 0; j < k; j++) s -= cs[j] * (double)W[(size_t)j * ldw + i];
    q[i] = (T)s;
  }
}

// y = sum_j c[j] W_j (+ y if add)
template <typename T, typename TY>
__global__ void lincomb(size_t n, int k, const T* __restrict__ W, size_t ldw,
                        const double* __restrict__ c, TY* __restrict__ y, int add) {
  __shared__ double cs[512];
  for (int j = threadIdx.x; j < k; j += blockDim.x) cs[j] = c[j];
  __syncthreads();
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double s = add ? (double)y[i] : 0.0;
    for (int j = 0; j < k; j++) s += cs[j] * (double)W[(size_t)j * ldw + i];
    y[i] = (TY)s;
  }
}

// x *= s, s = *sp (or 1/ *sp if invert) or the host value when sp is null
template <typename T>
__global__ void scale_kernel(size_t n, T* __restrict__ x, const double* __restrict__ sp,
                             double sv, int invert) {
  double s = sp ? (invert ? 1.0 / sp[0] : sp[0]) : sv;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    x[i] = (T)(s * (double)x[i]);
}

// y = a*x + b*y
template <typename TX, typename TY>
__global__ void axpby_kernel(size_t n, double a, const TX* __restrict__ x, double b,
                             TY* __restrict__ y) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = (TY)(a * (double)x[i] + (b == 0.0 ? 0.0 : b * (double)y[i]));
}

template <typename T>
__global__ void fill_kernel(size_t n, T* __restrict__ x, T v) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = v;
}

// partial[b] = sum (op 0) or max (op 1) of the block's share of x
__global__ __launch_bounds__(RED_THREADS) void reduce_partial(size_t n, const double* __restrict__ x,
                                                              int op, RedScratch* rs) {
  __shared__ double sh[RED_THREADS / 64];
  double acc = op ? -1e300 : 0.0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc = op ? fmax(acc, x[i]) : acc + x[i];
  double bs;
  if (op) {
    acc = wave_max(acc);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = acc;
    __syncthreads();
    bs = sh[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) bs = fmax(bs, sh[w]);
  } else {
    bs = block_sum(acc, sh);
  }
  if (threadIdx.x == 0) rs->partial[blockIdx.x] = bs;
}

__global__ void reduce_finalize(int nb, int op, const RedScratch* rs, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double t = rs->partial[0];
    for (int b = 1; b < nb; b++) t = op ? fmax(t, rs->partial[b]) : t + rs->partial[b];
    out[0] = t;
  }
}

// Givens update of Hessenberg column i (packed column-major, column i at
// H + i*(m+1)), as GMRES::solve KSM.cpp:864-895.  st[0] = |res[i+1]|.
__global__ void givens_kernel(int i, int m, double* __restrict__ H, double* __restrict__ Qc,
                              double* __restrict__ Qs, double* __restrict__ res,
                              double* __restrict__ st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* h = H + (size_t)i * (m + 1);
  for (int k = 0; k < i; k++) {
    double h1 = h[k], h2 = h[k + 1];
    h[k] = h1 * Qc[k] + h2 * Qs[k];
    h[k + 1] = -h1 * Qs[k] + h2 * Qc[k];
  }
  double h1 = h[i], h2 = h[i + 1];
  double sq = sqrt(h1 * h1 + h2 * h2);
  Qc[i] = h1 / sq;
  Qs[i] = h2 / sq;
  h[i] = h1 * Qc[i] + h2 * Qs[i];
  h[i + 1] = -h1 * Qs[i] + h2 * Qc[i];
  h1 = res[i];
  res[i] = h1 * Qc[i];
  res[i + 1] = -h1 * Qs[i];
  st[0] = fabs(res[i + 1]);
}

// Back-substitution of the upper-triangular system H y = res (KSM.cpp:912-918)
__global__ void backsolve_kernel(int niters, int m, const double* __restrict__ H,
                                 double* __restrict__ res) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = niters - 1; i >= 0; i--) {
    for (int j = i + 1; j < niters; j++) res[i] -= H[(size_t)j * (m + 1) + i] * res[j];
    res[i] /= H[(size_t)i * (m + 1) + i];
  }
}

// Copy of a device scalar with sign/scale: dst = s * src
__global__ void scalar_op(const double* __restrict__ src, double s, double* __restrict__ dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dst[0] = s * src[0];
}

// out[:, j] = sum_i W[:, i] P[i, j], j < nout <= 32 (P column-major nin x nout,
// staged in LDS as FP32): the basis change V_new = V P of a deflated restart.
// One thread per row, 32 accumulators.
__global__ __launch_bounds__(256) void basis_combine_kernel(size_t n, int nin,
                                                            const float* __restrict__ W,
                                                            size_t ldw, const double* __restrict__ P,
                                                            int nout, float* __restrict__ out,
                                                            size_t ldo) {
  extern __shared__ float sP[];  // [nin][32]
  for (int t = threadIdx.x; t < nin * 32; t += blockDim.x) {
    int i = t / 32, j = t % 32;
    sP[t] = (j < nout) ? (float)P[i + (size_t)nin * j] : 0.0f;
  }
  __syncthreads();
  for (size_t row = (size_t)blockIdx.x * blockDim.x + threadIdx.x; row < n;
       row += (size_t)gridDim.x * blockDim.x) {
    float acc[32];
#pragma unroll
    for (int j = 0; j < 32; j++) acc[j] = 0.0f;
    for (int i = 0; i < nin; i++) {
      float w = W[(size_t)i * ldw + row];
#pragma unroll
      for (int j = 0; j < 32; j++) acc[j] += w * sP[i * 32 + j];
    }
#pragma unroll
    for (int j = 0; j < 32; j++)
      if (j < nout) out[(size_t)j * ldo + row] = acc[j];
  }
}

// FP64 basis change: out[:, j] (+)= sum_i W[:, i] P[i, j], j < nout <= 32.
__global__ __launch_bounds__(256) void basis_combine64_kernel(size_t n, int nin,
                                                              const double* __restrict__ W,
                                                              size_t ldw,
                                                              const double* __restrict__ P,
                                                              int nout, double* __restrict__ out,
                                                              size_t ldo, int add) {
  extern __shared__ double sP64[];  // [nin][32]
  for (int t = threadIdx.x; t < nin * 32; t += blockDim.x) {
    int i = t / 32, j = t % 32;
    sP64[t] = (j < nout) ? P[i + (size_t)nin * j] : 0.0;
  }
  __syncthreads();
  for (size_t row = (size_t)blockIdx.x * blockDim.x + threadIdx.x; row < n;
       row += (size_t)gridDim.x * blockDim.x) {
    double acc[32];
#pragma unroll
    for (int j = 0; j < 32; j++) acc[j] = 0.0;
    for (int i = 0; i < nin; i++) {
      double w = W[(size_t)i * ldw + row];
#pragma unroll
      for (int j = 0; j < 32; j++) acc[j] += w * sP64[i * 32 + j];
    }
#pragma unroll
    for (int j = 0; j < 32; j++)
      if (j < nout) {
        size_t o = (size_t)j * ldo + row;
        out[o] = add ? out[o] + acc[j] : acc[j];
      }
  }
}

// Tall-skinny product out = X^T Y (X: n x p, Y: n x q, column-major with
// leading dimension n; p * q <= TSG_MAXPQ): each block stages 32-row tiles of
// every column in LDS and accumulates its share of the p x q sums; the
// partial sums are added in a second, deterministic launch.
#define TSG_ROWS 32
#define TSG_BLOCKS 512
#define TSG_MAXPQ 4096
__global__ __launch_bounds__(256) void tsgemm_partial(size_t n, int p, const double* __restrict__ X,
                                                      int q, const double* __restrict__ Y,
                                                      double* __restrict__ part) {
  extern __shared__ double tile[];  // [(p + q)][TSG_ROWS]
  const int pq = p * q;
  double acc[TSG_MAXPQ / 256];
#pragma unroll
  for (int t = 0; t < TSG_MAXPQ / 256; t++) acc[t] = 0.0;
  for (size_t r0 = (size_t)blockIdx.x * TSG_ROWS; r0 < n; r0 += (size_t)gridDim.x * TSG_ROWS) {
    const int nr = (int)((n - r0) < TSG_ROWS ? (n - r0) : TSG_ROWS);
    for (int t = threadIdx.x; t < (p + q) * TSG_ROWS; t += blockDim.x) {
      int v = t / TSG_ROWS, r = t % TSG_ROWS;
      double val = 0.0;
      if (r < nr) val = (v < p) ? X[(size_t)v * n + r0 + r] : Y[(size_t)(v - p) * n + r0 + r];
      tile[t] = val;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TSG_MAXPQ / 256; t++) {
      int e = threadIdx.x + 256 * t;
      if (e < pq) {
        int i = e % p, j = e / p;
        const double* xi = &tile[i * TSG_ROWS];
        const double* yj = &tile[(p + j) * TSG_ROWS];
        double s = 0.0;
        for (int r = 0; r < TSG_ROWS; r++) s += xi[r] * yj[r];
        acc[t] += s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < TSG_MAXPQ / 256; t++) {
    int e = threadIdx.x + 256 * t;
    if (e < pq) part[(size_t)blockIdx.x * pq + e] = acc[t];
  }
}

__global__ void tsgemm_finalize(int nb, int pq, const double* __restrict__ part,
                                double* __restrict__ out) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pq) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += part[(size_t)b * pq + e];
  out[e] = s;
}

static inline int vgrid(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

// Reduction partials of the allocation-free multi-dot (tacs_hip_mdot_noalloc):
// RED_BLOCKS x RED_MAXK doubles (512 KiB), one copy on every GPU the library
// runs on, held until the process ends.  Host code never passes this name to
// a kernel: the device address of each GPU's copy is resolved once, with that
// GPU current, by scratch_ptr.
static __device__ __align__(256) double scratch[RED_BLOCKS * RED_MAXK];
static_assert(sizeof(scratch) == sizeof(RedScratch), "scratch holds one RedScratch");

int tacs_set_error(const char* fmt, ...);

#define SCRATCH_MAX_DEVICES 64
static void* g_scratch_dev[SCRATCH_MAX_DEVICES];
static std::mutex g_scratch_mu;

// Device address of the current GPU's copy of scratch (resolved on the first
// call made with that GPU current, then cached).
static int scratch_ptr(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    tacs_set_error("tacs_hip_mdot_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    return (int)e;
  }
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
    tacs_set_error("tacs_hip_mdot_noalloc: device %d beyond the %d this library tracks", dev,
                   SCRATCH_MAX_DEVICES);
    return TACS_HIP_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = NULL;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      tacs_set_error("tacs_hip_mdot_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                     hipGetErrorString(e));
      return e != hipSuccess ? (int)e : TACS_HIP_ERR_ARG;
    }
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

static int mdot_launch(int prec, size_t n, int k, const void* W, size_t ldw, const void* q,
                       RedScratch* rs, double* h, int accumulate, hipStream_t stream) {
  const int nb = red_blocks(n);
#define MD_LAUNCH(T, KM)                                                                       \
  hipLaunchKernelGGL((mdot_partial<T, KM>), dim3(nb), dim3(RED_THREADS), 0, stream, n, k,      \
                     (const T*)W, ldw, (const T*)q, rs)
  if (prec == 64) {
    if (k <= 8) MD_LAUNCH(double, 8);
    else if (k <= 16) MD_LAUNCH(double, 16);
    else MD_LAUNCH(double, 32);
  } else {
    if (k <= 8) MD_LAUNCH(float, 8);
    else if (k <= 16) MD_LAUNCH(float, 16);
    else MD_LAUNCH(float, 32);
  }
#undef MD_LAUNCH
  hipLaunchKernelGGL(mdot_finalize, dim3(1), dim3(RED_THREADS), 0, stream, nb, k, rs, h,
                     accumulate);
  return (int)hipGetLastError();
}

// x[i] = a value in [-1, 1) from a 64-bit mix of (seed, i) (splitmix64 finaliser):
// a reproducible start vector with every frequency present.
__global__ void fill_hash_kernel(size_t n, double* __restrict__ x, unsigned long long seed) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

extern "C" {

size_t tacs_hip_red_scratch_bytes() { return sizeof(RedScratch); }

int tacs_hip_dot(int prec, size_t n, const void* wprev, const double* hprev, void* q,
                 const void* w, void* scratch, double* out, int take_sqrt, double* out_inv,
                 hipStream_t stream) {
  RedScratch* rs = (RedScratch*)scratch;
  const int nb = red_blocks(n);
  if (prec == 64)
    hipLaunchKernelGGL(dot_partial<double>, dim3(nb), dim3(RED_THREADS), 0, stream, n,
                       (const double*)wprev, hprev, (double*)q, (const double*)w, rs);
  else
    hipLaunchKernelGGL(dot_partial<float>, dim3(nb), dim3(RED_THREADS), 0, stream, n,
                       (const float*)wprev, hprev, (float*)q, (const float*)w, rs);
  hipLaunchKernelGGL(dot_finalize, dim3(1), dim3(RED_THREADS), 0, stream, nb, rs, out, take_sqrt,
                     out_inv);
  return (int)hipGetLastError();
}

// h[0..k) = W^T q, or h += W^T q when accumulate
int tacs_hip_mdot(int prec, size_t n, int k, const void* W, size_t ldw, const void* q,
                  void* scratch, double* h, int accumulate, hipStream_t stream) {
  if (k > RED_MAXK || k < 1) return TACS_HIP_ERR_ARG;
  return mdot_launch(prec, n, k, W, ldw, q, (RedScratch*)scratch, h, accumulate, stream);
}

// The same multi-dot with its partials in this file's static device array:
// allocates nothing, for callers that cannot allocate (graph capture,
// serving loops).
int tacs_hip_mdot_noalloc(int prec, size_t n, int k, const void* W, size_t ldw, const void* q,
                          double* h, int accumulate, hipStream_t stream) {
  if (k > RED_MAXK || k < 1) {
    tacs_set_error("tacs_hip_mdot_noalloc: k = %d vectors; its static scratch holds 1 to %d", k,
                   RED_MAXK);
    return TACS_HIP_ERR_ARG;
  }
  if (prec != 64 && prec != 32) {
    tacs_set_error("tacs_hip_mdot_noalloc: prec %d (64 or 32)", prec);
    return TACS_HIP_ERR_ARG;
  }
  void* rs = NULL;
  int e = scratch_ptr(&rs);
  if (e) return e;
  return mdot_launch(prec, n, k, W, ldw, q, (RedScratch*)rs, h, accumulate, stream);
}

int tacs_hip_cgs_update(int prec, size_t n, int k, const void* W, size_t ldw, const double* c,
                        const double* c2, void* q, hipStream_t stream) {
  if (k > RED_MAXK) return TACS_HIP_ERR_ARG;
  if (prec == 64)
    hipLaunchKernelGGL(cgs_update<double>, dim3(vgrid(n)), dim3(256), 0, stream, n, k,
                       (const double*)W, ldw, c, c2, (double*)q);
  else
    hipLaunchKernelGGL(cgs_update<float>, dim3(vgrid(n)), dim3(256), 0, stream, n, k,
                       (const float*)W, ldw, c, c2, (float*)q);
  return (int)hipGetLastError();
}

// y (+)= sum_j c[j] W_j ; prec_w: basis precision, prec_y: output precision
int tacs_hip_lincomb(int prec_w, int prec_y, size_t n, int k, const void* W, size_t ldw,
                     const double* c, void* y, int add, hipStream_t stream) {
  if (k > 512) return TACS_HIP_ERR_ARG;
  if (prec_w == 64 && prec_y == 64)
    hipLaunchKernelGGL((lincomb<double, double>), dim3(vgrid(n)), dim3(256), 0, stream, n, k,
                       (const double*)W, ldw, c, (double*)y, add);
  else if (prec_w == 32 && prec_y == 32)
    hipLaunchKernelGGL((lincomb<float, float>), dim3(vgrid(n)), dim3(256), 0, stream, n, k,
                       (const float*)W, ldw, c, (float*)y, add);
  else if (prec_w == 32 && prec_y == 64)
    hipLaunchKernelGGL((lincomb<float, double>), dim3(vgrid(n)), dim3(256), 0, stream, n, k,
                       (const float*)W, ldw, c, (double*)y, add);
  else
    return TACS_HIP_ERR_ARG;
  return (int)hipGetLastError();
}

int tacs_hip_scale(int prec, size_t n, void* x, const double* sp, double sv, int invert,
                   hipStream_t stream) {
  if (n == 0) return 0;
  if (prec == 64)
    hipLaunchKernelGGL(scale_kernel<double>, dim3(vgrid(n)), dim3(256), 0, stream, n, (double*)x,
                       sp, sv, invert);
  else
    hipLaunchKernelGGL(scale_kernel<float>, dim3(vgrid(n)), dim3(256), 0, stream, n, (float*)x,
                       sp, sv, invert);
  return (int)hipGetLastError();
}

// y = a*x + b*y with precisions px (of x) and py (of y)
int tacs_hip_axpby(int px, int py, size_t n, double a, const void* x, double b, void* y,
                   hipStream_t stream) {
  if (n == 0) return 0;
  if (px == 64 && py == 64)
    hipLaunchKernelGGL((axpby_kernel<double, double>), dim3(vgrid(n)), dim3(256), 0, stream, n, a,
                       (const double*)x, b, (double*)y);
  else if (px == 32 && py == 32)
    hipLaunchKernelGGL((axpby_kernel<float, float>), dim3(vgrid(n)), dim3(256), 0, stream, n, a,
                       (const float*)x, b, (float*)y);
  else if (px == 64 && py == 32)
    hipLaunchKernelGGL((axpby_kernel<double, float>), dim3(vgrid(n)), dim3(256), 0, stream, n, a,
                       (const double*)x, b, (float*)y);
  else
    hipLaunchKernelGGL((axpby_kernel<float, double>), dim3(vgrid(n)), dim3(256), 0, stream, n, a,
                       (const float*)x, b, (double*)y);
  return (int)hipGetLastError();
}

int tacs_hip_fill(int prec, size_t n, void* x, double v, hipStream_t stream) {
  if (n == 0) return 0;
  if (prec == 64)
    hipLaunchKernelGGL(fill_kernel<double>, dim3(vgrid(n)), dim3(256), 0, stream, n, (double*)x,
                       v);
  else if (prec == 32)
    hipLaunchKernelGGL(fill_kernel<float>, dim3(vgrid(n)), dim3(256), 0, stream, n, (float*)x,
                       (float)v);
  else
    hipLaunchKernelGGL(fill_kernel<int>, dim3(vgrid(n)), dim3(256), 0, stream, n, (int*)x, (int)v);
  return (int)hipGetLastError();
}

int tacs_hip_reduce(size_t n, const double* x, int op, void* scratch, double* out,
                    hipStream_t stream) {
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(reduce_partial, dim3(nb), dim3(RED_THREADS), 0, stream, n, x, op,
                     (RedScratch*)scratch);
  hipLaunchKernelGGL(reduce_finalize, dim3(1), dim3(64), 0, stream, nb, op,
                     (const RedScratch*)scratch, out);
  return (int)hipGetLastError();
}

int tacs_hip_givens(int i, int m, double* H, double* Qc, double* Qs, double* res, double* st,
                    hipStream_t stream) {
  hipLaunchKernelGGL(givens_kernel, dim3(1), dim3(64), 0, stream, i, m, H, Qc, Qs, res, st);
  return (int)hipGetLastError();
}

int tacs_hip_backsolve(int niters, int m, const double* H, double* res, hipStream_t stream) {
  hipLaunchKernelGGL(backsolve_kernel, dim3(1), dim3(64), 0, stream, niters, m, H, res);
  return (int)hipGetLastError();
}

int tacs_hip_basis_combine(size_t n, int nin, const float* W, size_t ldw, const double* P,
                           int nout, float* out, size_t ldo, hipStream_t stream) {
  if (n == 0 || nout == 0) return 0;
  if (nout > 32 || nin < 1 || nin > 512) return TACS_HIP_ERR_ARG;
  size_t lds = (size_t)nin * 32 * sizeof(float);
  hipLaunchKernelGGL(basis_combine_kernel, dim3(vgrid(n)), dim3(256), lds, stream, n, nin, W, ldw,
                     P, nout, out, ldo);
  return (int)hipGetLastError();
}

int tacs_hip_basis_combine64(size_t n, int nin, const double* W, size_t ldw, const double* P,
                             int nout, double* out, size_t ldo, int add, hipStream_t stream) {
  if (n == 0 || nout == 0 || nin == 0) return 0;
  if (nout > 32 || nin < 1 || nin > 256) return TACS_HIP_ERR_ARG;
  size_t lds = (size_t)nin * 32 * sizeof(double);
  hipLaunchKernelGGL(basis_combine64_kernel, dim3(vgrid(n)), dim3(256), lds, stream, n, nin, W,
                     ldw, P, nout, out, ldo, add);
  return (int)hipGetLastError();
}

size_t tacs_hip_tsgemm_scratch_bytes() { return (size_t)TSG_BLOCKS * TSG_MAXPQ * sizeof(double); }

int tacs_hip_tsgemm(size_t n, int p, const double* X, int q, const double* Y, double* scratch,
                    double* out, hipStream_t stream) {
  if (p <= 0 || q <= 0) return 0;
  if (p * q > TSG_MAXPQ || (size_t)(p + q) * TSG_ROWS * sizeof(double) > 65536)
    return TACS_HIP_ERR_ARG;
  size_t nb = (n + TSG_ROWS - 1) / TSG_ROWS;
  if (nb > TSG_BLOCKS) nb = TSG_BLOCKS;
  if (nb < 1) nb = 1;
  size_t lds = (size_t)(p + q) * TSG_ROWS * sizeof(double);
  hipLaunchKernelGGL(tsgemm_partial, dim3((int)nb), dim3(256), lds, stream, n, p, X, q, Y, scratch);
  hipLaunchKernelGGL(tsgemm_finalize, dim3((p * q + 255) / 256), dim3(256), 0, stream, (int)nb,
                     p * q, scratch, out);
  return (int)hipGetLastError();
}

int tacs_hip_scalar_op(const double* src, double s, double* dst, hipStream_t stream) {
  hipLaunchKernelGGL(scalar_op, dim3(1), dim3(64), 0, stream, src, s, dst);
  return (int)hipGetLastError();
}

int tacs_hip_fill_random(size_t n, double* x, unsigned long long seed, hipStream_t stream) {
  if (n == 0) return 0;
  size_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(fill_hash_kernel, dim3((unsigned)g), dim3(256), 0, stream, n, x, seed);
  return (int)hipGetLastError();
}

}  // extern "C"
