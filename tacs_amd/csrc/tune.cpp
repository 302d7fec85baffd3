// Tuning table of the kernels that come in several launch variants giving the
// same results (the reference's counterpart is the block-size-specialised
// kernel choice of BCSRMat::initBlockImpl, src/bpmat/BCSRMat.cpp:1028, made
// once from the block size; here the choice also depends on the problem
// size and is filled from measurements).  An entry maps (kernel, shape, size
// class) to a variant; -1 means the built-in default.  Shapes: sliced ELL
// product 10 b + kind (kind 0: FP64 values and vectors, 1: FP32 values with
// FP64 vectors, 2: FP32 values and vectors); multi right-hand-side product
// 100 (FP32 values) or 0 (FP64 values) + R; solid Jacobian the element order.
// Size classes by rows: 0 below 2^15, 1 below 2^19, 2 from 2^19.
#include <mutex>

#include "tacs_hip_common.h"

namespace {
constexpr int KMAX = 3, SMAX = 256, CMAX = 3;
signed char g_table[KMAX][SMAX][CMAX];
std::once_flag g_once;
std::mutex g_mu;

void init_table() {
  for (int k = 0; k < KMAX; k++)
    for (int s = 0; s < SMAX; s++)
      for (int c = 0; c < CMAX; c++) g_table[k][s][c] = -1;
}

// built-in defaults (measured on 128^3 hex8 and 512^2 Quad4 meshes)
int default_variant(int kernel, int shape) {
  if (kernel == TACS_TUNE_SELL_SPMV) {
    if (shape == 30) return 1;                // 3x3, FP64 values: U = 2
    if (shape == 31 || shape == 32) return 2;  // 3x3, FP32 values: U = 4
    return 0;                                  // 6x6: U = 1
  }
  if (kernel == TACS_TUNE_JACOBIAN && shape == 3) return 4;  // 27-node: element-block form
  return 0;
}
}  // namespace

extern "C" {

int tacs_tune_size_class(long rows) { return rows < (1L << 15) ? 0 : rows < (1L << 19) ? 1 : 2; }

int tacs_tune_variant(int kernel, int shape, long rows) {
  std::call_once(g_once, init_table);
  if (kernel < 0 || kernel >= KMAX || shape < 0 || shape >= SMAX) return 0;
  const int v = g_table[kernel][shape][tacs_tune_size_class(rows)];
  return v >= 0 ? v : default_variant(kernel, shape);
}

int tacs_tune_set(int kernel, int shape, int size_class, int variant) {
  std::call_once(g_once, init_table);
  if (kernel < 0 || kernel >= KMAX || shape < 0 || shape >= SMAX || size_class < 0 ||
      size_class >= CMAX || variant < -1 || variant > 127)
    return TACS_HIP_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  g_table[kernel][shape][size_class] = (signed char)variant;
  return 0;
}

int tacs_tune_get(int kernel, int shape, int size_class) {
  std::call_once(g_once, init_table);
  if (kernel < 0 || kernel >= KMAX || shape < 0 || shape >= SMAX || size_class < 0 ||
      size_class >= CMAX)
    return -1;
  return g_table[kernel][shape][size_class];
}

void tacs_tune_reset(void) {
  std::call_once(g_once, init_table);
  std::lock_guard<std::mutex> lk(g_mu);
  init_table();
}

}  // extern "C"
