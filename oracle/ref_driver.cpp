// Thin C-ABI driver around the reference's own header-only CPU oracle
// (kernel/host_srgemm.h) and test-input generator (utils/init_mat.h).
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/ from the
// sources where they lie under the reference tree; it is never shipped and the
// GPU path never loads it.  Used to pin the C restatement (host_srgemm.c) and
// to write the golden fixtures under tests/golden/.
#include <cstdint>
#include <cstdlib>

#include "kernel/host_srgemm.h"
#include "utils/init_mat.h"

namespace {
template <typename T> using binop = T (*)(T, T);

// (addop, mulop) pairs as tests/cuAsr_kernel_test.cu:72,... passes them.
bool pick(int op, binop<float> &add, binop<float> &mul) {
  switch (op) {
  case 0: add = mmo::minimum<float>; mul = mmo::plus<float>; return true;
  case 1: add = mmo::plus<float>; mul = mmo::multiplies<float>; return true;
  case 2: add = mmo::maximum<float>; mul = mmo::plus<float>; return true;
  case 3: add = mmo::maximum<float>; mul = mmo::minimum<float>; return true;
  case 4: add = mmo::maximum<float>; mul = mmo::multiplies<float>; return true;
  case 5: add = mmo::minimum<float>; mul = mmo::multiplies<float>; return true;
  case 6: add = mmo::minimum<float>; mul = mmo::maximum<float>; return true;
  case 8: add = mmo::plus<float>; mul = mmo::minussquare<float>; return true;
  default: return false;
  }
}
} // namespace

extern "C" {

int ref_host_srgemm_f32(int op, int M, int N, int K, float *A, float *B,
                        float *C, float *D, int do_epilogue) {
  binop<float> add, mul;
  if (!pick(op, add, mul)) return -1;
  mmo::host_srgemm<float>(M, N, K, A, M, B, K, C, M, D, do_epilogue != 0, add, mul);
  return 0;
}

int ref_host_srgemm_i32(int M, int N, int K, int32_t *A, int32_t *B,
                        int32_t *C, int32_t *D, int do_epilogue) {
  mmo::host_srgemm<int>(M, N, K, A, M, B, K, C, M, D, do_epilogue != 0,
                        mmo::bin_or<int>, mmo::bin_and<int>);
  return 0;
}

void ref_rng_init_matrix(float *buf, int len, int seed) {
  rng_init_matrix(buf, len, seed);
}

// The shape/seed sequence tests/cuAsr_kernel_test.cu:416-428 draws from an
// unseeded glibc rand(): M, N, K, then two test_cases(...) calls, each with
// three rand() arguments of which only the first (the A seed) is used.  The
// call expressions below have the same form, so this compiler evaluates the
// arguments in the same order a build of that test does.
static void take(int *dst, int a, int b, int c) { dst[0] = a; (void)b; (void)c; }
void ref_test_sequence(int n_tests, int *out /* n_tests x 5: M N K s1_noepi s1_epi */) {
  srand(1);
  for (int t = 0; t < n_tests; ++t) {
    int M = rand() % 256;
    int N = rand() % 256;
    int K = rand() % 256;
    out[t * 5 + 0] = M;
    out[t * 5 + 1] = N;
    out[t * 5 + 2] = K;
    take(&out[t * 5 + 3], rand(), rand(), rand());
    take(&out[t * 5 + 4], rand(), rand(), rand());
  }
}
}
