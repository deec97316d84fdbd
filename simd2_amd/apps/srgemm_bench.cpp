// Micro-benchmark of the nine semiring GEMMs over the C ABI: the shape list
// and output lines of the reference's micro_bench/srgemm_bench.cu:140-152
// (`<op>: <ms> ms` per shape, mean of 20 back-to-back calls), timed with HIP
// events on the null stream.
//
//   simd2_srgemm_bench [--max-gops G]   skip shapes above G * 1e9 semiring ops
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/simd2/srgemm.h"

namespace {

constexpr int NUM_ITR = 20;

const char *kNames[9] = {"min_plus", "plus_mul", "max_plus", "max_min", "max_mul",
                         "min_mul",  "min_max",  "or_and",   "minus_square"};

void fill(std::vector<float> &v, uint64_t seed) {
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<float> d(0.0f, 1.0f);
  for (auto &x : v) x = d(g);
}

bool check(hipError_t e, const char *what) {
  if (e != hipSuccess) std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
  return e == hipSuccess;
}

void bench_case(int M, int N, int K) {
  std::printf("Running microbench on A = %dx%d and B = %dx%d\n", M, K, K, N);
  std::vector<float> A((size_t)M * K), B((size_t)K * N), C((size_t)M * N);
  fill(A, 1);
  fill(B, 3080);
  fill(C, 3090);
  float *dA, *dB, *dC, *dD;
  int *iA, *iB, *iC, *iD;
  if (!check(hipMalloc(&dA, A.size() * 4), "hipMalloc") ||
      !check(hipMalloc(&dB, B.size() * 4), "hipMalloc") ||
      !check(hipMalloc(&dC, C.size() * 4), "hipMalloc") ||
      !check(hipMalloc(&dD, C.size() * 4), "hipMalloc"))
    std::exit(1);
  check(hipMemcpy(dA, A.data(), A.size() * 4, hipMemcpyHostToDevice), "copy");
  check(hipMemcpy(dB, B.data(), B.size() * 4, hipMemcpyHostToDevice), "copy");
  check(hipMemcpy(dC, C.data(), C.size() * 4, hipMemcpyHostToDevice), "copy");
  // or_and on 0/1 ints, as the reference's int drop-in takes them
  std::vector<int> Ai(A.size()), Bi(B.size()), Ci(C.size(), 0);
  for (size_t i = 0; i < A.size(); ++i) Ai[i] = A[i] < 0.02f;
  for (size_t i = 0; i < B.size(); ++i) Bi[i] = B[i] < 0.02f;
  if (!check(hipMalloc(&iA, Ai.size() * 4), "hipMalloc") ||
      !check(hipMalloc(&iB, Bi.size() * 4), "hipMalloc") ||
      !check(hipMalloc(&iC, Ci.size() * 4), "hipMalloc") ||
      !check(hipMalloc(&iD, Ci.size() * 4), "hipMalloc"))
    std::exit(1);
  check(hipMemcpy(iA, Ai.data(), Ai.size() * 4, hipMemcpyHostToDevice), "copy");
  check(hipMemcpy(iB, Bi.data(), Bi.size() * 4, hipMemcpyHostToDevice), "copy");
  check(hipMemcpy(iC, Ci.data(), Ci.size() * 4, hipMemcpyHostToDevice), "copy");

  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "event");
  check(hipEventCreate(&e1), "event");
  for (int op = 0; op < 9; ++op) {
    // row-major A (MxK) times B (KxN) is the column-major call (N, M, K, B, A)
    auto call = [&]() {
      if (op == SIMD2_OR_AND)
        return cuasr_orand_srsgemm(N, M, K, iB, N, iA, K, iC, N, iD, false, nullptr);
      return simd2_srgemm(op, N, M, K, dB, N, dA, K, dC, N, dD, N, 0, nullptr);
    };
    int st = call();  // warm-up
    check(hipDeviceSynchronize(), "sync");
    check(hipEventRecord(e0, nullptr), "record");
    for (int i = 0; i < NUM_ITR && st == SIMD2_SUCCESS; ++i) st = call();
    check(hipEventRecord(e1, nullptr), "record");
    check(hipEventSynchronize(e1), "sync");
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    if (st != SIMD2_SUCCESS)
      std::printf("simd2_%s: status %d (%s)\n", kNames[op], st, simd2_status_string(st));
    else
      std::printf("simd2_%s: %g ms  (%.1f GOPS)\n", kNames[op], ms / NUM_ITR,
                  2.0 * M * N * K / (ms / NUM_ITR * 1e-3) / 1e9);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(dD);
  (void)hipFree(iA); (void)hipFree(iB); (void)hipFree(iC); (void)hipFree(iD);
}

}  // namespace

int main(int argc, char **argv) {
  double max_gops = 3000.0;  // default: up to 8192^3-sized problems
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], "--max-gops")) max_gops = std::atof(argv[i + 1]);
  const int shapes[][3] = {{1024, 1024, 1024},    {2048, 2048, 2048},   {4096, 4096, 4096},
                           {8192, 8192, 8192},    {16384, 16384, 16384}, {2048, 1024, 512},
                           {4096, 2048, 1024},    {8192, 4096, 2048},   {16384, 8192, 4096},
                           {32768, 16384, 4096},  {65536, 32768, 16384}};
  for (const auto &s : shapes) {
    const double gops = 2.0 * s[0] * s[1] * s[2] / 1e9;
    if (gops > max_gops) {
      std::printf("skipping %dx%dx%d (%.0f G ops > --max-gops %.0f)\n", s[0], s[1], s[2], gops,
                  max_gops);
      continue;
    }
    bench_case(s[0], s[1], s[2]);
  }
  return 0;
}
