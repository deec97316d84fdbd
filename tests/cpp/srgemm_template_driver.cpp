// Caller of the C++ surface <simd2/srgemm.hpp>, the way host code written
// against the reference's templates calls it: simd2::srgemm<AddOp, MulOp, T>
// for all nine semirings (kernel/host_srgemm.h:16-30 shape, column-major as
// kernel/srgemm.cuh), the comp_update(int*) overload (kernel/converge.cuh:71-85)
// and the __half overloads of tensor_srgemm / f2h_device / h2f_device
// (kernel/tensor_srgemm.cuh:197-218, kernel/precision.cuh:21-33).
//
// Inputs are generated here with a fixed 64-bit LCG; inputs and outputs are
// written as raw little-endian files into <dir> with a manifest
// (<dir>/manifest.txt, one line per array: name dtype rows cols), and the
// test that runs this program checks them against the oracle.
//
//   simd2_template_driver <dir>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include <simd2/srgemm.hpp>

namespace {

uint64_t g_state = 0x5eed5eedULL;
uint32_t next_u32() {
  g_state = g_state * 6364136223846793005ULL + 1442695040888963407ULL;
  return (uint32_t)(g_state >> 33);
}
int uniform_int(int lo, int hi) { return lo + (int)(next_u32() % (uint32_t)(hi - lo + 1)); }

std::string g_dir;
FILE *g_manifest = nullptr;

template <typename T>
void save(const std::string &name, const std::vector<T> &v, int rows, int cols,
          const char *dtype) {
  const std::string path = g_dir + "/" + name + ".bin";
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(1);
  }
  std::fclose(f);
  std::fprintf(g_manifest, "%s %s %d %d\n", name.c_str(), dtype, rows, cols);
}

void ok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

template <typename T> T *to_device(const std::vector<T> &h) {
  T *d = nullptr;
  ok(hipMalloc(&d, h.size() * sizeof(T)), "hipMalloc");
  ok(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "copy in");
  return d;
}
template <typename T> std::vector<T> to_host(const T *d, size_t n) {
  std::vector<T> h(n);
  ok(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "copy out");
  return h;
}

// One float semiring through the template: column-major A (M x K, lda = M),
// B (K x N, ldb = K), C and D (M x N, ldc = M), epilogue on.
template <template <typename> class AddOp, template <typename> class MulOp>
void float_case(const char *name, int M, int N, int K, float lo, float hi, float hole,
                bool integer) {
  auto draw = [&](size_t n, bool holes) {
    std::vector<float> v(n);
    for (auto &x : v) {
      x = integer ? (float)uniform_int((int)lo, (int)hi)
                  : lo + (hi - lo) * (float)((next_u32() >> 8) + 1) / 16777216.0f;
      if (holes && next_u32() % 10 == 0) x = hole;
    }
    return v;
  };
  const bool holes = std::isinf(hole);
  std::vector<float> A = draw((size_t)M * K, holes), B = draw((size_t)K * N, holes),
                     C = draw((size_t)M * N, false);
  float *dA = to_device(A), *dB = to_device(B), *dC = to_device(C), *dD = nullptr;
  ok(hipMalloc(&dD, C.size() * sizeof(float)), "hipMalloc");
  const int st = simd2::srgemm<AddOp, MulOp, float>(M, N, K, dA, M, dB, K, dC, M, dD, true);
  if (st != SIMD2_SUCCESS) {
    std::fprintf(stderr, "%s: status %d\n", name, st);
    std::exit(1);
  }
  ok(hipDeviceSynchronize(), "sync");
  const std::string n(name);
  // column-major M x K is stored as a row-major K x M array
  save(n + "_A", A, K, M, "f32");
  save(n + "_B", B, N, K, "f32");
  save(n + "_C", C, N, M, "f32");
  save(n + "_D", to_host(dD, C.size()), N, M, "f32");
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  (void)hipFree(dD);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
    return 2;
  }
  g_dir = argv[1];
  g_manifest = std::fopen((g_dir + "/manifest.txt").c_str(), "w");
  if (!g_manifest) return 1;
  using namespace simd2;
  const float inf = INFINITY;
  const int M = 150, N = 97, K = 70;  // off every tile grid
  float_case<minimum, plus>("min_plus", M, N, K, 1, 100, inf, true);
  float_case<plus, multiplies>("plus_mul", M, N, K, -1, 1, 0.0f, false);
  float_case<maximum, plus>("max_plus", M, N, K, 1, 100, -inf, true);
  float_case<maximum, minimum>("max_min", M, N, K, 1, 100, -inf, true);
  float_case<maximum, multiplies>("max_mul", M, N, K, 0, 1, 0.0f, false);
  float_case<minimum, multiplies>("min_mul", M, N, K, 0, 1, 0.0f, false);
  float_case<minimum, maximum>("min_max", M, N, K, 1, 100, inf, true);
  float_case<plus, minussquare>("minus_square", M, N, K, -20, 20, 0.0f, true);

  // or_and on 0/1 ints
  {
    std::vector<int> A((size_t)M * K), B((size_t)K * N), C((size_t)M * N);
    for (auto &x : A) x = next_u32() % 16 == 0;
    for (auto &x : B) x = next_u32() % 16 == 0;
    for (auto &x : C) x = next_u32() % 64 == 0;
    int *dA = to_device(A), *dB = to_device(B), *dC = to_device(C), *dD = nullptr;
    ok(hipMalloc(&dD, C.size() * sizeof(int)), "hipMalloc");
    const int st = srgemm<binary_or, binary_and, int>(M, N, K, dA, M, dB, K, dC, M, dD, true);
    if (st != SIMD2_SUCCESS) {
      std::fprintf(stderr, "or_and: status %d\n", st);
      return 1;
    }
    ok(hipDeviceSynchronize(), "sync");
    save("or_and_A", A, K, M, "i32");
    save("or_and_B", B, N, K, "i32");
    save("or_and_C", C, N, M, "i32");
    save("or_and_D", to_host(dD, C.size()), N, M, "i32");
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dC);
    (void)hipFree(dD);
  }

  // comp_update(int*): changed flag and prev <- cur (kernel/converge.cuh:71-85)
  {
    const int R = 33, Cc = 65;
    std::vector<int> prev((size_t)R * Cc), cur((size_t)R * Cc);
    for (size_t i = 0; i < prev.size(); ++i) prev[i] = cur[i] = uniform_int(0, 1000);
    int *dp = to_device(prev), *dc = to_device(cur), *check = nullptr, check_h = 0;
    ok(hipMalloc(&check, sizeof(int)), "hipMalloc");
    const bool same = comp_update(dp, dc, check, &check_h, R, Cc);  // nothing changed
    cur[700] += 1;
    ok(hipMemcpy(dc, cur.data(), cur.size() * sizeof(int), hipMemcpyHostToDevice), "copy");
    const bool changed = comp_update(dp, dc, check, &check_h, R, Cc);
    std::vector<int> flags = {same ? 1 : 0, changed ? 1 : 0};
    save("comp_update_flags", flags, 1, 2, "i32");
    save("comp_update_cur", cur, R, Cc, "i32");
    save("comp_update_prev_after", to_host(dp, prev.size()), R, Cc, "i32");
    (void)hipFree(dp);
    (void)hipFree(dc);
    (void)hipFree(check);
  }

  // __half overloads: f2h_device, h2f_device, tensor_srgemm (row-major, dims % 16 == 0)
  {
    const int m = 48, n = 32, k = 64;
    std::vector<float> a((size_t)m * k), b((size_t)k * n), c((size_t)m * n);
    for (auto &x : a) x = (float)uniform_int(0, 200) * 0.25f;
    for (auto &x : b) x = (float)uniform_int(0, 200) * 0.25f;
    for (auto &x : c) x = (float)uniform_int(0, 400);
    float *da = to_device(a), *db = to_device(b), *dc = to_device(c), *dd = nullptr,
          *back = nullptr;
    __half *ha = nullptr, *hb = nullptr;
    ok(hipMalloc(&ha, a.size() * sizeof(__half)), "hipMalloc");
    ok(hipMalloc(&hb, b.size() * sizeof(__half)), "hipMalloc");
    ok(hipMalloc(&dd, c.size() * sizeof(float)), "hipMalloc");
    ok(hipMalloc(&back, a.size() * sizeof(float)), "hipMalloc");
    f2h_device(ha, da, (int)a.size());
    f2h_device(hb, db, (int)b.size());
    h2f_device(back, ha, (int)a.size());
    const int st = tensor_srgemm(ha, hb, dc, dd, m, n, k, true, SIMD2_MIN_PLUS);
    if (st != 0) {
      std::fprintf(stderr, "tensor_srgemm: status %d\n", st);
      return 1;
    }
    ok(hipDeviceSynchronize(), "sync");
    save("tensor_A", a, m, k, "f32");
    save("tensor_B", b, k, n, "f32");
    save("tensor_C", c, m, n, "f32");
    save("tensor_D", to_host(dd, c.size()), m, n, "f32");
    save("tensor_A_roundtrip", to_host(back, a.size()), m, k, "f32");
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dc);
    (void)hipFree(dd);
    (void)hipFree(back);
    (void)hipFree(ha);
    (void)hipFree(hb);
  }
  std::fclose(g_manifest);
  std::printf("template driver: wrote %s/manifest.txt\n", g_dir.c_str());
  return 0;
}
