// simd2_app -- the reference's seven repeated-squaring graph drivers
// (apps/{apsp,aplp,mcp,maxrp,minrp,mst,gtc}/cuASR/*_cuASR.cu) rebuilt on
// libsimd2's C ABI, with the same inputs, fills and output line.
//
//   simd2_app <app> -f [file]           edge list on stdin: "v e" then e lines "u v w"
//                                        (mst: a PBBS weighted edge array file,
//                                        apps/data/from_pbbs.h:31-57)
//   simd2_app <app> v density [weight]  generated graph: rgg_1d (gtc: rgg_1d_directed,
//                                        weight 10), seed 7 (apps/data/graph_gen.h:33-81)
//   options: --reps N (default 10, the reference's NUM_ITR)  --loop reference|fused
//            --check threshold|exact  --dump FILE (final matrix, raw row-major)
//            --warmup N (default 0: untimed solves before the timed ones, so that the
//            first launch of each kernel -- code loading -- stays out of the mean)
//
// Output: the mean milliseconds per solve (first line, as the reference prints it),
// then "iters <n>" and "checksum <x>" (utils/check_sum.h:3-18).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/simd2/converge.h"
#include "../../include/simd2/srgemm.h"

namespace {

struct AppSpec {
  const char *name;
  int op;
  bool epilogue;
  int check;  // comp_update flavour the reference app uses
};

const AppSpec kApps[] = {
    {"apsp", SIMD2_MIN_PLUS, true, SIMD2_CHECK_THRESHOLD},
    {"aplp", SIMD2_MAX_PLUS, true, SIMD2_CHECK_APLP},
    {"mcp", SIMD2_MAX_MIN, true, SIMD2_CHECK_THRESHOLD},
    {"maxrp", SIMD2_MAX_MUL, true, SIMD2_CHECK_THRESHOLD},
    {"minrp", SIMD2_MIN_MUL, true, SIMD2_CHECK_THRESHOLD},
    {"mst", SIMD2_MIN_MAX, false, SIMD2_CHECK_THRESHOLD},
    {"gtc", SIMD2_OR_AND, true, SIMD2_CHECK_EXACT},
};

// Random graph generators with the reference's draw order (glibc rand(),
// weight draw before edge draw), so a given (v, density, weight, seed)
// reproduces the reference's matrix.
void rgg_undirected(std::vector<float> &g, int v, float density, float weight, int seed) {
  srand(seed);
  for (int i = 0; i < v; i++)
    for (int j = 0; j < v; j++) {
      if (i == j) {
        g[(size_t)i * v + j] = 0.0f;
        continue;
      }
      if (j > i) {
        const float w = (float)rand() / ((float)RAND_MAX / weight);
        const float has = (float)rand() / (float)(RAND_MAX / 1);
        g[(size_t)i * v + j] = has < density ? w : FLT_MAX;
        g[(size_t)j * v + i] = g[(size_t)i * v + j];
      }
    }
}

void rgg_directed(std::vector<float> &g, int v, float density, float weight, int seed) {
  srand(seed);
  for (int i = 0; i < v; i++)
    for (int j = 0; j < v; j++) {
      if (i == j) {
        g[(size_t)i * v + j] = 0.0f;
        continue;
      }
      const float w = (float)rand() / ((float)RAND_MAX / weight);
      const float has = (float)rand() / (float)(RAND_MAX / 1);
      g[(size_t)i * v + j] = has < density ? w : FLT_MAX;
    }
}

// Edge list from stdin with the app's fill (the reference mains).
int read_stdin(const std::string &app, std::vector<float> &g) {
  int v = 0, e = 0;
  if (!(std::cin >> v >> e) || v <= 0) return -1;
  float fill = FLT_MAX;
  if (app == "aplp" || app == "mcp") fill = -std::numeric_limits<float>::max();
  if (app == "maxrp") fill = 0.0f;
  g.assign((size_t)v * v, fill);
  for (int i = 0; i < e; ++i) {
    int a, b;
    float w;
    if (!(std::cin >> a >> b >> w)) return -1;
    if (a < 0 || b < 0 || a >= v || b >= v) return -1;
    g[(size_t)a * v + b] = w;
  }
  if (app == "apsp")
    for (int i = 0; i < v; ++i) g[(size_t)i * v + i] = 0.0f;
  if (app == "mcp")
    for (int i = 0; i < v; ++i) g[(size_t)i * v + i] = -FLT_MAX;
  return v;
}

// PBBS weighted edge array (header line, then "a b w"), symmetric, absent
// off-diagonal edges FLT_MAX, diagonal 0.
int read_pbbs(const std::string &fname, std::vector<float> &g) {
  std::ifstream in(fname);
  if (!in) return -1;
  std::string line;
  std::getline(in, line);
  std::vector<std::tuple<int, int, float>> edges;
  int vmax = 0;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    int a, b;
    float w;
    if (!(iss >> a >> b >> w)) break;
    if (a < 0 || b < 0) return -1;
    edges.emplace_back(a, b, w);
    vmax = std::max(vmax, std::max(a, b));
  }
  const int v = vmax + 1;
  g.assign((size_t)v * v, 0.0f);
  for (auto &t : edges) {
    g[(size_t)std::get<0>(t) * v + std::get<1>(t)] = std::get<2>(t);
    g[(size_t)std::get<1>(t) * v + std::get<0>(t)] = std::get<2>(t);
  }
  for (int i = 0; i < v; ++i)
    for (int j = 0; j < v; ++j)
      if (i != j && g[(size_t)i * v + j] == 0.0f) g[(size_t)i * v + j] = FLT_MAX;
  return v;
}

double check_sum(const std::vector<float> &x) {
  double s = 0;
  for (float f : x)
    if (std::fabs(f) < FLT_MAX - 100) s += std::fabs(f);
  return std::sqrt(s);
}

#define HIP_OK(call)                                                             \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
                   __FILE__, __LINE__);                                          \
      std::exit(2);                                                              \
    }                                                                            \
  } while (0)

// One solve as the reference times it: buffers allocated and filled before
// the clock starts, the loop timed, the result copied back after.
template <typename T>
double solve(const AppSpec &a, int v, std::vector<T> &m, int loop, int check, int *iters) {
  using clk = std::chrono::high_resolution_clock;
  T *x = nullptr, *w = nullptr;
  int *flag = nullptr;
  const size_t bytes = (size_t)v * v * sizeof(T);
  HIP_OK(hipMalloc(&x, bytes));
  HIP_OK(hipMalloc(&w, bytes));
  HIP_OK(hipMalloc(&flag, sizeof(int)));
  HIP_OK(hipMemcpy(x, m.data(), bytes, hipMemcpyHostToDevice));
  HIP_OK(hipDeviceSynchronize());
  const auto t0 = clk::now();
  int st;
  if constexpr (std::is_same<T, int>::value)
    st = simd2_converge_i32(v, x, v, w, v, flag, 500, loop, nullptr, iters);
  else
    st = simd2_converge_f32(a.op, v, x, v, w, v, flag, a.epilogue ? 1 : 0, check, 500, loop,
                            nullptr, iters);
  HIP_OK(hipDeviceSynchronize());
  const auto t1 = clk::now();
  if (st != SIMD2_SUCCESS && st != SIMD2_ERROR_NOT_CONVERGED) {
    std::fprintf(stderr, "converge failed: %s\n", simd2_status_string(st));
    std::exit(2);
  }
  HIP_OK(hipMemcpy(m.data(), x, bytes, hipMemcpyDeviceToHost));
  HIP_OK(hipFree(x));
  HIP_OK(hipFree(w));
  HIP_OK(hipFree(flag));
  return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

void usage() {
  std::fprintf(stderr,
               "usage: simd2_app <apsp|aplp|mcp|maxrp|minrp|mst|gtc> (-f [file] | v density "
               "[weight]) [--reps N] [--loop reference|fused] [--check threshold|exact] "
               "[--dump FILE]\n");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    usage();
    return 1;
  }
  const std::string app = argv[1];
  const AppSpec *spec = nullptr;
  for (const auto &a : kApps)
    if (app == a.name) spec = &a;
  if (!spec) {
    usage();
    return 1;
  }
  int reps = 10, warmup = 0, loop = SIMD2_CONVERGE_REFERENCE, check = spec->check;
  std::string dump, file;
  std::vector<std::string> pos;
  for (int i = 2; i < argc; ++i) {
    const std::string s = argv[i];
    if (s == "--reps" && i + 1 < argc) reps = std::max(1, std::atoi(argv[++i]));
    else if (s == "--warmup" && i + 1 < argc) warmup = std::max(0, std::atoi(argv[++i]));
    else if (s == "--loop" && i + 1 < argc) loop = std::string(argv[++i]) == "fused" ? SIMD2_CONVERGE_FUSED : SIMD2_CONVERGE_REFERENCE;
    else if (s == "--check" && i + 1 < argc) check = std::string(argv[++i]) == "exact" ? SIMD2_CHECK_EXACT : SIMD2_CHECK_THRESHOLD;
    else if (s == "--dump" && i + 1 < argc) dump = argv[++i];
    else pos.push_back(s);
  }
  std::vector<float> g;
  int v = -1;
  if (!pos.empty() && pos[0] == "-f") {
    if (app == "mst") {
      if (pos.size() < 2) {
        usage();
        return 1;
      }
      v = read_pbbs(pos[1], g);
    } else {
      v = read_stdin(app, g);
    }
  } else if (pos.size() >= 2) {
    v = std::atoi(pos[0].c_str());
    const float density = (float)std::atof(pos[1].c_str());
    if (v <= 0 || density < 0 || density > 1) {
      std::fprintf(stderr, "bad generator arguments\n");
      return 1;
    }
    g.assign((size_t)v * v, 0.0f);
    if (app == "gtc") rgg_directed(g, v, density, 10.0f, 7);
    else rgg_undirected(g, v, density, pos.size() >= 3 ? (float)std::atof(pos[2].c_str()) : 10.0f, 7);
  }
  if (v <= 0) {
    std::fprintf(stderr, "could not read the graph\n");
    return 1;
  }

  double total = 0;
  int iters = 0;
  double cs = 0;
  if (app == "gtc") {
    // apps/gtc/cuASR/gtc_cuASR.cu:96-111: reachability matrix, diagonal 1.
    std::vector<int> t0((size_t)v * v), t;
    for (int i = 0; i < v; ++i)
      for (int j = 0; j < v; ++j)
        t0[(size_t)i * v + j] = (i == j) || g[(size_t)i * v + j] < FLT_MAX - 100;
    for (int r = 0; r < warmup + reps; ++r) {
      t = t0;
      const double ms = solve(*spec, v, t, loop, SIMD2_CHECK_EXACT, &iters);
      if (r >= warmup) total += ms;
    }
    for (int x : t) cs += x;
    if (!dump.empty()) {
      std::ofstream o(dump, std::ios::binary);
      o.write(reinterpret_cast<const char *>(t.data()), (std::streamsize)(t.size() * 4));
    }
  } else {
    std::vector<float> m;
    for (int r = 0; r < warmup + reps; ++r) {
      m = g;
      const double ms = solve(*spec, v, m, loop, check, &iters);
      if (r >= warmup) total += ms;
    }
    cs = check_sum(m);
    if (!dump.empty()) {
      std::ofstream o(dump, std::ios::binary);
      o.write(reinterpret_cast<const char *>(m.data()), (std::streamsize)(m.size() * 4));
    }
  }
  std::printf("%f\n", total / reps);
  std::printf("iters %d\n", iters);
  std::printf("checksum %f\n", cs);
  return 0;
}
