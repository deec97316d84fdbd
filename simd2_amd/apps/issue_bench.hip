// VALU issue-rate micro-benchmark: wave-instructions per CU per cycle that
// gfx950 sustains for the instruction mixes of the semiring kernels' inner
// steps, at 1..8 waves per SIMD.  These rates are the ceilings the srgemm
// kernels are measured against (a kernel issuing its mix at the measured rate
// cannot go faster without fewer instructions per semiring operation).
//
// Mixes (each on 8 independent register chains, no memory traffic):
//   add        v_add_f32
//   min3       v_min3_f32
//   min_plus   2 x v_add_f32 + v_min3_f32   (min_plus / max_plus inner step)
//   max_mul    2 x v_mul_f32 + v_max3_f32   (max_mul / min_mul)
//   max_min    2 x v_min_f32 + v_max3_f32   (max_min / min_max)
//   sub_fmac   v_sub_f32 + v_fmac_f32       (direct minus_square)
//   min_i32    v_min_i32
//   min3_i32   v_min3_i32
//   max_min_i32  2 x v_min_i32 + v_max3_i32 (max_min / min_max on order-preserving
//              integer keys of the floats)
//   and_or     v_and_or_b32                 (packed or_and: 32 terms per lane)
//   mfma_i8    v_mfma_i32_16x16x64_i8       (the int8 minus_square form's
//              products; rate x 32768 ops x CUs x clock = ops/s)
//   lds_b128   ds_read_b128, 16 bytes a lane from 1 KB blocks of a 32 KB LDS
//              array (rate x 1024 = LDS bytes read per CU per cycle)
//
// One launch per (mix, W) of 4 W waves per CU (one or two workgroups per CU).
// Every wave reads the shader clock (s_memtime) around its loop and its
// hardware CU id (XCC and HW_ID registers); for each CU that held exactly
// 4 W waves, rate = 4 W x instructions per wave / (last end - first start
// over its waves), and the median over those CUs is reported.  Also printed:
// the rate implied by the event time at the device's reported clock.
//
//   simd2_issue_bench [--json]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

namespace {

constexpr int ITERS = 2048;
constexpr int CHAINS = 8;

enum Mix { ADD, MIN3, MIN_PLUS, MAX_MUL, MAX_MIN, SUB_FMAC, MIN_I32, MIN3_I32, MAX_MIN_I32, AND_OR,
           MFMA_I8, LDS_B128, NMIX };
const char *kMixName[NMIX] = {"add",      "min3",    "min_plus", "max_mul",     "max_min",
                              "sub_fmac", "min_i32", "min3_i32", "max_min_i32", "and_or",
                              "mfma_i8",  "lds_b128"};
const int kInstrPerChainIter[NMIX] = {1, 1, 3, 3, 3, 2, 1, 1, 3, 1, 1, 1};
typedef int i32x4 __attribute__((ext_vector_type(4)));

// One loop iteration = one asm block over the 8 chains (the compiler puts an
// s_nop after every asm statement, so one block per iteration keeps that to
// one per 8-24 VALU instructions; s_nop does not occupy the VALU).  Chain c
// computes the two products of a k-pair into temporaries and folds them into
// its accumulator m_c, as the srgemm kernels' inner step does.
#define S2_CH(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)
#define S2_ADD(c) "v_add_f32 %[m" #c "], %[m" #c "], %[a]\n\t"
#define S2_MIN3(c) "v_min3_f32 %[m" #c "], %[m" #c "], %[a], %[b]\n\t"
#define S2_MIN_PLUS(c)                                                                      \
  "v_add_f32 %[t0], %[y" #c "], %[a]\n\tv_add_f32 %[t1], %[y" #c "], %[b]\n\t"            \
  "v_min3_f32 %[m" #c "], %[m" #c "], %[t0], %[t1]\n\t"
#define S2_MAX_MUL(c)                                                                       \
  "v_mul_f32 %[t0], %[y" #c "], %[a]\n\tv_mul_f32 %[t1], %[y" #c "], %[b]\n\t"            \
  "v_max3_f32 %[m" #c "], %[m" #c "], %[t0], %[t1]\n\t"
#define S2_MAX_MIN(c)                                                                       \
  "v_min_f32 %[t0], %[y" #c "], %[a]\n\tv_min_f32 %[t1], %[y" #c "], %[b]\n\t"            \
  "v_max3_f32 %[m" #c "], %[m" #c "], %[t0], %[t1]\n\t"
#define S2_MIN_I32(c) "v_min_i32 %[m" #c "], %[m" #c "], %[a]\n\t"
#define S2_MIN3_I32(c) "v_min3_i32 %[m" #c "], %[m" #c "], %[a], %[b]\n\t"
#define S2_MAX_MIN_I32(c)                                                                   \
  "v_min_i32 %[t0], %[y" #c "], %[a]\n\tv_min_i32 %[t1], %[y" #c "], %[b]\n\t"            \
  "v_max3_i32 %[m" #c "], %[m" #c "], %[t0], %[t1]\n\t"
#define S2_AND_OR(c) "v_and_or_b32 %[m" #c "], %[a], %[y" #c "], %[m" #c "]\n\t"
#define S2_SUB_FMAC(c)                                                                      \
  "v_sub_f32 %[t0], %[y" #c "], %[a]\n\tv_fmac_f32 %[m" #c "], %[t0], %[t0]\n\t"
#define S2_OUT_M                                                                            \
  [m0] "+v"(m[0]), [m1] "+v"(m[1]), [m2] "+v"(m[2]), [m3] "+v"(m[3]), [m4] "+v"(m[4]),      \
      [m5] "+v"(m[5]), [m6] "+v"(m[6]), [m7] "+v"(m[7]), [t0] "=&v"(t0), [t1] "=&v"(t1)
#define S2_IN_Y                                                                             \
  [y0] "v"(y[0]), [y1] "v"(y[1]), [y2] "v"(y[2]), [y3] "v"(y[3]), [y4] "v"(y[4]),           \
      [y5] "v"(y[5]), [y6] "v"(y[6]), [y7] "v"(y[7]), [a] "v"(a), [b] "v"(b)

template <int MIX>
__global__ void issue_kernel(float seed, float *sink, unsigned long long *t, int waves) {
  static_assert(CHAINS == 8, "the asm blocks name 8 chains");
  float y[CHAINS], m[CHAINS];
  float t0 = 0.f, t1 = 0.f;
#pragma unroll
  for (int c = 0; c < CHAINS; ++c) {
    y[c] = seed * c;
    m[c] = seed - c;
  }
  const float a = seed * 0.5f, b = seed * 0.25f;
  i32x4 acc[CHAINS];  // MFMA_I8 accumulators / LDS_B128 read sums
#pragma unroll
  for (int c = 0; c < CHAINS; ++c) acc[c] = i32x4{c, 0, 0, 0};
  const int tid = (int)__builtin_amdgcn_workitem_id_x();
  __shared__ __attribute__((aligned(16))) int lds[MIX == LDS_B128 ? 8192 : 4];
  if constexpr (MIX == LDS_B128) {
    for (int i = tid; i < 8192; i += (int)__builtin_amdgcn_workgroup_size_x()) lds[i] = i * 7;
    __syncthreads();
  }
  __builtin_amdgcn_s_barrier();
  const unsigned long long t0c = __builtin_amdgcn_s_memtime();
  for (int it = 0; it < ITERS; ++it) {
    if constexpr (MIX == MFMA_I8) {
      const i32x4 av = {(int)seed, it, 3, 5}, bv = {7, (int)seed, it, 1};
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) acc[c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bv, acc[c], 0, 0, 0);
    } else if constexpr (MIX == LDS_B128) {
      // block (wave + c + it) % 32 of 1 KB, lane l at 16 l: in bounds, no conflicts
#pragma unroll
      for (int c = 0; c < CHAINS; ++c)
        acc[c] ^= *reinterpret_cast<const i32x4 *>(
            &lds[((tid / 64 + c + it) % 32) * 256 + (tid % 64) * 4]);
    } else if constexpr (MIX == ADD)
      asm volatile(S2_CH(S2_ADD) : S2_OUT_M : S2_IN_Y);
    else if constexpr (MIX == MIN3)
      asm volatile(S2_CH(S2_MIN3) : S2_OUT_M : S2_IN_Y);
    else if constexpr (MIX == MIN_PLUS)
      asm volatile(S2_CH(S2_MIN_PLUS) : S2_OUT_M : S2_IN_Y);
    else if constexpr (MIX == MAX_MUL)
      asm volatile(S2_CH(S2_MAX_MUL) : S2_OUT_M : S2_IN_Y);
    else if constexpr (MIX == MAX_MIN)
      asm volatile(S2_CH(S2_MAX_MIN) : S2_OUT_M : S2_IN_Y);
    else if constexpr (MIX == SUB_FMAC)
      asm volatile(S2_CH(S2_SUB_FMAC) : S2_OUT_M : S2_IN_Y);
    else if constexpr (MIX == MIN_I32)
      asm volatile(S2_CH(S2_MIN_I32) : S2_OUT_M : S2_IN_Y);
    else if constexpr (MIX == MIN3_I32)
      asm volatile(S2_CH(S2_MIN3_I32) : S2_OUT_M : S2_IN_Y);
    else if constexpr (MIX == AND_OR)
      asm volatile(S2_CH(S2_AND_OR) : S2_OUT_M : S2_IN_Y);
    else
      asm volatile(S2_CH(S2_MAX_MIN_I32) : S2_OUT_M : S2_IN_Y);
  }
  const unsigned long long t1c = __builtin_amdgcn_s_memtime();
  float s = t0 + t1;
#pragma unroll
  for (int c = 0; c < CHAINS; ++c) s += m[c] + (float)(acc[c][0] ^ acc[c][3]);
  if (s == 1234.5f) sink[0] = s;  // keeps the chains live; never true for the seeds used
  const int lane = (int)__builtin_amdgcn_workitem_id_x() % 64;
  const int wave = (int)__builtin_amdgcn_workitem_id_x() / 64;
  // HW_ID bits 15:8 = SE, SH and CU of this wave; XCC_ID bits 3:0 = its XCD.
  const unsigned hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);
  const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20);
  if (lane == 0) {
    const size_t w = (size_t)__builtin_amdgcn_workgroup_id_x() * waves + wave;
    t[3 * w] = t0c;
    t[3 * w + 1] = t1c;
    t[3 * w + 2] = (xcc << 8) | ((hw >> 8) & 0xffu);
  }
}

bool ok(hipError_t e, const char *what) {
  if (e != hipSuccess) std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
  return e == hipSuccess;
}

template <int MIX>
bool run(int cus, int wps, int clock_khz, bool json) {
  const int wgs = (wps + 3) / 4;  // workgroups per CU (at most 16 waves each)
  const int waves = 4 * wps / wgs;
  const int threads = waves * 64;
  const size_t nwaves = (size_t)cus * 4 * wps;
  unsigned long long *t = nullptr;
  float *sink = nullptr;
  if (!ok(hipMalloc(&t, nwaves * 3 * sizeof(unsigned long long)), "hipMalloc") ||
      !ok(hipMalloc(&sink, 4), "hipMalloc"))
    return false;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  // warm-up, then the timed launch
  issue_kernel<MIX><<<cus * wgs, threads>>>(1.0f, sink, t, waves);
  (void)hipEventRecord(e0);
  issue_kernel<MIX><<<cus * wgs, threads>>>(1.0f, sink, t, waves);
  (void)hipEventRecord(e1);
  if (!ok(hipEventSynchronize(e1), "kernel")) return false;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  std::vector<unsigned long long> h(nwaves * 3);
  if (!ok(hipMemcpy(h.data(), t, h.size() * sizeof(h[0]), hipMemcpyDeviceToHost), "copy"))
    return false;
  const double instr_per_wave = (double)ITERS * CHAINS * kInstrPerChainIter[MIX];
  struct Span {
    int waves = 0;
    unsigned long long lo = ~0ull, hi = 0;
  };
  std::map<unsigned long long, Span> per_cu;
  for (size_t w = 0; w < nwaves; ++w) {
    Span &sp = per_cu[h[3 * w + 2]];
    ++sp.waves;
    sp.lo = std::min(sp.lo, h[3 * w]);
    sp.hi = std::max(sp.hi, h[3 * w + 1]);
  }
  std::vector<double> rate;
  for (const auto &kv : per_cu)
    if (kv.second.waves == 4 * wps && kv.second.hi > kv.second.lo)
      rate.push_back(4.0 * wps * instr_per_wave / (double)(kv.second.hi - kv.second.lo));
  if (rate.empty()) {
    std::fprintf(stderr, "%s W=%d: no CU held exactly %d waves\n", kMixName[MIX], wps, 4 * wps);
    return false;
  }
  std::sort(rate.begin(), rate.end());
  const double med = rate[rate.size() / 2];
  const double evt = (double)nwaves * instr_per_wave / cus / (ms * 1e-3 * clock_khz * 1e3);
  if (json)
    std::printf("{\"mix\": \"%s\", \"waves_per_simd\": %d, \"rate\": %.4f, "
                "\"rate_event_clock\": %.4f, \"ms\": %.4f, \"cus_measured\": %zu}\n",
                kMixName[MIX], wps, med, evt, ms, rate.size());
  else
    std::printf("%-9s waves/SIMD %d: %.3f wave-instr/CU/cycle (shader clock, %zu CUs), %.3f at "
                "the event time and %d MHz, %.3f ms\n",
                kMixName[MIX], wps, med, rate.size(), evt, clock_khz / 1000, ms);
  (void)hipFree(t);
  (void)hipFree(sink);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  const bool json = argc > 1 && std::strcmp(argv[1], "--json") == 0;
  hipDeviceProp_t p;
  if (!ok(hipGetDeviceProperties(&p, 0), "hipGetDeviceProperties")) return 1;
  const int cus = p.multiProcessorCount;
  if (!json) std::printf("%s: %d CUs, %d MHz\n", p.gcnArchName, cus, p.clockRate / 1000);
  for (int wps : {1, 2, 4, 8}) {
    if (!run<ADD>(cus, wps, p.clockRate, json) || !run<MIN3>(cus, wps, p.clockRate, json) ||
        !run<MIN_PLUS>(cus, wps, p.clockRate, json) ||
        !run<MAX_MUL>(cus, wps, p.clockRate, json) ||
        !run<MAX_MIN>(cus, wps, p.clockRate, json) ||
        !run<SUB_FMAC>(cus, wps, p.clockRate, json) ||
        !run<MIN_I32>(cus, wps, p.clockRate, json) ||
        !run<MIN3_I32>(cus, wps, p.clockRate, json) ||
        !run<MAX_MIN_I32>(cus, wps, p.clockRate, json) ||
        !run<AND_OR>(cus, wps, p.clockRate, json) || !run<MFMA_I8>(cus, wps, p.clockRate, json) ||
        !run<LDS_B128>(cus, wps, p.clockRate, json))
      return 1;
  }
  return 0;
}
