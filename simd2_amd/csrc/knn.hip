// k-nearest neighbours by pairwise squared L2 -- the PLD app's use of the
// minus_square semiring (apps/pld/cuASR/knncuda.cu:399-549) -- on gfx950.
//
// 1. knn_select_kernel: the reference's post-processing of a materialised
//    distance matrix (modified_insertion_sort_linear + compute_sqrt_linear,
//    knncuda.cu:170-237): per query the k smallest (distance, index) pairs in
//    ascending order, ties kept in index order, optionally square-rooted.
//    Layout as the reference: dist[q + r * Q], outputs [j * Q + q].  Two waves
//    scan two halves of the references for the same 64 queries (coalesced
//    rows), each keeping a sorted register list, and merge through LDS.
//    k <= 32.
//
// 2. The precision form (no counterpart in the reference):
//    * prep_kernel: FP32 points -> BF16 high and low parts (zero-padded to a
//      multiple of 32 dims) and FP32 squared norms;
//    * cand_kernel: ||q||^2 + ||r||^2 - 2 q.r with the dot product on BF16 MFMA
//      (v_mfma_f32_16x16x32_bf16, FP32 accumulate; hi.hi + hi.lo + lo.hi, three
//      MFMAs per k-step, for ~16 significant bits).  A workgroup owns 64
//      queries (each wave 16, their A fragments held in registers for the whole
//      scan) and one slice of the references, streamed through LDS in 64-row
//      tiles.  The 64x64 distance tile goes to LDS, where each thread scans 16
//      references for one query and keeps the 16 smallest in registers
//      (through a small pending buffer merged in lockstep).  The
//      Q x R matrix is never written; each query leaves 64 candidates per
//      reference slice;
//    * rerank_kernel: one wave per query picks the 32 best candidates by the
//      approximate distance, recomputes them exactly in FP32 ((q - r)^2 summed
//      by the wave), ranks by (distance, index), and for candidates whose FP32
//      distance lies within a relative 1e-5 of another one near the top k --
//      where cancellation or rounding could order them wrongly -- recomputes the
//      distance in FP64 and ranks those by it.  FP64 is used nowhere else.
#include <float.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "internal.hpp"

namespace simd2 {

namespace {

// ---------------------------------------------------------------------------
// 1. top-k of a materialised distance matrix (reference layout)
// ---------------------------------------------------------------------------
constexpr int SEL_KMAX = 32;
constexpr int SEL_Q = 64;      // queries per workgroup (one per lane)
constexpr int SEL_PARTS = 2;   // waves splitting the references

// Stable insertion into a sorted register list: strictly smaller distances
// move ahead, equal ones stay behind (earlier index first); while the list
// holds fewer than k entries every element enters.  `thr` caches entry k-1
// so no register is indexed dynamically.
template <int KMAX>
__device__ __forceinline__ void sorted_insert(float (&ld)[KMAX], int (&li)[KMAX], int &n,
                                              float &thr, int k, float d, int idx) {
  if (n >= k && !(d < thr)) return;
  float cd = d;
  int ci = idx;
  float last = thr;
  // The new element goes to the first position holding a strictly larger
  // distance (or the first free one); every entry after it then shifts one
  // place, unconditionally, so equal distances keep their order.
  bool shifting = false;
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    if (j < k) {
      const bool sw = shifting || ((j < n) ? (cd < ld[j]) : (j == n));
      const float td = ld[j];
      const int ti = li[j];
      if (sw) {
        ld[j] = cd;
        li[j] = ci;
        cd = td;
        ci = ti;
        shifting = true;
      }
      if (j == k - 1) last = ld[j];
    }
  }
  if (n < k) ++n;
  thr = last;
}

template <int KMAX>
__global__ __launch_bounds__(SEL_Q * SEL_PARTS) void knn_select_kernel(
    const float *__restrict__ dist, int Q, int R, int k, float *out_d, int *out_i, int do_sqrt) {
  __shared__ float sd[SEL_PARTS][KMAX][SEL_Q];
  __shared__ int si[SEL_PARTS][KMAX][SEL_Q];
  __shared__ int sn[SEL_PARTS][SEL_Q];
  const int lane = tid_x() % 64;
  const int part = tid_x() / 64;
  const int q = bid_x() * SEL_Q + lane;
  const int chunk = (R + SEL_PARTS - 1) / SEL_PARTS;
  const int r0 = part * chunk;
  const int r1 = min(R, r0 + chunk);
  float ld[KMAX];
  int li[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    ld[j] = FLT_MAX;
    li[j] = -1;
  }
  int n = 0;
  float thr = FLT_MAX;
  if (q < Q) {
    int r = r0;
    for (; r + 8 <= r1; r += 8) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = dist[(int64_t)(r + u) * Q + q];
#pragma unroll
      for (int u = 0; u < 8; ++u) sorted_insert<KMAX>(ld, li, n, thr, k, v[u], r + u);
    }
    for (; r < r1; ++r) sorted_insert<KMAX>(ld, li, n, thr, k, dist[(int64_t)r * Q + q], r);
  }
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    sd[part][j][lane] = ld[j];
    si[part][j][lane] = li[j];
  }
  sn[part][lane] = n;
  __syncthreads();
  if (part != 0 || q >= Q) return;
  // Merge: part 0 holds the lower indices, so on equal distances it wins,
  // which keeps index order.
  int p0 = 0, p1 = 0;
  const int n0 = sn[0][lane], n1 = sn[1][lane];
  for (int j = 0; j < k; ++j) {
    const bool has0 = p0 < n0, has1 = p1 < n1;
    if (!has0 && !has1) break;
    const float d0 = has0 ? sd[0][p0][lane] : 0.f;
    const float d1 = has1 ? sd[1][p1][lane] : 0.f;
    const bool take0 = has0 && (!has1 || !(d1 < d0));
    const float bd = take0 ? d0 : d1;
    const int bi = take0 ? si[0][p0][lane] : si[1][p1][lane];
    if (take0) ++p0; else ++p1;
    out_d[(int64_t)j * Q + q] = do_sqrt ? sqrtf(bd) : bd;
    out_i[(int64_t)j * Q + q] = bi;
  }
}

// ---------------------------------------------------------------------------
// 2. precision form
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// One wave per point: the FP32 coordinates split into a BF16 high part and a
// BF16 low part (x - hi), dims zero-padded to dpad, stored [point][hi | lo],
// and the FP32 squared norm (wave-summed).
__global__ __launch_bounds__(256) void prep_kernel(const float *__restrict__ P, int64_t ldp,
                                                   int n, int d, int dpad, __bf16 *Pb,
                                                   float *norm) {
  const int lane = tid_x() % 64;
  const int p = bid_x() * 4 + tid_x() / 64;
  if (p >= n) return;
  float s = 0.f;
  for (int j = lane; j < dpad; j += 64) {
    const float x = j < d ? P[(int64_t)p * ldp + j] : 0.f;
    s += x * x;
    const __bf16 hi = __float2bfloat16(x);
    Pb[(int64_t)p * 2 * dpad + j] = hi;
    Pb[(int64_t)p * 2 * dpad + dpad + j] = __float2bfloat16(x - __bfloat162float(hi));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) norm[p] = s;
}

constexpr int CQ = 64;    // queries per workgroup
constexpr int CR = 64;    // references per tile
constexpr int CLIST = 16; // candidates kept per thread
constexpr int CPEND = 4;  // pending candidates per thread before a merge

// q.r ~ qh.rh + qh.rl + ql.rh: three BF16 MFMAs per k-step carry ~16
// significant bits, which the norm expansion needs (centres ~N(0, 10^2) make
// ||q||^2 + ||r||^2 ~ 100x the neighbour distances); a single BF16 product
// leaves errors of ~10 on distances ~250 and loses true neighbours.
template <int DPAD>
__global__ __launch_bounds__(256) void cand_kernel(const __bf16 *__restrict__ Qb,
                                                   const __bf16 *__restrict__ Rb,
                                                   const float *__restrict__ qn,
                                                   const float *__restrict__ rn, int Q, int R,
                                                   int slice, float *cand_d, int *cand_i,
                                                   int ncand) {
  constexpr int KS = DPAD / 32;             // MFMA k-steps
  constexpr int RLD = 2 * DPAD + 8;         // LDS row: hi | lo, +16 bytes of padding
  constexpr int CHUNKS = CR * 2 * DPAD / 8; // 16-byte chunks per reference tile
  constexpr int PER = CHUNKS / 256;         // chunks per thread
  __shared__ __attribute__((aligned(16))) __bf16 Rt[CR][RLD];
  __shared__ float Dt[CQ][CR + 1];
  __shared__ float Rn[CR];

  const int tid = tid_x();
  const int lane = tid % 64;
  const int wave = tid / 64;
  const int q0 = bid_x() * CQ;
  const int rb0 = bid_y() * slice;
  const int rb1 = min(R, rb0 + slice);

  // A fragments of this wave's 16 queries: lane l holds query l%16, dims
  // 8*(l/16) .. +7 of every 32-wide k-step, high and low parts.
  bf16x8 ah[KS], al[KS];
  {
    const int q = q0 + wave * 16 + lane % 16;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      if (q < Q) {
        const __bf16 *base = Qb + (int64_t)q * 2 * DPAD + ks * 32 + 8 * (lane / 16);
        ah[ks] = *reinterpret_cast<const bf16x8 *>(base);
        al[ks] = *reinterpret_cast<const bf16x8 *>(base + DPAD);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          ah[ks][e] = (__bf16)0.0f;
          al[ks][e] = (__bf16)0.0f;
        }
      }
    }
  }
  // Norms of the 4 queries whose accumulator rows this lane holds.
  float qnv[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int q = q0 + wave * 16 + 4 * (lane / 16) + v;
    qnv[v] = q < Q ? qn[q] : 0.f;
  }

  // Candidate list of thread (query tid/4, tile quarter tid%4).
  float ld[CLIST];
  int li[CLIST];
#pragma unroll
  for (int c = 0; c < CLIST; ++c) {
    ld[c] = __builtin_inff();
    li[c] = -1;
  }
  float lmax = __builtin_inff();
  int lpos = 0;
  const int sq = tid / 4, sh = tid % 4;
  // Candidates pass `dd < lmax` into a pending buffer of CPEND (a few
  // instructions per reference); the buffers are merged into the lists, in
  // arrival order and against the then-current lmax, when some lane's buffer
  // is full -- all lanes at once -- so the list update does not run for the
  // whole wave every time any single lane has a candidate.  The lists come
  // out exactly as with one-at-a-time insertion: lmax only falls.
  float pd[CPEND];
  int pi[CPEND];
  int pc = 0;
  auto insert = [&](float dd, int idx) {
    if (dd < lmax) {
#pragma unroll
      for (int c = 0; c < CLIST; ++c)
        if (c == lpos) {
          ld[c] = dd;
          li[c] = idx;
        }
      lmax = ld[0];
      lpos = 0;
#pragma unroll
      for (int c = 1; c < CLIST; ++c)
        if (ld[c] > lmax) {
          lmax = ld[c];
          lpos = c;
        }
    }
  };
  auto flush = [&]() {
#pragma unroll
    for (int b = 0; b < CPEND; ++b)
      if (b < pc) insert(pd[b], pi[b]);
    pc = 0;
  };

  // Register prefetch of the next reference tile (global -> registers while
  // the current tile is multiplied and scanned, registers -> LDS after).
  uint4 pre[PER];
  float pren = 0.f;
  auto fetch = [&](int r0) {
#pragma unroll
    for (int h = 0; h < PER; ++h) {
      const int f = tid + 256 * h;
      const int rr = f / (2 * DPAD / 8);
      const int cc = (f % (2 * DPAD / 8)) * 8;
      const int r = r0 + rr;
      pre[h] = make_uint4(0u, 0u, 0u, 0u);
      if (r < rb1) pre[h] = *reinterpret_cast<const uint4 *>(Rb + (int64_t)r * 2 * DPAD + cc);
    }
    if (tid < CR) pren = (r0 + tid < rb1) ? rn[r0 + tid] : 0.f;
  };
  auto stage = [&]() {
#pragma unroll
    for (int h = 0; h < PER; ++h) {
      const int f = tid + 256 * h;
      const int rr = f / (2 * DPAD / 8);
      const int cc = (f % (2 * DPAD / 8)) * 8;
      *reinterpret_cast<uint4 *>(&Rt[rr][cc]) = pre[h];
    }
    if (tid < CR) Rn[tid] = pren;
  };

  if (rb0 < rb1) fetch(rb0);
  for (int r0 = rb0; r0 < rb1; r0 += CR) {
    __syncthreads();  // previous tile fully multiplied and scanned
    stage();
    __syncthreads();
    if (r0 + CR < rb1) fetch(r0 + CR);
    f32x4 acc[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) acc[rb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        const __bf16 *row = &Rt[rb * 16 + lane % 16][ks * 32 + 8 * (lane / 16)];
        const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(row);
        const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(row + DPAD);
        acc[rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[ks], bh, acc[rb], 0, 0, 0);
        acc[rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bl, acc[rb], 0, 0, 0);
        acc[rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bh, acc[rb], 0, 0, 0);
      }
    }
    // Approximate distances into the LDS tile; references past the slice end
    // get +inf so they never become candidates.
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      const int rr = rb * 16 + lane % 16;
      const bool valid = r0 + rr < rb1;
      const float rnv = Rn[rr];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float dd = valid ? qnv[v] + rnv - 2.0f * acc[rb][v] : __builtin_inff();
        Dt[wave * 16 + 4 * (lane / 16) + v][rr] = dd;
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int j = 0; j < CR / 4; ++j) {
      const float dd = Dt[sq][sh * (CR / 4) + j];
      if (dd < lmax) {
        const int idx = r0 + sh * (CR / 4) + j;
#pragma unroll
        for (int b = 0; b < CPEND; ++b)
          if (b == pc) {
            pd[b] = dd;
            pi[b] = idx;
          }
        ++pc;
      }
      if (__any(pc == CPEND)) flush();
    }
  }
  flush();
  const int q = q0 + sq;
  if (q < Q) {
    const int64_t base = (int64_t)q * ncand + (int64_t)bid_y() * (4 * CLIST) + sh * CLIST;
#pragma unroll
    for (int c = 0; c < CLIST; ++c) {
      cand_d[base + c] = ld[c];
      cand_i[base + c] = li[c];
    }
  }
}

constexpr int RR_M = 32;        // candidates re-ranked exactly per query
constexpr float RR_TOL = 1e-5f; // relative FP32 gap treated as a near tie

__device__ __forceinline__ bool key_less(float d, int i, float e, int j) {
  return d < e || (d == e && i < j);
}
__device__ __forceinline__ bool key_less64(double d, int i, double e, int j) {
  return d < e || (d == e && i < j);
}

__global__ __launch_bounds__(256) void rerank_kernel(const float *__restrict__ Qf, int64_t ldq,
                                                     const float *__restrict__ Rf, int64_t ldr,
                                                     int Q, int d, int k,
                                                     const float *__restrict__ cand_d,
                                                     const int *__restrict__ cand_i, int ncand,
                                                     float *out_d, int *out_i, int do_sqrt,
                                                     int *stats) {
  const int lane = tid_x() % 64;
  const int q = bid_x() * 4 + tid_x() / 64;
  if (q >= Q) return;
  const float *cd = cand_d + (int64_t)q * ncand;
  const int *ci = cand_i + (int64_t)q * ncand;

  // Step 1: the RR_M best candidates by (approximate distance, index).
  int sel = -1;  // lane m < RR_M receives the m-th choice
  float last_d = -__builtin_inff();
  int last_i = -1;
  for (int m = 0; m < RR_M; ++m) {
    float bd = __builtin_inff();
    int bi = 0x7fffffff;
    for (int c = lane; c < ncand; c += 64) {
      const float dd = cd[c];
      const int ii = ci[c];
      if (ii >= 0 && key_less(last_d, last_i, dd, ii) && key_less(dd, ii, bd, bi)) {
        bd = dd;
        bi = ii;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const float od = __shfl_xor(bd, off);
      const int oi = __shfl_xor(bi, off);
      if (key_less(od, oi, bd, bi)) {
        bd = od;
        bi = oi;
      }
    }
    if (bi == 0x7fffffff) break;  // fewer candidates than RR_M
    if (lane == m) sel = bi;
    last_d = bd;
    last_i = bi;
  }

  // Step 2: exact FP32 distances, one candidate at a time across the wave.
  float d32 = __builtin_inff();
  for (int m = 0; m < RR_M; ++m) {
    const int r = __shfl(sel, m);
    if (r < 0) break;
    float s = 0.f;
    for (int j = lane; j < d; j += 64) {
      const float t = Qf[(int64_t)q * ldq + j] - Rf[(int64_t)r * ldr + j];
      s += t * t;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == m) d32 = s;
  }
  const int my_i = sel >= 0 ? sel : 0x7fffffff;

  // Step 3: rank by (FP32 distance, index).
  int rank = 0;
  for (int m = 0; m < RR_M; ++m) {
    const float od = __shfl(d32, m);
    const int oi = __shfl(my_i, m);
    if (m != lane && key_less(od, oi, d32, my_i)) ++rank;
  }
  // Step 4: near ties that touch the top k get an FP64 distance.
  bool flag = false;
  for (int m = 0; m < RR_M; ++m) {
    const float od = __shfl(d32, m);
    const int orank = __shfl(rank, m);
    const int oi = __shfl(my_i, m);
    if (lane < RR_M && m != lane && sel >= 0 && oi != 0x7fffffff &&
        fabsf(od - d32) <= RR_TOL * fmaxf(d32, 1e-30f) && (rank <= k || orank <= k))
      flag = true;
  }
  double d64 = (double)d32;
  const uint64_t fl = __ballot(flag);
  if (fl) {
    for (int m = 0; m < RR_M; ++m) {
      if (!((fl >> m) & 1ull)) continue;
      const int r = __shfl(sel, m);
      double s = 0.0;
      for (int j = lane; j < d; j += 64) {
        const double t = (double)Qf[(int64_t)q * ldq + j] - (double)Rf[(int64_t)r * ldr + j];
        s += t * t;
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == m) d64 = s;
    }
    rank = 0;
    for (int m = 0; m < RR_M; ++m) {
      const double od = __shfl(d64, m);
      const int oi = __shfl(my_i, m);
      if (m != lane && key_less64(od, oi, d64, my_i)) ++rank;
    }
    if (stats && lane == 0) {
      atomicAdd(&stats[0], 1);
      atomicAdd(&stats[1], (int)__popcll(fl));
    }
  }
  if (lane < RR_M && sel >= 0 && rank < k) {
    const float outd = flag ? (float)d64 : d32;
    out_d[(int64_t)q * k + rank] = do_sqrt ? sqrtf(outd) : outd;
    out_i[(int64_t)q * k + rank] = sel;
  }
}

int dpad_of(int d) { return ((d + 31) / 32) * 32; }

}  // namespace

hipError_t launch_knn_select(const float *dist, int Q, int R, int k, float *out_d, int *out_i,
                             int do_sqrt, hipStream_t s) {
  const dim3 grid((Q + SEL_Q - 1) / SEL_Q), block(SEL_Q * SEL_PARTS);
  if (k <= 16)
    knn_select_kernel<16><<<grid, block, 0, s>>>(dist, Q, R, k, out_d, out_i, do_sqrt);
  else
    knn_select_kernel<SEL_KMAX><<<grid, block, 0, s>>>(dist, Q, R, k, out_d, out_i, do_sqrt);
  return hipGetLastError();
}

size_t knn_bf16_workspace(int Q, int R, int d, int *slices_out) {
  const int dpad = dpad_of(d);
  const int qblocks = (Q + CQ - 1) / CQ;
  // Enough reference slices to fill the GPU in whole rounds: the kernel
  // holds 2 workgroups per CU (its registers), 512 at a time.
  int slices = (1024 + qblocks - 1) / qblocks;
  slices = slices < 1 ? 1 : (slices > 8 ? 8 : slices);
  const int max_slices = (R + CR - 1) / CR;
  if (slices > max_slices) slices = max_slices < 1 ? 1 : max_slices;
  if (slices_out) *slices_out = slices;
  const size_t ncand = (size_t)slices * 4 * CLIST;
  size_t bytes = 0;
  bytes += ((size_t)Q * dpad * 4 + 255) / 256 * 256;  // hi | lo BF16
  bytes += ((size_t)R * dpad * 4 + 255) / 256 * 256;
  bytes += ((size_t)Q * 4 + 255) / 256 * 256;
  bytes += ((size_t)R * 4 + 255) / 256 * 256;
  bytes += ((size_t)Q * ncand * 4 + 255) / 256 * 256;
  bytes += ((size_t)Q * ncand * 4 + 255) / 256 * 256;
  return bytes;
}

hipError_t launch_knn_bf16(const float *query, int64_t ldq, const float *ref, int64_t ldr, int Q,
                           int R, int d, int k, float *out_d, int *out_i, int do_sqrt,
                           void *workspace, int *stats, hipStream_t s) {
  int slices = 1;
  knn_bf16_workspace(Q, R, d, &slices);
  const int dpad = dpad_of(d);
  const int ncand = slices * 4 * CLIST;
  char *w = static_cast<char *>(workspace);
  auto take = [&](size_t bytes) {
    char *p = w;
    w += (bytes + 255) / 256 * 256;
    return p;
  };
  __bf16 *Qb = reinterpret_cast<__bf16 *>(take((size_t)Q * dpad * 4));
  __bf16 *Rb = reinterpret_cast<__bf16 *>(take((size_t)R * dpad * 4));
  float *qn = reinterpret_cast<float *>(take((size_t)Q * 4));
  float *rn = reinterpret_cast<float *>(take((size_t)R * 4));
  float *cd = reinterpret_cast<float *>(take((size_t)Q * ncand * 4));
  int *ci = reinterpret_cast<int *>(take((size_t)Q * ncand * 4));

  prep_kernel<<<(Q + 3) / 4, 256, 0, s>>>(query, ldq, Q, d, dpad, Qb, qn);
  prep_kernel<<<(R + 3) / 4, 256, 0, s>>>(ref, ldr, R, d, dpad, Rb, rn);
  int slice = (R + slices - 1) / slices;
  slice = (slice + CR - 1) / CR * CR;
  const dim3 grid((Q + CQ - 1) / CQ, slices);
  switch (dpad) {
  case 32: cand_kernel<32><<<grid, 256, 0, s>>>(Qb, Rb, qn, rn, Q, R, slice, cd, ci, ncand); break;
  case 64: cand_kernel<64><<<grid, 256, 0, s>>>(Qb, Rb, qn, rn, Q, R, slice, cd, ci, ncand); break;
  case 96: cand_kernel<96><<<grid, 256, 0, s>>>(Qb, Rb, qn, rn, Q, R, slice, cd, ci, ncand); break;
  case 128: cand_kernel<128><<<grid, 256, 0, s>>>(Qb, Rb, qn, rn, Q, R, slice, cd, ci, ncand); break;
  case 160: cand_kernel<160><<<grid, 256, 0, s>>>(Qb, Rb, qn, rn, Q, R, slice, cd, ci, ncand); break;
  default: return hipErrorInvalidValue;
  }
  rerank_kernel<<<(Q + 3) / 4, 256, 0, s>>>(query, ldq, ref, ldr, Q, d, k, cd, ci, ncand, out_d,
                                           out_i, do_sqrt, stats);
  return hipGetLastError();
}

}  // namespace simd2
