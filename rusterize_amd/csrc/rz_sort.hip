// rz_sort.hip -- exclusive scan and stable LSD radix sort of (key, value)
// pairs, written for the rasteriser's binning: every launch is a plain
// kernel (no memset / host-side state), so a sequence of them can be
// captured in a HIP graph and replayed.
#pragma once

#include "rz_kernels.h"

namespace rzk {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;                              // per thread
constexpr int kScanBlock = kScanThreads * kScanItems;      // 2048 elements per block

// Block-wide exclusive scan of one value per thread (256 threads); returns
// the thread's exclusive prefix and writes the block total to *total.
template <typename T>
__device__ inline T block_exclusive_scan(T v, T* s_tmp /* [kScanThreads / 64] */, T* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;  // inclusive scan inside the wave
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const T y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_tmp[wave] = x;
  __syncthreads();
  T wave_pre = 0, sum = 0;
  for (int w = 0; w < kScanThreads / 64; ++w) {
    const T t = s_tmp[w];
    if (w < wave) wave_pre += t;
    sum += t;
  }
  __syncthreads();
  *total = sum;
  return wave_pre + x - v;
}

// Pass 1: the sum of every block of kScanBlock inputs.
template <typename In, typename Out>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const In* in, uint64_t n, Out* bsum) {
  __shared__ Out s_tmp[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanBlock + (uint64_t)threadIdx.x * kScanItems;
  Out v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) v += (Out)in[base + k];
  Out total;
  (void)block_exclusive_scan<Out>(v, s_tmp, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Pass 2: exclusive scan of the block sums in place, one block.
template <typename Out>
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(Out* bsum, uint64_t nb) {
  __shared__ Out s_tmp[kScanThreads / 64];
  const uint64_t per = (nb + kScanThreads - 1) / kScanThreads;
  const uint64_t b0 = (uint64_t)threadIdx.x * per, b1 = min(nb, b0 + per);
  Out v = 0;
  for (uint64_t b = b0; b < b1; ++b) v += bsum[b];
  Out total;
  Out run = block_exclusive_scan<Out>(v, s_tmp, &total);
  for (uint64_t b = b0; b < b1; ++b) {
    const Out t = bsum[b];
    bsum[b] = run;
    run += t;
  }
}

// Pass 3: scan of every block plus its block offset.
template <typename In, typename Out>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const In* in, uint64_t n, const Out* bsum,
                                                             Out* out) {
  __shared__ Out s_tmp[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanBlock + (uint64_t)threadIdx.x * kScanItems;
  Out vals[kScanItems];
  Out v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    vals[k] = base + k < n ? (Out)in[base + k] : (Out)0;
    v += vals[k];
  }
  Out total;
  Out run = block_exclusive_scan<Out>(v, s_tmp, &total) + (bsum ? bsum[blockIdx.x] : (Out)0);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += vals[k];
  }
}

inline uint64_t scan_blocks(uint64_t n) { return std::max<uint64_t>(1, (n + kScanBlock - 1) / kScanBlock); }

// Exclusive prefix sums out[i] = in[0] + ... + in[i-1] over n elements.
// tmp: scan_blocks(n) elements of Out.  Up to one block of inputs is one
// launch (no block sums).
template <typename In, typename Out>
void exclusive_scan(const In* in, Out* out, uint64_t n, Out* tmp, hipStream_t s) {
  if (n == 0) return;
  const uint64_t nb = scan_blocks(n);
  if (nb == 1) {
    hipLaunchKernelGGL((k_scan_apply<In, Out>), dim3(1), dim3(kScanThreads), 0, s, in, n, (const Out*)nullptr, out);
    return;
  }
  hipLaunchKernelGGL((k_scan_reduce<In, Out>), dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, tmp);
  hipLaunchKernelGGL((k_scan_blocks<Out>), dim3(1), dim3(kScanThreads), 0, s, tmp, nb);
  hipLaunchKernelGGL((k_scan_apply<In, Out>), dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n,
                     (const Out*)tmp, out);
}

// ------------------------------------------------------------ radix sort --

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;                          // keys per thread
constexpr int kSortBlock = kSortThreads * kSortRounds;   // 4096 keys per block
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

inline uint64_t sort_blocks(uint64_t n) { return std::max<uint64_t>(1, (n + kSortBlock - 1) / kSortBlock); }

// Digit histogram of every block, digit-major: hist[d * nb + b].
__global__ __launch_bounds__(kSortThreads) void k_rs_hist(const uint32_t* keys, uint64_t n, int shift,
                                                           uint32_t* hist, uint64_t nb) {
  __shared__ uint32_t s_h[kRadix];
  s_h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kSortBlock;
  for (int r = 0; r < kSortRounds; ++r) {
    const uint64_t i = base + (uint64_t)r * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & (kRadix - 1)], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = s_h[threadIdx.x];
}

// Stable scatter of one digit pass: keys of a block are ranked in index order
// (earlier rounds first, then lower waves, then lower lanes of the same
// digit, found with eight ballots).
__global__ __launch_bounds__(kSortThreads) void k_rs_scatter(const uint32_t* keys, const uint32_t* vals,
                                                              uint64_t n, int shift, const uint32_t* off,
                                                              uint64_t nb, uint32_t* okeys, uint32_t* ovals) {
  __shared__ uint32_t s_run[kRadix];       // keys of each digit placed by earlier rounds
  __shared__ uint32_t s_base[kRadix];      // the block's first output slot of each digit
  __shared__ uint32_t s_wc[4][kRadix];     // this round's count of each digit per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  s_run[threadIdx.x] = 0;
  s_base[threadIdx.x] = off[(uint64_t)threadIdx.x * nb + blockIdx.x];
  for (int w = 0; w < 4; ++w) s_wc[w][threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kSortBlock;
  for (int r = 0; r < kSortRounds; ++r) {
    const uint64_t i = base + (uint64_t)r * kSortThreads + threadIdx.x;
    const bool valid = i < n;
    uint32_t key = 0, val = 0, d = 0;
    if (valid) {
      key = keys[i];
      val = vals[i];
      d = (key >> shift) & (kRadix - 1);
    }
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t ones = __ballot(valid && bit);
      peers &= bit ? ones : ~ones;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    if (valid && rank == 0) s_wc[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pre = 0;
      for (int w = 0; w < wave; ++w) pre += s_wc[w][d];
      const uint32_t pos = s_base[d] + s_run[d] + pre + rank;
      okeys[pos] = key;
      ovals[pos] = val;
    }
    __syncthreads();
    const uint32_t t = threadIdx.x;  // one thread per digit
    s_run[t] += s_wc[0][t] + s_wc[1][t] + s_wc[2][t] + s_wc[3][t];
    s_wc[0][t] = s_wc[1][t] = s_wc[2][t] = s_wc[3][t] = 0;
    __syncthreads();
  }
}

// Stable sort of n pairs by key bits [0, end_bit), ping-ponging between
// (ka, va) and (kb, vb); returns true when the result is in (kb, vb).
// hist: kRadix * sort_blocks(n) words; scan_tmp: scan_blocks of that.
inline bool radix_sort_pairs(uint32_t* ka, uint32_t* va, uint32_t* kb, uint32_t* vb, uint64_t n,
                             int end_bit, uint32_t* hist, uint32_t* off, uint32_t* scan_tmp,
                             hipStream_t s) {
  const uint64_t nb = sort_blocks(n);
  bool in_b = false;
  for (int shift = 0; shift < end_bit; shift += kRadixBits) {
    const uint32_t* sk = in_b ? kb : ka;
    const uint32_t* sv = in_b ? vb : va;
    uint32_t* dk = in_b ? ka : kb;
    uint32_t* dv = in_b ? va : vb;
    hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)nb), dim3(kSortThreads), 0, s, sk, n, shift, hist, nb);
    exclusive_scan<uint32_t, uint32_t>(hist, off, (uint64_t)kRadix * nb, scan_tmp, s);
    hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nb), dim3(kSortThreads), 0, s, sk, sv, n, shift,
                       (const uint32_t*)off, nb, dk, dv);
    in_b = !in_b;
  }
  return in_b;
}

}  // namespace rzk
