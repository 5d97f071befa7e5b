// radix.hip -- stable LSD radix sort of (key, value) pairs for gfx950.
//
// Used for the particle re-sort by cell (32- or 64-bit keys: segment rank
// above the cell key) and for the reference-order scatter of the exact FP64
// mode (64-bit keys: node, original particle index, stencil slot).
//
// One pass per 8-bit digit, three kernels per pass:
//   k_rs_hist    per tile of 4096 keys, the count of every digit value,
//                stored digit-major (counts[d * ntiles + tile]);
//   k_rs_scan    one block per digit: exclusive scan of its row over the
//                tiles, and the digit's total;
//   k_rs_scatter per tile: the tile's base offset of every digit (scan of
//                the totals plus its row entry), then 16 rounds of 256 keys
//                in index order; inside a round every wave ranks its lanes
//                among equal digits with ballots (lanes of lower index
//                first), the four waves' counts are prefix-summed per digit
//                in LDS, and every key goes to base + wave prefix + rank.
// Equal digits keep their input order (tiles, rounds, waves and lanes are
// all visited in index order), so the sort is stable and the permutation is
// the one a stable CPU sort of the same keys gives.
//
// The sort reads the input pair arrays and alternates between the two
// buffers; the return value says which holds the result.  Scratch:
// radix_scratch_words(n) unsigned words, no allocation inside (so a sort can
// be enqueued in captured streams).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace mpmk {

namespace {
constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
constexpr int RS_RADIX = 256;
constexpr int RS_WAVES = RS_BLOCK / 64;

template <typename K>
__device__ __forceinline__ unsigned digit_of(K key, int shift) {
  return (unsigned)(key >> shift) & (RS_RADIX - 1);
}

template <typename K>
__global__ void __launch_bounds__(RS_BLOCK) k_rs_hist(const K* __restrict__ keys, long n, int shift,
                                                     unsigned* __restrict__ counts, int ntiles) {
  __shared__ unsigned h[RS_RADIX];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long base = (long)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int it = 0; it < RS_ITEMS; it++) {
    const long idx = base + (long)it * RS_BLOCK + threadIdx.x;
    if (idx < n) atomicAdd(&h[digit_of(keys[idx], shift)], 1u);
  }
  __syncthreads();
  counts[(long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of one digit's row of tile counts; the row total to tot[d]
__global__ void __launch_bounds__(RS_BLOCK) k_rs_scan(unsigned* __restrict__ counts, int ntiles,
                                                     unsigned* __restrict__ tot) {
  __shared__ unsigned s[RS_BLOCK];
  unsigned* row = counts + (long)blockIdx.x * ntiles;
  unsigned carry = 0;
  constexpr int PER = 4;
  for (int t0 = 0; t0 < ntiles; t0 += RS_BLOCK * PER) {
    unsigned v[PER], sum = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int idx = t0 + threadIdx.x * PER + k;
      v[k] = idx < ntiles ? row[idx] : 0u;
      sum += v[k];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < RS_BLOCK; off <<= 1) {   // inclusive Hillis-Steele scan
      const unsigned add = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    unsigned excl = s[threadIdx.x] - sum + carry;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int idx = t0 + threadIdx.x * PER + k;
      if (idx < ntiles) row[idx] = excl;
      excl += v[k];
    }
    carry += s[RS_BLOCK - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

template <typename K>
__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(const K* __restrict__ kin, const int* __restrict__ vin,
                                                        K* __restrict__ kout, int* __restrict__ vout, long n,
                                                        int shift, const unsigned* __restrict__ counts,
                                                        const unsigned* __restrict__ tot, int ntiles) {
  __shared__ unsigned base[RS_RADIX];
  __shared__ unsigned wc[RS_WAVES][RS_RADIX];
  __shared__ unsigned sc[RS_RADIX];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // digit bases: exclusive scan of the digit totals, plus this tile's row entry
  {
    const unsigned my = tot[t];
    sc[t] = my;
    __syncthreads();
    for (int off = 1; off < RS_RADIX; off <<= 1) {
      const unsigned add = t >= off ? sc[t - off] : 0u;
      __syncthreads();
      sc[t] += add;
      __syncthreads();
    }
    base[t] = sc[t] - my + counts[(long)t * ntiles + blockIdx.x];
  }
  const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const long tile0 = (long)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ITEMS; r++) {
    const long idx = tile0 + (long)r * RS_BLOCK + t;
    const bool valid = idx < n;
    K key = 0;
    int val = 0;
    unsigned d = 0;
    if (valid) {
      key = kin[idx];
      val = vin[idx];
      d = digit_of(key, shift);
    }
    // lanes of this wave with the same digit
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long bb = __ballot(valid && bit);
      m &= bit ? bb : ~bb;
    }
    const unsigned rank = (unsigned)__popcll(m & lt);
#pragma unroll
    for (int w = 0; w < RS_WAVES; w++) wc[w][t] = 0u;
    __syncthreads();
    if (valid && rank == 0) wc[wave][d] = (unsigned)__popcll(m);
    __syncthreads();
    unsigned total;
    {   // per digit: exclusive prefix over the waves
      unsigned acc = 0;
#pragma unroll
      for (int w = 0; w < RS_WAVES; w++) {
        const unsigned c = wc[w][t];
        wc[w][t] = acc;
        acc += c;
      }
      total = acc;
    }
    __syncthreads();
    if (valid) {
      const unsigned pos = base[d] + wc[wave][d] + rank;
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    base[t] += total;
    __syncthreads();
  }
}
}  // namespace

size_t radix_scratch_words(long n) {
  const long ntiles = (n + RS_TILE - 1) / RS_TILE;
  return (size_t)RS_RADIX * (size_t)(ntiles > 0 ? ntiles : 1) + RS_RADIX;
}

template <typename K>
int radix_sort_pairs(K* k0, K* k1, int* v0, int* v1, long n, int begin_bit, int end_bit, unsigned* scratch,
                     hipStream_t s) {
  if (n <= 0) return 0;
  const int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
  unsigned* counts = scratch;
  unsigned* tot = scratch + (size_t)RS_RADIX * ntiles;
  K* kin = k0;
  K* kout = k1;
  int* vin = v0;
  int* vout = v1;
  int parity = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist<K>, dim3(ntiles), dim3(RS_BLOCK), 0, s, kin, n, shift, counts, ntiles);
    hipLaunchKernelGGL(k_rs_scan, dim3(RS_RADIX), dim3(RS_BLOCK), 0, s, counts, ntiles, tot);
    hipLaunchKernelGGL(k_rs_scatter<K>, dim3(ntiles), dim3(RS_BLOCK), 0, s, kin, vin, kout, vout, n, shift, counts,
                       tot, ntiles);
    K* tk = kin; kin = kout; kout = tk;
    int* tv = vin; vin = vout; vout = tv;
    parity ^= 1;
  }
  return parity;
}

template int radix_sort_pairs<unsigned>(unsigned*, unsigned*, int*, int*, long, int, int, unsigned*, hipStream_t);
template int radix_sort_pairs<unsigned long long>(unsigned long long*, unsigned long long*, int*, int*, long, int,
                                                  int, unsigned*, hipStream_t);
}  // namespace mpmk
