// radix_sort.hip -- stable LSD radix sort of (uint32 key, uint32 value)
// pairs for gfx950: the sortParticles stage (src/functions.cu:336-341).
//
// 8-bit digits, one pass per digit of the key range actually used (keys are
// cell indices, so 24-bit keys need 3 passes).  Each pass is
//   upsweep   : per-tile digit counts (wave64 ballot match, no atomics);
//   scan_rows : per-digit exclusive scan over tiles (+ digit totals);
//   downsweep : stable tile ranking (ballot match + popcount prefix), a
//               scatter into LDS in sorted order, then ordered writes so that
//               consecutive lanes write consecutive addresses of a digit run.
// Tiles are 256 threads x 16 keys.  Inside a tile, wave w owns the 1024
// consecutive keys [w*1024, (w+1)*1024) and visits them 64 at a time in
// input order, so ranks follow input order: the sort is stable.
#include "dem_internal.h"

namespace dem {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kItems = 16;
constexpr uint32_t kTile = kThreads * kItems;
constexpr int kBits = 8;
constexpr int kRadix = 1 << kBits;
static_assert(kRadix == kThreads, "one thread per digit in the per-digit phases");

// Lanes of this wave (within `valid`) whose digit equals this lane's digit.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
  uint64_t peers = valid;
#pragma unroll
  for (int b = 0; b < kBits; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, kWave);
    if (lane >= (uint32_t)o) x += y;
  }
  return x;
}

// Exclusive scan over the 256 threads of the block.  Contains barriers; every
// thread of the block must call it.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* wsum, uint32_t& total) {
  const uint32_t lane = __lane_id(), w = threadIdx.x / kWave;
  const uint32_t inc = wave_inclusive_scan(x);
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kWaves; ++k) {
    const uint32_t t = wsum[k];
    off += (k < (int)w) ? t : 0u;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return off + inc - x;
}

__global__ __launch_bounds__(kThreads) void k_upsweep(const uint32_t* __restrict__ keys, uint32_t n,
                                                      uint32_t shift, uint32_t* __restrict__ block_counts,
                                                      uint32_t num_tiles) {
  __shared__ uint32_t cnt[kWaves][kRadix];
  for (int k = threadIdx.x; k < kWaves * kRadix; k += kThreads) (&cnt[0][0])[k] = 0;
  __syncthreads();
  const uint32_t lane = __lane_id(), w = threadIdx.x / kWave, tile = blockIdx.x;
  const uint32_t base = tile * kTile + w * (kWave * kItems);
  uint32_t key[kItems];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint32_t idx = base + it * kWave + lane;
    key[it] = idx < n ? keys[idx] : 0u;
  }
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint32_t idx = base + it * kWave + lane;
    const bool ok = idx < n;
    const uint64_t valid = __ballot(ok);
    const uint32_t d = (key[it] >> shift) & (kRadix - 1);
    const uint64_t peers = match_digit(d, valid);
    if (ok && lane == (uint32_t)(__ffsll((unsigned long long)peers) - 1)) cnt[w][d] += (uint32_t)__popcll(peers);
  }
  __syncthreads();
  const uint32_t d = threadIdx.x;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kWaves; ++k) s += cnt[k][d];
  block_counts[(size_t)d * num_tiles + tile] = s;
}

__global__ __launch_bounds__(kThreads) void k_scan_rows(uint32_t* __restrict__ block_counts,
                                                        uint32_t num_tiles,
                                                        uint32_t* __restrict__ row_totals) {
  __shared__ uint32_t wsum[kWaves];
  uint32_t* row = block_counts + (size_t)blockIdx.x * num_tiles;
  const uint32_t per = (num_tiles + kThreads - 1) / kThreads;
  const uint32_t beg = min(threadIdx.x * per, num_tiles), end = min(beg + per, num_tiles);
  uint32_t s = 0;
  for (uint32_t t = beg; t < end; ++t) s += row[t];
  uint32_t total;
  uint32_t run = block_exclusive_scan(s, wsum, total);
  for (uint32_t t = beg; t < end; ++t) {
    const uint32_t v = row[t];
    row[t] = run;
    run += v;
  }
  if (threadIdx.x == 0) row_totals[blockIdx.x] = total;
}

__global__ __launch_bounds__(kThreads) void k_downsweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift,
    const uint32_t* __restrict__ block_offsets, const uint32_t* __restrict__ row_totals,
    uint32_t num_tiles) {
  __shared__ uint32_t s_keys[kTile];
  __shared__ uint32_t s_vals[kTile];
  __shared__ uint32_t cnt[kWaves][kRadix];
  __shared__ uint32_t g_start[kRadix];
  __shared__ uint32_t wsum[kWaves];

  const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid / kWave, tile = blockIdx.x;
  for (int k = tid; k < kWaves * kRadix; k += kThreads) (&cnt[0][0])[k] = 0;
  uint32_t unused_total;
  // Start of each digit's run in the output: scan of the digit totals.
  const uint32_t digit_base = block_exclusive_scan(row_totals[tid], wsum, unused_total);

  const uint32_t base = tile * kTile + w * (kWave * kItems);
  uint32_t key[kItems], val[kItems], rank[kItems];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint32_t idx = base + it * kWave + lane;
    if (idx < n) {
      key[it] = keys_in[idx];
      val[it] = vals_in[idx];
    } else {
      key[it] = 0u;
      val[it] = 0u;
    }
  }

  const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint32_t idx = base + it * kWave + lane;
    const bool ok = idx < n;
    const uint64_t valid = __ballot(ok);
    const uint32_t d = (key[it] >> shift) & (kRadix - 1);
    const uint64_t peers = match_digit(d, valid);
    rank[it] = 0;
    if (ok) {
      const uint32_t before = cnt[w][d];
      rank[it] = before + (uint32_t)__popcll(peers & lt);
      if (lane == (uint32_t)(__ffsll((unsigned long long)peers) - 1))
        cnt[w][d] = before + (uint32_t)__popcll(peers);
    }
  }
  __syncthreads();

  {
    const uint32_t d = tid;
    const uint32_t c0 = cnt[0][d], c1 = cnt[1][d], c2 = cnt[2][d], c3 = cnt[3][d];
    uint32_t tile_total;
    const uint32_t local_start = block_exclusive_scan(c0 + c1 + c2 + c3, wsum, tile_total);
    cnt[0][d] = local_start;
    cnt[1][d] = local_start + c0;
    cnt[2][d] = local_start + c0 + c1;
    cnt[3][d] = local_start + c0 + c1 + c2;
    g_start[d] = digit_base + block_offsets[(size_t)d * num_tiles + tile] - local_start;
  }
  __syncthreads();

#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint32_t idx = base + it * kWave + lane;
    if (idx < n) {
      const uint32_t d = (key[it] >> shift) & (kRadix - 1);
      const uint32_t p = cnt[w][d] + rank[it];
      if (p < kTile) {  // always true for a consistent ranking; guards the store
        s_keys[p] = key[it];
        s_vals[p] = val[it];
      }
    }
  }
  __syncthreads();

  const uint32_t tile_n = min(kTile, n - tile * kTile);
  for (uint32_t s = tid; s < tile_n; s += kThreads) {
    const uint32_t k = s_keys[s];
    const uint32_t g = g_start[(k >> shift) & (kRadix - 1)] + s;
    if (g < n) {  // always true for a consistent ranking; guards the store
      keys_out[g] = k;
      vals_out[g] = s_vals[s];
    }
  }
}

}  // namespace

int sort_passes(uint32_t key_bits) { return (int)((key_bits + kBits - 1) / kBits); }

uint32_t sort_tiles(uint32_t n) { return (n + kTile - 1) / kTile; }

size_t sort_block_count_elems(uint32_t n) { return (size_t)kRadix * sort_tiles(n); }

hipError_t launch_sort(hipStream_t s, uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits,
                       const SortScratch& scratch, bool* result_in_alt) {
  *result_in_alt = false;
  const int passes = sort_passes(key_bits);
  if (n <= 1 || passes == 0) return hipSuccess;
  if (n > scratch.capacity || !scratch.keys_alt || !scratch.vals_alt || !scratch.block_counts ||
      !scratch.row_totals)
    return hipErrorInvalidValue;
  const uint32_t tiles = sort_tiles(n);
  uint32_t *ki = keys, *vi = vals, *ko = scratch.keys_alt, *vo = scratch.vals_alt;
  for (int p = 0; p < passes; ++p) {
    const uint32_t shift = (uint32_t)(p * kBits);
    hipLaunchKernelGGL(k_upsweep, dim3(tiles), dim3(kThreads), 0, s, ki, n, shift, scratch.block_counts, tiles);
    hipLaunchKernelGGL(k_scan_rows, dim3(kRadix), dim3(kThreads), 0, s, scratch.block_counts, tiles,
                       scratch.row_totals);
    hipLaunchKernelGGL(k_downsweep, dim3(tiles), dim3(kThreads), 0, s, ki, vi, ko, vo, n, shift,
                       scratch.block_counts, scratch.row_totals, tiles);
    uint32_t* t = ki; ki = ko; ko = t;
    t = vi; vi = vo; vo = t;
  }
  *result_in_alt = (passes & 1) != 0;
  return hipGetLastError();
}

}  // namespace dem
