// radix_sort.hip -- stable LSD radix sort of (uint32 key, uint32 value)
// pairs for gfx950: the sortParticles stage (src/functions.cu:336-341).
//
// 8-bit digits, one pass per digit of the key range actually used (keys are
// cell indices, so 24-bit keys need 3 passes).  Each pass is
//   upsweep   : per-tile digit counts (LDS atomics on per-wave counters);
//   scan_rows : per-digit exclusive scan over tiles (+ digit totals);
//   downsweep : stable tile ranking, a scatter into LDS in sorted order, then
//               ordered writes so that consecutive lanes write consecutive
//               addresses of a digit run.
// Tiles are 4096 keys.  Inside a tile, each wave owns a run of consecutive
// keys and visits them 64 at a time in input order, and the waves' counts are
// combined in wave order, so ranks follow input order: the sort is stable.
//
// A lane's rank among the earlier lanes of its wave with the same digit comes
// from the wave's peer mask: every lane ORs its lane bit into the wave's LDS
// word for its digit, reads the word back (the lanes sharing its digit) and
// clears it.  The LDS instructions of one wave execute in issue order, so the
// read sees every lane's OR and the next group's ORs see the cleared words.
// A group whose lanes all share one digit (the high digits of nearly sorted
// keys) skips the LDS round trip.  The counting and ranking cost a few
// instructions a key, so the passes stream at close to the memory rate.
#include "dem_internal.h"

namespace dem {
namespace {

constexpr int kThreads = 256;  // the row scan
constexpr int kWaves = kThreads / kWave;
constexpr uint32_t kTile = kSortTile;
constexpr int kUThreads = 512;  // the upsweep
constexpr int kUWaves = kUThreads / kWave;
constexpr int kItems = (int)kTile / kUThreads;
static_assert(kItems * kUThreads == (int)kTile && kItems % 4 == 0, "upsweep threads cover a tile, 4 keys a load");
constexpr int kBits = 8;
constexpr int kRadix = 1 << kBits;
static_assert(kRadix == kThreads, "one thread per digit in the per-digit phases");
static_assert(kWaves * kRadix <= (int)kTile, "the peer masks fit in the tile's LDS staging");

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

// True when every lane with ok has digit d0 (the first lane's digit; the
// lanes with ok are a prefix of the wave).
__device__ __forceinline__ bool wave_one_digit(uint32_t d, bool ok, uint32_t& d0) {
  d0 = __builtin_amdgcn_readfirstlane(d);
  return __ballot(ok && d != d0) == 0;
}

// Lanes of this wave (within valid, the lanes with ok) whose digit equals
// this lane's; 0 for a lane without ok.  tab: this wave's kRadix words, zero
// on entry and on return.
__device__ __forceinline__ uint64_t peer_mask(uint32_t d, bool ok, uint64_t valid, uint64_t* tab) {
  uint32_t d0;
  if (wave_one_digit(d, ok, d0)) return ok ? valid : 0ull;
  uint64_t peers = 0;
  if (ok) __hip_atomic_fetch_or(&tab[d], 1ull << __lane_id(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  if (ok) peers = __hip_atomic_load(&tab[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  // one lane a digit clears its word (fewer bank conflicts than every lane)
  if (ok && __lane_id() == (uint32_t)(__ffsll((unsigned long long)peers) - 1))
    __hip_atomic_store(&tab[d], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  return peers;
}

// Loads of 4 consecutive words at once (16-byte loads stream faster than
// one word a lane) where the array is 16-byte aligned and the 4 words lie
// below n; the caller checks the alignment.
template <bool kVec>
__device__ __forceinline__ uint4 load4(const uint32_t* __restrict__ a, uint32_t i, uint32_t n) {
  if (kVec && i + 4u <= n) return *reinterpret_cast<const uint4*>(a + i);
  uint4 v;
  v.x = i < n ? a[i] : 0u;
  v.y = i + 1u < n ? a[i + 1u] : 0u;
  v.z = i + 2u < n ? a[i + 2u] : 0u;
  v.w = i + 3u < n ? a[i + 3u] : 0u;
  return v;
}

__device__ __forceinline__ uint32_t word(const uint4& v, int j) { return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w; }

// Per-tile digit counts.  The order of the keys does not matter here, so
// each lane counts 4 consecutive keys at a time.
template <bool kVec>
__global__ __launch_bounds__(kUThreads) void k_upsweep(const uint32_t* __restrict__ keys, uint32_t n,
                                                       uint32_t shift, uint32_t* __restrict__ block_counts,
                                                       uint32_t num_tiles) {
  __shared__ uint32_t cnt[kUWaves][kRadix];
  for (int k = threadIdx.x; k < kUWaves * kRadix; k += kUThreads) (&cnt[0][0])[k] = 0;
  __syncthreads();
  const uint32_t lane = __lane_id(), w = threadIdx.x / kWave, tile = blockIdx.x;
  const uint32_t base = tile * kTile + w * (kWave * kItems);
  constexpr int kQuads = kItems / 4;
  uint4 kv[kQuads];
#pragma unroll
  for (int q = 0; q < kQuads; ++q) kv[q] = load4<kVec>(keys, base + 4u * (q * kWave + lane), n);
#pragma unroll
  for (int q = 0; q < kQuads; ++q) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t idx = base + 4u * (q * kWave + lane) + j;
      const bool ok = idx < n;
      const uint32_t d = (word(kv[q], j) >> shift) & (kRadix - 1);
      uint32_t d0;
      if (wave_one_digit(d, ok, d0)) {
        const uint64_t valid = __ballot(ok);
        if (lane == 0 && valid) atomicAdd(&cnt[w][d0], (uint32_t)__popcll(valid));
      } else if (ok) {
        atomicAdd(&cnt[w][d], 1u);
      }
    }
  }
  __syncthreads();
  const uint32_t d = threadIdx.x;
  if (d < (uint32_t)kRadix) {
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kUWaves; ++k) s += cnt[k][d];
    block_counts[(size_t)d * num_tiles + tile] = s;
  }
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

// The downsweep: 512-thread blocks, each taking tiles blockIdx.x,
// blockIdx.x + gridDim.x, ...  Wave w of a block owns the tile's keys
// [w*1024, (w+1)*1024): it loads them 4 consecutive words a lane, passes
// them through its part of the LDS staging, and ranks them 64 at a time in
// input order.  The loads of a block's next tile are issued before the
// current tile's write-out.
constexpr int kDThreads = 512;
constexpr int kDWaves = kDThreads / kWave;
constexpr int kDItems = (int)kTile / kDThreads;
constexpr int kDQuads = kDItems / 4;
static_assert(kDItems * kDThreads == (int)kTile && kDItems % 4 == 0, "downsweep tiles are the upsweep's tiles");
static_assert(kDWaves * kRadix * 2 <= (int)kTile, "the peer masks fit in the keys' half of the LDS staging");

template <bool kVec>
__device__ __forceinline__ void load_tile(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
                                          uint32_t base, uint32_t n, uint4* kq, uint4* vq) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int q = 0; q < kDQuads; ++q) {
    const uint32_t i = base + 4u * (q * kWave + lane);
    kq[q] = load4<kVec>(keys_in, i, n);
    vq[q] = load4<kVec>(vals_in, i, n);
  }
}

// (two blocks a CU: at most 128 VGPRs)
template <bool kVec>
__global__ __launch_bounds__(kDThreads, 4) void k_downsweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift,
    const uint32_t* __restrict__ block_offsets, const uint32_t* __restrict__ row_totals,
    uint32_t num_tiles) {
  // the tile's keys and values: in load order (keys in the first half,
  // values in the second), then, after the ranking (for which its first
  // kDWaves * kRadix * 8 bytes hold the waves' peer masks), in sorted order
  // as (value << 32 | key)
  __shared__ uint64_t stage[kTile];
  __shared__ uint32_t cnt[kDWaves][kRadix];
  __shared__ uint32_t g_start[kRadix];
  __shared__ uint32_t wsum[kDWaves];
  __shared__ uint32_t s_base[kRadix];
  uint32_t* s_keys = reinterpret_cast<uint32_t*>(stage);
  uint32_t* s_vals = s_keys + kTile;

  const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid / kWave;
  uint64_t* tab = stage + w * kRadix;
  uint32_t* wk = s_keys + w * (kWave * kDItems);
  uint32_t* wv = s_vals + w * (kWave * kDItems);
  // Start of each digit's run in the output: exclusive scan of the digit
  // totals (the first 256 threads, one digit each: four waves).
  if (tid < (uint32_t)kRadix) {
    const uint32_t x = row_totals[tid];
    const uint32_t inc = wave_inclusive_scan(x);
    if (lane == kWave - 1) wsum[w] = inc;
    s_base[tid] = inc - x;
  }
  __syncthreads();
  if (tid < (uint32_t)kRadix) {
    uint32_t off = 0;
    for (uint32_t k = 0; k < w; ++k) off += wsum[k];
    s_base[tid] += off;
  }

  const uint64_t lt = (1ull << lane) - 1ull;
  uint4 kq[kDQuads], vq[kDQuads];
  uint32_t key[kDItems], val[kDItems], rank[kDItems];
  uint32_t tile = blockIdx.x;
  if (tile < num_tiles) load_tile<kVec>(keys_in, vals_in, tile * kTile + w * (kWave * kDItems), n, kq, vq);
  for (; tile < num_tiles; tile += gridDim.x) {
    const uint32_t base = tile * kTile + w * (kWave * kDItems);
    __syncthreads();  // the previous tile's write-out has read the staging
    // lane l holds keys 4(q*64 + l) .. +3 of the wave's run; the ranking
    // visits key it*64 + l in its item it
#pragma unroll
    for (int q = 0; q < kDQuads; ++q) {
      *reinterpret_cast<uint4*>(wk + 4u * (q * kWave + lane)) = kq[q];
      *reinterpret_cast<uint4*>(wv + 4u * (q * kWave + lane)) = vq[q];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kDItems; ++it) key[it] = wk[it * kWave + lane];
    __syncthreads();  // every wave has read its keys before the peer masks are cleared
    for (int k = tid; k < kDWaves * kRadix; k += kDThreads) {
      (&cnt[0][0])[k] = 0;
      stage[k] = 0ull;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kDItems; ++it) {
      const uint32_t idx = base + it * kWave + lane;
      const bool ok = idx < n;
      const uint64_t valid = __ballot(ok);
      const uint32_t d = (key[it] >> shift) & (kRadix - 1);
      const uint64_t peers = peer_mask(d, ok, valid, tab);
      rank[it] = 0;
      if (ok) {
        const uint32_t before = cnt[w][d];
        rank[it] = before + (uint32_t)__popcll(peers & lt);
        if (lane == (uint32_t)(__ffsll((unsigned long long)peers) - 1))
          cnt[w][d] = before + (uint32_t)__popcll(peers);
      }
    }
    __syncthreads();
    // per digit: the tile's waves in order (the first 256 threads), then the
    // digit's start in the tile (exclusive scan over digits)
    if (tid < (uint32_t)kRadix) {
      const uint32_t d = tid;
      uint32_t run = 0;
#pragma unroll
      for (int k = 0; k < kDWaves; ++k) {
        const uint32_t c = cnt[k][d];
        cnt[k][d] = run;
        run += c;
      }
      const uint32_t inc = wave_inclusive_scan(run);
      if (lane == kWave - 1) wsum[w] = inc;
      g_start[d] = inc - run;  // the digit's start in the tile, before the waves' offsets
    }
    __syncthreads();
    if (tid < (uint32_t)kRadix) {
      const uint32_t d = tid;
      uint32_t off = 0;
      for (uint32_t k = 0; k < w; ++k) off += wsum[k];
      const uint32_t local_start = g_start[d] + off;
#pragma unroll
      for (int k = 0; k < kDWaves; ++k) cnt[k][d] += local_start;
      g_start[d] = s_base[d] + block_offsets[(size_t)d * num_tiles + tile] - local_start;
    }
    // the values wait in the staging (the peer masks use only the keys'
    // half) until the scatter, which overwrites it
#pragma unroll
    for (int it = 0; it < kDItems; ++it) val[it] = wv[it * kWave + lane];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kDItems; ++it) {
      const uint32_t idx = base + it * kWave + lane;
      if (idx < n) {
        const uint32_t d = (key[it] >> shift) & (kRadix - 1);
        const uint32_t p = cnt[w][d] + rank[it];
        if (p < kTile)  // always true for a consistent ranking; guards the store
          stage[p] = ((uint64_t)val[it] << 32) | key[it];
      }
    }
    __syncthreads();
    // the next tile's loads go out before this tile's writes
    const uint32_t next = tile + gridDim.x;
    if (next < num_tiles) load_tile<kVec>(keys_in, vals_in, next * kTile + w * (kWave * kDItems), n, kq, vq);
    const uint32_t tile_n = min(kTile, n - tile * kTile);
    for (uint32_t s = tid; s < tile_n; s += kDThreads) {
      const uint64_t kv = stage[s];
      const uint32_t k = (uint32_t)kv;
      const uint32_t g = g_start[(k >> shift) & (kRadix - 1)] + s;
      if (g < n) {  // always true for a consistent ranking; guards the store
        keys_out[g] = k;
        vals_out[g] = (uint32_t)(kv >> 32);
      }
    }
  }
}

__global__ void k_copy_pairs(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint32_t* __restrict__ da,
                             uint32_t* __restrict__ db, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    da[i] = a[i];
    db[i] = b[i];
  }
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace

int sort_passes(uint32_t key_bits) { return (int)((key_bits + kBits - 1) / kBits); }

uint32_t sort_tiles(uint32_t n) { return (n + kTile - 1) / kTile; }

size_t sort_block_count_elems(uint32_t n) { return (size_t)kRadix * sort_tiles(n); }

hipError_t launch_sort_to(hipStream_t s, uint32_t* keys_in, uint32_t* vals_in, uint32_t* keys_out,
                          uint32_t* vals_out, uint32_t n, uint32_t key_bits, const SortScratch& scratch) {
  if (n == 0) return hipSuccess;
  const int passes = sort_passes(key_bits);
  const dim3 copy_grid((n + 255) / 256);
  if (n == 1 || passes == 0) {
    if (keys_out != keys_in || vals_out != vals_in)
      hipLaunchKernelGGL(k_copy_pairs, copy_grid, dim3(256), 0, s, keys_in, vals_in, keys_out, vals_out, n);
    return hipGetLastError();
  }
  if (n > scratch.capacity || !scratch.keys_alt || !scratch.vals_alt || !scratch.block_counts ||
      !scratch.row_totals || (keys_out == keys_in) != (vals_out == vals_in))
    return hipErrorInvalidValue;
  const uint32_t tiles = sort_tiles(n);
  // two downsweep blocks a CU (256 CUs), each taking tiles in turn
  const uint32_t down_blocks = std::min(tiles, 512u);
  // Intermediate passes alternate between the scratch arrays and the second
  // scratch arrays, or the input arrays where there are none.
  uint32_t* bk[2] = {scratch.keys_alt, scratch.keys_alt2 ? scratch.keys_alt2 : keys_in};
  uint32_t* bv[2] = {scratch.vals_alt, scratch.vals_alt2 ? scratch.vals_alt2 : vals_in};
  uint32_t *ki = keys_in, *vi = vals_in;
  bool copy_back = false;
  for (int p = 0; p < passes; ++p) {
    uint32_t *ko = bk[p & 1], *vo = bv[p & 1];
    if (p == passes - 1) {
      if (keys_out != ki) {
        ko = keys_out;
        vo = vals_out;
      } else {
        copy_back = true;  // the output is this pass's input: sort into the other buffer
        ko = ki == bk[0] ? bk[1] : bk[0];
        vo = vi == bv[0] ? bv[1] : bv[0];
      }
    }
    const uint32_t shift = (uint32_t)(p * kBits);
    const bool vec = aligned16(ki) && aligned16(vi);
    if (vec) hipLaunchKernelGGL(k_upsweep<true>, dim3(tiles), dim3(kUThreads), 0, s, ki, n, shift, scratch.block_counts, tiles);
    else hipLaunchKernelGGL(k_upsweep<false>, dim3(tiles), dim3(kUThreads), 0, s, ki, n, shift, scratch.block_counts, tiles);
    hipLaunchKernelGGL(k_scan_rows, dim3(kRadix), dim3(kThreads), 0, s, scratch.block_counts, tiles,
                       scratch.row_totals);
    if (vec)
      hipLaunchKernelGGL(k_downsweep<true>, dim3(down_blocks), dim3(kDThreads), 0, s, ki, vi, ko, vo, n, shift,
                         scratch.block_counts, scratch.row_totals, tiles);
    else
      hipLaunchKernelGGL(k_downsweep<false>, dim3(down_blocks), dim3(kDThreads), 0, s, ki, vi, ko, vo, n, shift,
                         scratch.block_counts, scratch.row_totals, tiles);
    ki = ko;
    vi = vo;
  }
  if (copy_back) hipLaunchKernelGGL(k_copy_pairs, copy_grid, dim3(256), 0, s, ki, vi, keys_out, vals_out, n);
  return hipGetLastError();
}

}  // namespace dem
