// resort.hip -- stable re-sort of cell keys for a state that is already in
// the sorted order of its previous keys (the sortParticles stage,
// src/functions.cu:336-341, of every step after the first).
//
// After a step the particles sit in the order of the previous step's sorted
// keys prev_key (reorder wrote them so and integrate works in place), and
// only the few particles that crossed a cell boundary ("movers") have a new
// key that differs from prev_key.  The stable sort of the new keys by
// position is the ascending order of the pairs (key[i], i).  The stayers
// keep key = prev_key, which is non-decreasing in i, so they already appear
// in that order; the result is the merge of the stayers with the movers
// sorted by (key, position):
//   stayer i -> slot (stayers before i) + (movers whose (key, pos) < (key[i], i))
//   mover r  -> slot r + (stayers q whose (prev_key[q], q) < (key, pos))
// where r is the mover's rank among the sorted movers.  Every slot is
// written exactly once, and the result equals the full radix sort for any
// number of movers; the cost grows with the number of movers.
//
// The movers' keys, the mover bit mask (bit i%32 of word i/32) and the
// movers per tile of kResortTile particles come from the previous step's
// integrate, which hashes the positions it has just written.  Launches:
//   movers  : one workgroup: per-tile offsets of the movers from the tile
//             counts; for the tiles that have movers, the movers before each
//             32-particle word and the compacted (key, position) list; then
//             the sort of the movers: a bitonic sort of the unique 64-bit
//             (key << 32 | position) values in LDS up to 4096 movers, a
//             stable chunked LSD radix sort through global memory beyond;
//   place   : the two slot formulas above, one particle per lane.  A
//             stayer's mover count is a binary search in the window of
//             sorted movers that falls between the first and last keys of
//             its block's 256 particles (staged in LDS); a
//             mover's stayer count comes from the previous step's cell
//             ranges (cellStart/cellEnd of prev_key) or a binary search
//             over prev_key.
#include <algorithm>

#include "dem_internal.h"

namespace dem {
namespace {

constexpr uint32_t kWordsPerTile = kResortTile / 32;  // 128
constexpr int kSortThreads = 1024;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr uint32_t kBitonicMax = 4096;
constexpr int kPlaceThreads = 256;
constexpr uint32_t kPlaceItems = 8;  // stayers per thread
constexpr uint32_t kPlaceTile = kPlaceThreads * kPlaceItems;
constexpr uint32_t kMoverBlocks = 32;
constexpr uint32_t kWindowMax = 1024;
constexpr int kDigitBits = 8;
constexpr int kDigits = 1 << kDigitBits;

__device__ __forceinline__ uint64_t composite(uint32_t key, uint32_t pos) { return ((uint64_t)key << 32) | pos; }

// Movers among positions [0, p): the tile's offset where p's tile has no
// movers (its word prefixes are not written), else the word prefix plus the
// movers before p in its word.  All four words are loaded at once (one
// round trip); the prefix is used only where it was written.
__device__ __forceinline__ uint32_t movers_before(uint32_t p, uint32_t n, uint32_t m, const uint32_t* __restrict__ bits,
                                                  const uint32_t* __restrict__ word_prefix,
                                                  const uint32_t* __restrict__ tile_off) {
  if (p >= n) return m;
  const uint32_t t = p / kResortTile, w = p >> 5, b = p & 31u;
  const uint32_t o = tile_off[t], o1 = tile_off[t + 1u];
  const uint32_t wp = word_prefix[w], bw = bits[w];
  if (o1 == o) return o;
  return wp + (uint32_t)__popc(bw & ((1u << b) - 1u));
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

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}

// Lanes of this wave (within `valid`) whose digit equals this lane's digit.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
  uint64_t peers = valid;
#pragma unroll
  for (int b = 0; b < kDigitBits; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// Exclusive scan of one value per thread over a 1024-thread block.
__device__ __forceinline__ uint32_t block_scan_1024(uint32_t x, uint32_t* wsum) {
  const uint32_t lane = __lane_id(), w = threadIdx.x / kWave;
  const uint32_t inc = wave_inclusive_scan(x);
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0;
  for (int k = 0; k < (int)w; ++k) off += wsum[k];
  __syncthreads();
  return off + inc - x;
}

// Compare-exchange of bitonic stage (k, j) for element i held in a register;
// its partner i ^ j (j < 64) is held by lane ^ j of the same wave.
__device__ __forceinline__ uint64_t bitonic_lane(uint64_t x, uint32_t i, uint32_t k, uint32_t j) {
  const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  const uint64_t y = ((uint64_t)(uint32_t)__shfl_xor((int)hi, (int)j, kWave) << 32) |
                     (uint32_t)__shfl_xor((int)lo, (int)j, kWave);
  const bool keep_min = ((i & j) == 0) == ((i & k) == 0);
  return keep_min ? (x < y ? x : y) : (x < y ? y : x);
}

// One workgroup: the movers, compacted in position order, and their stable
// sort by key.
//  * tile_off[t] = movers in tiles before t (tile_off[tiles] = the count),
//    from the per-tile counts, which are then cleared for the next step's
//    integrate (it accumulates into them); the offsets and the list of the
//    tiles that have movers stay in LDS up to kLdsTiles tiles;
//  * for every tile that has movers (one wave per tile): the movers before
//    each of its 32-particle words (word_prefix; not written for the other
//    tiles, where tile_off gives the answer) and its movers, in position
//    order, as (key << 32 | position) values in LDS up to kBitonicMax movers,
//    else appended to (mkey[0], mpos[0]);
//  * the sort: a bitonic sort of those unique values up to kBitonicMax movers
//    (the stages whose pairs lie within a wave's 64 values in registers, by
//    lane shuffles; only the wider ones through LDS, with a barrier each), a
//    stable chunked LSD radix sort through global memory beyond.
// info[0] = count, info[1] = the buffer (0 or 1) that holds the sorted movers.
constexpr uint32_t kLdsTiles = 2048;
__global__ __launch_bounds__(kSortThreads) void k_resort_movers(
    uint32_t* __restrict__ tile_count, uint32_t tiles, const uint32_t* __restrict__ bits,
    const uint32_t* __restrict__ key, uint32_t n, uint32_t* __restrict__ tile_off, uint32_t* __restrict__ tile_list,
    uint32_t* __restrict__ word_prefix, uint32_t* mkey0, uint32_t* mpos0, uint32_t* mkey1, uint32_t* mpos1,
    uint32_t key_bits, uint32_t* __restrict__ info, uint32_t* __restrict__ host_count,
    uint32_t* __restrict__ shift_count) {
  __shared__ uint64_t v[kBitonicMax];
  __shared__ uint32_t wsum[kSortWaves];
  __shared__ uint32_t wcnt[kSortWaves][kDigits];
  __shared__ uint32_t base[kDigits];
  __shared__ uint32_t ctot[kDigits];
  __shared__ uint32_t s_off[kLdsTiles];
  __shared__ uint32_t s_list[kLdsTiles];
  __shared__ uint32_t s_carry, s_nz;
  const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid / kWave;
  const bool lds_tiles = tiles <= kLdsTiles;
  // the in-place placement that follows appends to these lists, under a
  // new epoch
  if (shift_count) {
    if (tid < 2) shift_count[tid] = 0;
    if (tid == 2) shift_count[2] = shift_count[2] + 1u;
  }
  if (tid == 0) {
    s_carry = 0;
    s_nz = 0;
  }
  __syncthreads();
  for (uint32_t c0 = 0; c0 < tiles; c0 += kSortThreads) {
    const uint32_t t = c0 + tid;
    const uint32_t cnt = t < tiles ? tile_count[t] : 0u;
    const uint32_t ex = block_scan_1024(cnt, wsum);
    const uint32_t carry = s_carry;
    if (t < tiles) {
      tile_off[t] = carry + ex;
      tile_count[t] = 0;
      if (lds_tiles) s_off[t] = carry + ex;
      if (cnt) {
        const uint32_t slot = atomicAdd(&s_nz, 1u);
        if (lds_tiles) s_list[slot] = t;
        else tile_list[slot] = t;
      }
    }
    __syncthreads();
    if (tid == kSortThreads - 1u) s_carry = carry + ex + cnt;
    __syncthreads();
  }
  const uint32_t m = s_carry, nz = s_nz;
  if (tid == 0) {
    tile_off[tiles] = m;
    info[0] = m;
    if (host_count) *host_count = m;
  }
  const bool in_lds = m <= kBitonicMax;
  const uint32_t nwords = (n + 31u) / 32u;
  for (uint32_t k = w; k < nz; k += kSortWaves) {
    const uint32_t t = lds_tiles ? s_list[k] : tile_list[k];
    const uint32_t off = lds_tiles ? s_off[t] : tile_off[t];
    const uint32_t wa = t * kWordsPerTile + lane, wb = wa + (uint32_t)kWave;
    const uint32_t ba = wa < nwords ? bits[wa] : 0u, bb = wb < nwords ? bits[wb] : 0u;
    const uint32_t ca = (uint32_t)__popc(ba), cb = (uint32_t)__popc(bb);
    const uint32_t ia = wave_inclusive_scan(ca), ib = wave_inclusive_scan(cb);
    const uint32_t tot_a = __shfl(ia, kWave - 1, kWave);
    uint32_t ra = off + ia - ca, rb = off + tot_a + ib - cb;
    if (wa < nwords) word_prefix[wa] = ra;
    if (wb < nwords) word_prefix[wb] = rb;
    for (uint32_t x = ba; x; x &= x - 1u) {
      const uint32_t p = wa * 32u + (uint32_t)(__ffs(x) - 1);
      const uint32_t kp = key[p];
      if (in_lds) {
        v[ra] = composite(kp, p);
      } else {
        mkey0[ra] = kp;
        mpos0[ra] = p;
      }
      ++ra;
    }
    for (uint32_t x = bb; x; x &= x - 1u) {
      const uint32_t p = wb * 32u + (uint32_t)(__ffs(x) - 1);
      const uint32_t kp = key[p];
      if (in_lds) {
        v[rb] = composite(kp, p);
      } else {
        mkey0[rb] = kp;
        mpos0[rb] = p;
      }
      ++rb;
    }
  }
  if (m == 0u) {
    if (tid == 0) info[1] = 0;
    return;
  }

  if (in_lds) {
    uint32_t P = 1;
    while (P < m) P <<= 1;
    for (uint32_t i = m + tid; i < P; i += kSortThreads) v[i] = ~0ull;  // padding sorts last
    __syncthreads();  // the compacted values and the padding are read by every wave below
    constexpr int kR = (int)(kBitonicMax / kSortThreads);  // values a thread holds
    // a wave works on its r-th group of 64 values only when the group starts
    // below P (uniform over the wave: the shuffles need every lane)
    const uint32_t wave_first = tid - lane;
    uint64_t x[kR];
    // stages k = 2 .. 64: every pair lies within one wave's 64 values
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      const uint32_t i = tid + (uint32_t)r * kSortThreads;
      x[r] = i < P ? v[i] : ~0ull;
    }
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      if (wave_first + (uint32_t)r * kSortThreads >= P) continue;
      for (uint32_t k = 2; k <= 64u && k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) x[r] = bitonic_lane(x[r], tid + (uint32_t)r * kSortThreads, k, j);
    }
    // wider stages: pairs across waves through LDS, the rest in registers
    for (uint32_t k = 128; k <= P; k <<= 1) {
#pragma unroll
      for (int r = 0; r < kR; ++r) {
        const uint32_t i = tid + (uint32_t)r * kSortThreads;
        if (i < P) v[i] = x[r];
      }
      __syncthreads();
      for (uint32_t j = k >> 1; j >= 64u; j >>= 1) {
        for (uint32_t i = tid; i < P; i += kSortThreads) {
          const uint32_t l = i ^ j;
          if (l > i) {
            const uint64_t a = v[i], b = v[l];
            const bool up = (i & k) == 0;
            if ((a > b) == up) {
              v[i] = b;
              v[l] = a;
            }
          }
        }
        __syncthreads();
      }
#pragma unroll
      for (int r = 0; r < kR; ++r) {
        const uint32_t i = tid + (uint32_t)r * kSortThreads;
        x[r] = i < P ? v[i] : ~0ull;
      }
      __syncthreads();  // every value is read before the next k writes
#pragma unroll
      for (int r = 0; r < kR; ++r) {
        if (wave_first + (uint32_t)r * kSortThreads >= P) continue;
        for (uint32_t j = 32; j > 0; j >>= 1) x[r] = bitonic_lane(x[r], tid + (uint32_t)r * kSortThreads, k, j);
      }
    }
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      const uint32_t i = tid + (uint32_t)r * kSortThreads;
      if (i < m) {
        mkey0[i] = (uint32_t)(x[r] >> 32);
        mpos0[i] = (uint32_t)x[r];
      }
    }
    if (tid == 0) info[1] = 0;
    return;
  }
  __syncthreads();  // the compacted movers are read by the whole workgroup below

  // stable LSD radix sort, one digit per pass, chunks of 1024 in order
  uint32_t* kb[2] = {mkey0, mkey1};
  uint32_t* pb[2] = {mpos0, mpos1};
  const uint32_t passes = (key_bits + kDigitBits - 1) / kDigitBits;
  uint32_t src = 0;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (uint32_t pass = 0; pass < passes; ++pass) {
    const uint32_t shift = pass * kDigitBits;
    if (tid < (uint32_t)kDigits) base[tid] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < m; i += kSortThreads) atomicAdd(&base[(kb[src][i] >> shift) & (kDigits - 1)], 1u);
    __syncthreads();
    {
      const uint32_t x = tid < (uint32_t)kDigits ? base[tid] : 0u;
      const uint32_t ex = block_scan_1024(x, wsum);
      if (tid < (uint32_t)kDigits) base[tid] = ex;
      __syncthreads();
    }
    for (uint32_t c0 = 0; c0 < m; c0 += kSortThreads) {
      for (uint32_t k = tid; k < kSortWaves * kDigits; k += kSortThreads) (&wcnt[0][0])[k] = 0;
      __syncthreads();
      const uint32_t i = c0 + tid;
      const bool ok = i < m;
      const uint32_t key = ok ? kb[src][i] : 0u, pos = ok ? pb[src][i] : 0u;
      const uint32_t d = (key >> shift) & (kDigits - 1);
      const uint64_t peers = match_digit(d, __ballot(ok));
      const uint32_t rank = (uint32_t)__popcll(peers & lt);
      if (ok && lane == (uint32_t)(__ffsll((unsigned long long)peers) - 1)) wcnt[w][d] = (uint32_t)__popcll(peers);
      __syncthreads();
      if (tid < (uint32_t)kDigits) {
        uint32_t run = 0;
        for (int k = 0; k < kSortWaves; ++k) {
          const uint32_t c = wcnt[k][tid];
          wcnt[k][tid] = run;
          run += c;
        }
        ctot[tid] = run;  // this chunk's count of digit tid
      }
      __syncthreads();
      if (ok) {
        const uint32_t slot = base[d] + wcnt[w][d] + rank;
        if (slot < m) {  // always true for a consistent ranking; guards the store
          kb[src ^ 1u][slot] = key;
          pb[src ^ 1u][slot] = pos;
        }
      }
      __syncthreads();
      if (tid < (uint32_t)kDigits) base[tid] += ctot[tid];
      __syncthreads();
    }
    src ^= 1u;
  }
  if (tid == 0) info[1] = src;
}

// Count of sorted movers whose (key, pos) is below x, in [lo, hi).
__device__ __forceinline__ uint32_t movers_below(const uint32_t* __restrict__ sk, const uint32_t* __restrict__ sp,
                                                 uint32_t lo, uint32_t hi, uint64_t x) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (composite(sk[mid], sp[mid]) < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Insertion point of a mover's pair (k, p) among the previous sorted pairs
// (prev_key[q], q): the number of them that are smaller.  From the cell
// ranges of prev_key when valid (the cell's own range, else the start of the
// first non-empty cell among the next 8, all loaded at once), else a binary
// search over prev_key.
constexpr uint32_t kNextCells = 8;
__device__ __forceinline__ uint32_t insertion_point(uint32_t k, uint32_t p, const uint32_t* __restrict__ prev_key,
                                                    const uint32_t* __restrict__ prev_cell_start,
                                                    const uint32_t* __restrict__ prev_cell_end, uint32_t num_cells,
                                                    int cells_valid, uint32_t n) {
  if (cells_valid && k < num_cells) {
    const uint32_t s0 = prev_cell_start[k], e0 = prev_cell_end[k];
    uint32_t next[kNextCells];
#pragma unroll
    for (uint32_t d = 0; d < kNextCells; ++d) next[d] = k + 1u + d < num_cells ? prev_cell_start[k + 1u + d] : kEmpty;
    if (s0 != kEmpty) return p < s0 ? s0 : (p >= e0 ? e0 : p);
#pragma unroll
    for (uint32_t d = 0; d < kNextCells; ++d)
      if (next[d] != kEmpty) return next[d];
  }
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint32_t q = prev_key[mid];
    if (q < k || (q == k && mid < p)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kPlaceThreads) void k_resort_place(
    const uint32_t* __restrict__ prev_key, const uint32_t* __restrict__ bits,
    const uint32_t* __restrict__ word_prefix, const uint32_t* __restrict__ tile_off, const uint32_t* mkey0, const uint32_t* mpos0, const uint32_t* mkey1,
    const uint32_t* mpos1, const uint32_t* __restrict__ info, const uint32_t* __restrict__ prev_cell_start,
    const uint32_t* __restrict__ prev_cell_end, uint32_t num_cells, int cells_valid, uint32_t* __restrict__ out_key,
    uint32_t* __restrict__ out_idx, uint32_t* __restrict__ clear, uint32_t clear_n, uint32_t n) {
  __shared__ uint64_t win[kWindowMax];
  __shared__ uint32_t s_lo, s_hi;
  const uint32_t m = info[0];
  const uint32_t* sk = info[1] ? mkey1 : mkey0;
  const uint32_t* sp = info[1] ? mpos1 : mpos0;
  const uint32_t tid = threadIdx.x;

  if (blockIdx.x < kMoverBlocks) {
    // movers: slot = rank + stayers before their (key, pos)
    for (uint32_t r = blockIdx.x * kPlaceThreads + tid; r < m; r += kMoverBlocks * kPlaceThreads) {
      const uint32_t k = sk[r], p = sp[r];
      const uint32_t P = insertion_point(k, p, prev_key, prev_cell_start, prev_cell_end, num_cells, cells_valid, n);
      const uint32_t slot = r + P - movers_before(P, n, m, bits, word_prefix, tile_off);
      if (slot < n) {  // always true for a consistent input; guards the store
        out_key[slot] = k;
        out_idx[slot] = p;
      }
    }
    return;
  }

  // stayers: a tile of kPlaceTile particles per block, kPlaceItems per
  // thread at a stride of kPlaceThreads (coalesced).  The movers that sort
  // among the tile's stayers are staged in LDS: all of them when there are
  // at most kWindowMax (no search in global memory at all), else the window
  // [lo, hi) found by two binary searches.
  const uint32_t tile = blockIdx.x - kMoverBlocks;
  const uint32_t tiles = gridDim.x - kMoverBlocks;
  const uint32_t first = tile * kPlaceTile;
  if (first >= n) return;  // (the grid has one block per tile below n)
  const uint32_t last = min(first + kPlaceTile, n) - 1u;
  // all loads of the thread's items first (every slot is written, and none
  // of them depends on the movers); the block's particles lie in one tile of
  // kResortTile: without movers there, its word prefixes are not written
  // and tile_off gives them
  const uint32_t rt = first / kResortTile, tile_base = tile_off[rt];
  const bool tile_movers = tile_off[rt + 1u] != tile_base;
  uint32_t kk[kPlaceItems], bw[kPlaceItems], wp[kPlaceItems];
#pragma unroll
  for (uint32_t it = 0; it < kPlaceItems; ++it) {
    const uint32_t i = first + it * kPlaceThreads + tid;
    const bool valid = i <= last;
    kk[it] = valid ? prev_key[i] : 0u;  // a stayer's key (movers are skipped below)
    bw[it] = valid ? bits[i >> 5] : 0xFFFFFFFFu;
    const uint32_t wraw = valid ? word_prefix[i >> 5] : 0u;
    wp[it] = valid ? (tile_movers ? wraw : tile_base) : 0u;
  }
  if (clear) {
    // this tile's share of the next step's cellStart reset
    const uint32_t per = (clear_n + tiles - 1u) / tiles;
    const uint32_t c0 = min(tile * per, clear_n), c1 = min(c0 + per, clear_n);
    for (uint32_t k = c0 + tid; k < c1; k += kPlaceThreads) clear[k] = kEmpty;
  }
  uint32_t lo = 0, hi = m;
  if (m > kWindowMax) {
    if (tid == 0) s_lo = movers_below(sk, sp, 0, m, composite(prev_key[first], first));
    if (tid == 1) s_hi = movers_below(sk, sp, 0, m, composite(prev_key[last], last));
    __syncthreads();
    lo = s_lo;
    hi = s_hi;
  }
  const bool staged = hi - lo <= kWindowMax;
  if (staged)
    for (uint32_t j = tid; j < hi - lo; j += kPlaceThreads) win[j] = composite(sk[lo + j], sp[lo + j]);
  __syncthreads();
#pragma unroll
  for (uint32_t it = 0; it < kPlaceItems; ++it) {
    const uint32_t i = first + it * kPlaceThreads + tid;
    const uint32_t b = i & 31u;
    if (i > last || ((bw[it] >> b) & 1u)) continue;  // past the end, or a mover
    const uint64_t x = composite(kk[it], i);
    uint32_t below;
    if (staged) {
      uint32_t a = 0, e = hi - lo;
      while (a < e) {
        const uint32_t mid = (a + e) >> 1;
        if (win[mid] < x) a = mid + 1;
        else e = mid;
      }
      below = lo + a;
    } else {
      below = movers_below(sk, sp, lo, hi, x);
    }
    // i - movers_before(i) + below
    const uint32_t slot = i - (wp[it] + (uint32_t)__popc(bw[it] & ((1u << b) - 1u))) + below;
    if (slot < n) {  // always true for a consistent input; guards the store
      out_key[slot] = kk[it];
      out_idx[slot] = i;
    }
  }
}

// ---- in-place re-sort ------------------------------------------------------

// Wave-aggregated append: one atomic per wave, slots in lane order.  Every
// lane of the wave must call it (uniform control flow).
__device__ __forceinline__ uint32_t wave_append(bool emit, uint32_t* counter) {
  const uint64_t b = __ballot(emit);
  if (b == 0) return 0;
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)b) - 1);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(b));
  base = __shfl(base, (int)leader, kWave);
  return base + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
}

// Sorted movers whose key is below k (keys are non-decreasing in rank); sk
// in global memory or LDS.
__device__ __forceinline__ uint32_t keys_below(const uint32_t* sk, uint32_t m, uint32_t k) {
  uint32_t lo = 0, hi = m;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sk[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <bool kRadius>
__device__ __forceinline__ void stage(const ShiftScratch& sh, const ShiftState& st, uint32_t at, uint32_t from,
                                      uint32_t to, uint32_t key, uint32_t epoch) {
  sh.tag[to] = ((uint64_t)epoch << 32) | key;
  sh.dst[at] = to;
  sh.src[at] = from;
  sh.key[at] = key;
  sh.pos[at] = st.pos[from];
  sh.vel[at] = st.vel[from];
  sh.theta[at] = st.theta[from];
  sh.omega[at] = st.omega[from];
  sh.id[at] = st.id[from];
  sh.type[at] = st.type[from];
  if (kRadius) sh.radius[at] = st.radius[from];
}

// The placement of k_resort_place, listing only what changes: every mover
// (its key changed) and every stayer whose slot differs from its position,
// with its state staged, plus the old cells of movers that end up empty.
// A stayer tile is skipped whole when none of its stayers can move: no
// mover lies in it, no mover's pair sorts among its pairs, and the movers
// sorting below it equal the movers before it.
template <bool kRadius>
__global__ __launch_bounds__(kPlaceThreads) void k_shift_place(
    ShiftState st, ShiftScratch sh, const uint32_t* __restrict__ bits,
    const uint32_t* __restrict__ word_prefix, const uint32_t* __restrict__ tile_off, const uint32_t* mkey0, const uint32_t* mpos0, const uint32_t* mkey1,
    const uint32_t* mpos1, const uint32_t* __restrict__ info, uint32_t num_cells, uint32_t n) {
  __shared__ uint64_t win[kWindowMax];
  __shared__ uint32_t s_lo, s_hi, s_bf, s_bl;
  const uint32_t m = info[0];
  if (m == 0) return;
  const uint32_t* sk = info[1] ? mkey1 : mkey0;
  const uint32_t* sp = info[1] ? mpos1 : mpos0;
  const uint32_t* prev_key = st.skey;
  const uint32_t epoch = sh.count[2];
  const uint32_t tid = threadIdx.x;

  if (blockIdx.x < kMoverBlocks) {
    // the sorted movers' keys in LDS when they fit (the arriving counts of
    // the old cells are binary searches over them)
    uint32_t* mk = reinterpret_cast<uint32_t*>(win);
    const bool keys_lds = m <= 2u * kWindowMax;
    if (keys_lds)
      for (uint32_t j = tid; j < m; j += kPlaceThreads) mk[j] = sk[j];
    __syncthreads();
    for (uint32_t r0 = blockIdx.x * kPlaceThreads; r0 < m; r0 += kMoverBlocks * kPlaceThreads) {
      const uint32_t r = r0 + tid;
      const bool ok = r < m;
      uint32_t k = 0, p = 0, slot = n, ko = kEmpty;
      if (ok) {
        k = keys_lds ? mk[r] : sk[r];
        p = sp[r];
        if (p < n) {
          ko = prev_key[p];  // the old cell, loaded with the insertion point's words
          const uint32_t P = insertion_point(k, p, prev_key, st.cell_start, st.cell_end, num_cells, 1, n);
          slot = r + P - movers_before(P, n, m, bits, word_prefix, tile_off);
        }
      }
      const bool emit = ok && p < n && slot < n;
      const uint32_t at = wave_append(emit, &sh.count[0]);
      if (emit && at < sh.capacity) stage<kRadius>(sh, st, at, p, slot, k, epoch);
      // the mover's old cell is left empty when all its particles moved
      // out and none moved in (checked by the cell's first mover only)
      bool emptied = false;
      if (emit && ko < num_cells) {
        const uint32_t s0 = st.cell_start[ko], e0 = st.cell_end[ko];
        if (s0 != kEmpty && s0 <= p && p < e0) {
          const uint32_t mb0 = movers_before(s0, n, m, bits, word_prefix, tile_off);
          const uint32_t mbp = movers_before(p, n, m, bits, word_prefix, tile_off);
          const uint32_t mbe = movers_before(e0, n, m, bits, word_prefix, tile_off);
          if (mbp == mb0) {
            const uint32_t entering = keys_lds ? keys_below(mk, m, ko + 1u) - keys_below(mk, m, ko)
                                               : keys_below(sk, m, ko + 1u) - keys_below(sk, m, ko);
            emptied = e0 - s0 == mbe - mb0 && entering == 0;
          }
        }
      }
      const uint32_t ea = wave_append(emptied, &sh.count[1]);
      if (emptied && ea < sh.capacity) sh.emptied[ea] = ko;
    }
    return;
  }

  const uint32_t tile = blockIdx.x - kMoverBlocks;
  const uint32_t first = tile * kPlaceTile;
  if (first >= n) return;
  const uint32_t last = min(first + kPlaceTile, n) - 1u;
  // the block's words that do not depend on the movers go out with the
  // window's loads: its first and last keys, the movers before its ends,
  // and (the block's particles lie in one tile of kResortTile; without
  // movers there, its word prefixes are not written and tile_off gives
  // them) the tile's offsets
  const uint32_t key_first = prev_key[first], key_last = prev_key[last];
  const uint32_t mb_first = movers_before(first, n, m, bits, word_prefix, tile_off);
  const uint32_t mb_end = movers_before(last + 1u, n, m, bits, word_prefix, tile_off);
  const uint32_t rt = first / kResortTile, tile_base = tile_off[rt];
  const bool tile_movers = tile_off[rt + 1u] != tile_base;
  uint32_t lo = 0, hi = m;
  if (m > kWindowMax) {
    if (tid == 0) s_lo = movers_below(sk, sp, 0, m, composite(key_first, first));
    if (tid == 1) s_hi = movers_below(sk, sp, 0, m, composite(key_last, last));
    __syncthreads();
    lo = s_lo;
    hi = s_hi;
  }
  const bool staged = hi - lo <= kWindowMax;
  if (staged)
    for (uint32_t j = tid; j < hi - lo; j += kPlaceThreads) win[j] = composite(sk[lo + j], sp[lo + j]);
  __syncthreads();
  if (tid < 2) {
    // movers sorting below the tile's first and last pairs
    uint32_t b;
    if (m > kWindowMax) {
      b = tid == 0 ? lo : hi;
    } else {
      const uint32_t q = tid == 0 ? first : last;
      const uint64_t x = composite(tid == 0 ? key_first : key_last, q);
      uint32_t a = 0, e = m;
      while (a < e) {
        const uint32_t mid = (a + e) >> 1;
        if (win[mid] < x) a = mid + 1;
        else e = mid;
      }
      b = a;
    }
    if (tid == 0) s_bf = b;
    else s_bl = b;
  }
  __syncthreads();
  if (mb_first == mb_end && s_bf == s_bl && s_bf == mb_first) return;  // uniform over the block
  uint32_t kk[kPlaceItems], bw[kPlaceItems], wp[kPlaceItems];
#pragma unroll
  for (uint32_t it = 0; it < kPlaceItems; ++it) {
    const uint32_t i = first + it * kPlaceThreads + tid;
    const bool valid = i <= last;
    kk[it] = valid ? prev_key[i] : 0u;  // a stayer's key (movers are skipped below)
    bw[it] = valid ? bits[i >> 5] : 0xFFFFFFFFu;
    wp[it] = valid ? (tile_movers ? word_prefix[i >> 5] : tile_base) : 0u;
  }
#pragma unroll
  for (uint32_t it = 0; it < kPlaceItems; ++it) {
    const uint32_t i = first + it * kPlaceThreads + tid;
    const uint32_t b = i & 31u;
    const bool stayer = i <= last && !((bw[it] >> b) & 1u);
    uint32_t slot = i;
    if (stayer) {
      const uint64_t x = composite(kk[it], i);
      uint32_t below;
      if (staged) {
        uint32_t a = 0, e = hi - lo;
        while (a < e) {
          const uint32_t mid = (a + e) >> 1;
          if (win[mid] < x) a = mid + 1;
          else e = mid;
        }
        below = lo + a;
      } else {
        below = movers_below(sk, sp, lo, hi, x);
      }
      slot = i - (wp[it] + (uint32_t)__popc(bw[it] & ((1u << b) - 1u))) + below;
    }
    const bool emit = stayer && slot != i && slot < n;
    const uint32_t at = wave_append(emit, &sh.count[0]);
    if (emit && at < sh.capacity) stage<kRadius>(sh, st, at, i, slot, kk[it], epoch);
  }
}

// The new sorted key of slot q: from its tag if this step listed it, else
// its unchanged key (which no lane writes).
__device__ __forceinline__ uint32_t new_key(const uint32_t* skey, const uint64_t* __restrict__ tag, uint32_t q,
                                            uint32_t epoch) {
  const uint64_t t = tag[q];
  return (uint32_t)(t >> 32) == epoch ? (uint32_t)t : skey[q];
}

// The boundary test of reorderDataAndFindCellStartD (src/particles_kernel.
// cuh:189-212) at slot q of the new sorted keys (q == n: the last cell's end).
__device__ __forceinline__ void cell_record(const uint32_t* skey, const uint64_t* __restrict__ tag, uint32_t epoch,
                                            uint32_t q, uint32_t* cell_start, uint32_t* cell_end, uint32_t num_cells,
                                            uint32_t n) {
  if (q == n) {
    const uint32_t h = new_key(skey, tag, n - 1u, epoch);
    if (h < num_cells) cell_end[h] = n;
    return;
  }
  const uint32_t h = new_key(skey, tag, q, epoch);
  if (q == 0) {
    if (h < num_cells) cell_start[h] = 0;
    return;
  }
  const uint32_t hp = new_key(skey, tag, q - 1u, epoch);
  if (h != hp) {
    if (h < num_cells) cell_start[h] = q;
    if (hp < num_cells) cell_end[hp] = q;
  }
}

// Staged state -> new slots; cellStart of emptied cells -> EMPTY; the cell
// boundaries at every listed new slot and the slot after it (the only places
// where the key sequence changed; two lanes may write the same record, with
// the same value).
template <bool kRadius>
__global__ __launch_bounds__(256) void k_shift_apply(ShiftState st, ShiftScratch sh, uint32_t num_cells, uint32_t n) {
  const uint32_t listed = sh.count[0];
  const uint32_t S = min(listed, sh.capacity), E = min(sh.count[1], sh.capacity), epoch = sh.count[2];
  const uint32_t stride = gridDim.x * blockDim.x;
  if (sh.host_listed && blockIdx.x == 0 && threadIdx.x == 0) *sh.host_listed = listed;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < S; k += stride) {
    const uint32_t d = sh.dst[k];
    if (d >= n) continue;
    st.pos[d] = sh.pos[k];
    st.vel[d] = sh.vel[k];
    st.theta[d] = sh.theta[k];
    st.omega[d] = sh.omega[k];
    const uint32_t id = sh.id[k];
    st.id[d] = id;
    if (id < n) st.loc[id] = d;
    st.type[d] = sh.type[k];
    if (kRadius) st.radius[d] = sh.radius[k];
    st.skey[d] = sh.key[k];
    cell_record(st.skey, sh.tag, epoch, d, st.cell_start, st.cell_end, num_cells, n);
    cell_record(st.skey, sh.tag, epoch, d + 1u, st.cell_start, st.cell_end, num_cells, n);
  }
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < E; k += stride) {
    const uint32_t c = sh.emptied[k];
    if (c < num_cells) st.cell_start[c] = kEmpty;
  }
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ index, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) index[i] = i;
}

__global__ __launch_bounds__(256) void k_shift_index(const uint32_t* __restrict__ dst,
                                                     const uint32_t* __restrict__ src,
                                                     const uint32_t* __restrict__ count, uint32_t capacity,
                                                     uint32_t* __restrict__ index, uint32_t n) {
  const uint32_t S = min(count[0], capacity);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < S; k += gridDim.x * blockDim.x) {
    const uint32_t d = dst[k];
    if (d < n) index[d] = src[k];
  }
}

// Workgroups of the list kernels (grid-stride over a count known only on
// the device): enough to cover a large list in a few passes.
uint32_t list_blocks(uint32_t n) { return std::max(1u, std::min(128u, (n + 2047u) / 2048u)); }
// Workgroups of the apply kernel: a step may list most of the particles
// (bursts of movers shift long runs of them), so one per 4,096 particles,
// at most four per CU.
uint32_t apply_blocks(uint32_t n) { return std::max(1u, std::min(1024u, (n + 4095u) / 4096u)); }

}  // namespace

hipError_t launch_resort_shift_plan(hipStream_t s, const uint32_t* key, const ShiftState& st, const ShiftScratch& sh,
                                    uint32_t num_cells, uint32_t key_bits, const ResortScratch& r, uint32_t n) {
  if (n == 0) return hipSuccess;
  if (n > r.capacity || !r.bits || !r.word_prefix || !r.tile_count || !r.tile_off || !r.tile_list || !r.mkey[0] ||
      !r.mkey[1] || !r.mpos[0] || !r.mpos[1] || !r.info || key_bits > 32 || n > sh.capacity || !sh.count || !sh.dst || !sh.src || !sh.key ||
      !sh.emptied || !sh.tag || !sh.pos || !sh.vel || !sh.theta || !sh.omega || !sh.id || !sh.type || !st.pos || !st.vel ||
      !st.theta || !st.omega || !st.id || !st.type || !st.skey || !st.cell_start || !st.cell_end ||
      (st.radius && !sh.radius))
    return hipErrorInvalidValue;
  const uint32_t tiles = (n + kResortTile - 1u) / kResortTile;
  hipLaunchKernelGGL(k_resort_movers, dim3(1), dim3(kSortThreads), 0, s, r.tile_count, tiles, r.bits, key, n,
                     r.tile_off, r.tile_list, r.word_prefix, r.mkey[0], r.mpos[0], r.mkey[1], r.mpos[1], key_bits,
                     r.info, r.host_count, sh.count);
  const dim3 grid(kMoverBlocks + (n + kPlaceTile - 1u) / kPlaceTile);
  if (st.radius)
    hipLaunchKernelGGL(k_shift_place<true>, grid, dim3(kPlaceThreads), 0, s, st, sh, r.bits, r.word_prefix,
                       r.tile_off, r.mkey[0], r.mpos[0], r.mkey[1], r.mpos[1], r.info, num_cells, n);
  else
    hipLaunchKernelGGL(k_shift_place<false>, grid, dim3(kPlaceThreads), 0, s, st, sh, r.bits, r.word_prefix,
                       r.tile_off, r.mkey[0], r.mpos[0], r.mkey[1], r.mpos[1], r.info, num_cells, n);
  return hipGetLastError();
}

hipError_t launch_resort_shift_apply(hipStream_t s, const ShiftState& st, const ShiftScratch& sh, uint32_t num_cells,
                                     uint32_t n) {
  if (n == 0) return hipSuccess;
  if (n > sh.capacity || !sh.count || !sh.dst || !sh.tag || !st.skey || !st.loc || !st.cell_start ||
      !st.cell_end || (st.radius && !sh.radius))
    return hipErrorInvalidValue;
  const dim3 grid(apply_blocks(n));
  if (st.radius) hipLaunchKernelGGL(k_shift_apply<true>, grid, dim3(256), 0, s, st, sh, num_cells, n);
  else hipLaunchKernelGGL(k_shift_apply<false>, grid, dim3(256), 0, s, st, sh, num_cells, n);
  return hipGetLastError();
}

hipError_t launch_shift_index(hipStream_t s, const ShiftScratch& sh, uint32_t* index, uint32_t n, bool identity_done) {
  if (n == 0) return hipSuccess;
  if (n > sh.capacity || !sh.count || !sh.dst || !sh.src || !index) return hipErrorInvalidValue;
  if (!identity_done) hipLaunchKernelGGL(k_iota, dim3((n + 255u) / 256u), dim3(256), 0, s, index, n);
  hipLaunchKernelGGL(k_shift_index, dim3(list_blocks(n)), dim3(256), 0, s, sh.dst, sh.src, sh.count, sh.capacity,
                     index, n);
  return hipGetLastError();
}

hipError_t launch_resort(hipStream_t s, const uint32_t* key, const uint32_t* prev_key, const uint32_t* prev_cell_start,
                         const uint32_t* prev_cell_end, uint32_t num_cells, bool cells_valid, uint32_t key_bits,
                         uint32_t* out_key, uint32_t* out_idx, const ResortScratch& r, uint32_t n, uint32_t* clear,
                         uint32_t clear_n) {
  if (n == 0) return hipSuccess;
  if (n > r.capacity || !r.bits || !r.word_prefix || !r.tile_count || !r.tile_off || !r.tile_list || !r.mkey[0] ||
      !r.mkey[1] || !r.mpos[0] || !r.mpos[1] || !r.info || key_bits > 32)
    return hipErrorInvalidValue;
  const uint32_t tiles = (n + kResortTile - 1u) / kResortTile;
  hipLaunchKernelGGL(k_resort_movers, dim3(1), dim3(kSortThreads), 0, s, r.tile_count, tiles, r.bits, key, n,
                     r.tile_off, r.tile_list, r.word_prefix, r.mkey[0], r.mpos[0], r.mkey[1], r.mpos[1], key_bits,
                     r.info, r.host_count, nullptr);
  const uint32_t place_tiles = (n + kPlaceTile - 1u) / kPlaceTile;
  hipLaunchKernelGGL(k_resort_place, dim3(kMoverBlocks + place_tiles), dim3(kPlaceThreads), 0, s, prev_key, r.bits,
                     r.word_prefix, r.tile_off, r.mkey[0], r.mpos[0], r.mkey[1], r.mpos[1], r.info, prev_cell_start,
                     prev_cell_end, num_cells, cells_valid ? 1 : 0, out_key, out_idx, clear, clear_n, n);
  return hipGetLastError();
}

}  // namespace dem
