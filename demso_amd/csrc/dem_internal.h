// dem_internal.h -- types shared by the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/demso.h"

namespace dem {

constexpr uint32_t kEmpty = DEM_EMPTY_CELL;
constexpr int kWave = 64;
constexpr int kRowWords = DEM_ROW_WORDS;

// Material row as the kernels read it (48 bytes, one row per type).  Fields
// of ParticleProperties (src/main.cuh:39-54) that the step uses, plus the
// areal density for per-particle radii and the owning instance.
struct alignas(16) Material {
  float radius, mass, inertia;
  float kn, kt, cn;
  float bdamp, mu, attract;
  float areal;
  uint32_t inst, pad;
};

// Instance geometry (32 bytes): wall box, hashing extent, grid, key base.
struct alignas(16) Instance {
  float box[2];
  float extent[2];
  uint32_t grid[2];
  uint32_t cell_base, pad;
};

// Uniform step parameters, passed by value to every kernel (kernarg ->
// scalar registers).  inst0 is instance 0 inlined for the single-system path.
struct StepConst {
  float dt, gx, gy, bkn, bkt;
  uint32_t flags;       // DEM_GRID_CLAMP | DEM_OPEN_TOP
  uint32_t attract;     // any material has a non-zero attraction coefficient
  uint32_t num_types;
  uint32_t num_instances;
  uint32_t cells;       // cells of all instances: every cell index is below it
  const Material* mat;
  const Instance* inst;
  // Per ordered pair of materials (row tA * num_types + tB): the series
  // combinations kn, kt, cn and the mean mu of collideSpheres
  // (src/particles_kernel.cuh:262-268), computed once on the host with the
  // same float operations (seriesAssociation, :36-41).
  const float4* pair;
  // For the adjoint, per ordered pair (6 doubles at 6 * (tA * num_types +
  // tB)): d(series)/d(own value) and d(series)/d(partner's value) of kn, kt
  // and cn, in double; null above kDPairMaxTypes material rows.
  const double* dpair;
  Instance inst0;
};
constexpr uint32_t kDPairMaxTypes = 1024;
// Material rows per context (the pair table holds up to kMaxTypes^2 entries).
constexpr uint32_t kMaxTypes = DEM_MAX_TYPES;

// The controlled particle's state (device) and the per-particle reaction
// vectors the pair with it writes.
struct ControlOut {
  const dem_controlled* state = nullptr;
  float* react_x = nullptr;
  float* react_y = nullptr;
  float* react_m = nullptr;
};
constexpr uint32_t kReactBlock = 512;  // the reference's reduction block (src/functions.cu:475)

// Workgroups are dispatched to the 8 XCDs round-robin (XCD = workgroup id
// mod 8), each XCD with its own L2.  The collide kernels gather the
// neighbour rows of their particles, which lie a few workgroups away in
// sorted order; giving each XCD one contiguous eighth of the workgroups (a
// bijection of the ids) keeps those rows in the L2 of the XCD that reads
// them.  Returns the workgroup's position in that order.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t x = b % kXcds, local = b / kXcds, q = nb / kXcds, r = nb % kXcds;
  return x * q + min(x, r) + local;
}

// Launch helpers (defined in the .hip files).
// per: particles a lane (1, 2 or 4) of hash, reorder and integrate.
hipError_t launch_hash(const StepConst& c, hipStream_t s, int block, const float2* pos,
                       const uint32_t* type, uint32_t* hash, uint32_t* index, uint32_t n, int per = 1);
hipError_t launch_fill(hipStream_t s, uint32_t* p, uint32_t value, uint32_t count);

// ---- re-sort of a state already sorted by its previous keys --------------
// Particles per tile of the mover bit mask and counts.
constexpr uint32_t kResortTile = 4096;
struct ResortScratch {
  uint32_t* bits = nullptr;         // ceil(n/32) mover bits
  uint32_t* word_prefix = nullptr;  // ceil(n/32) movers before each word (written for tiles with movers)
  uint32_t* tile_count = nullptr;   // tiles (accumulated by integrate, cleared by the movers kernel)
  uint32_t* tile_off = nullptr;     // tiles + 1: movers before each tile
  uint32_t* tile_list = nullptr;    // tiles: the tiles that have movers
  uint32_t* mkey[2] = {};           // n each: mover keys (compacted, then sorted)
  uint32_t* mpos[2] = {};           // n each: mover positions
  uint32_t* info = nullptr;         // [0] mover count, [1] buffer holding the sorted movers
  uint32_t* host_count = nullptr;   // device address of a host-mapped mirror of the mover count (or null)
  uint32_t capacity = 0;
};
// Stable sort of the new keys of a state whose order is sorted by prev_key,
// from the mover bits, counts and mover keys written with them (NextKeys):
// writes the sorted keys and the source index of every sorted slot.  Equal
// to a stable sort of key by position for any number of movers; its cost
// grows with the number of movers.  With clear, the placement kernel also
// writes kEmpty to clear[0, clear_n) (the next step's cellStart reset).
// prev_cell_start/prev_cell_end are the
// cell ranges of prev_key when cells_valid (a faster lookup), else unused.
hipError_t launch_resort(hipStream_t s, const uint32_t* key, const uint32_t* prev_key, const uint32_t* prev_cell_start,
                         const uint32_t* prev_cell_end, uint32_t num_cells, bool cells_valid, uint32_t key_bits,
                         uint32_t* out_key, uint32_t* out_idx, const ResortScratch& r, uint32_t n,
                         uint32_t* clear = nullptr, uint32_t clear_n = 0);
// In-place re-sort.  When a state is re-sorted where it lies (instead of
// being reordered into the other buffer), only the particles whose slot or
// key changes are listed, staged and written back:
//   plan  : compact + movers as above, then the placement of the movers and
//           of the stayers whose slot changes, each appended to the list
//           (new slot, old slot, key) with its state copied into staging,
//           and the cells that lose all their particles;
//   apply : the staged state written to the new slots (sorted keys and loc
//           too), cellStart cleared for the emptied cells, and the cell
//           boundaries re-evaluated at every new slot and the slot after it
//           (one kernel: the new key of a listed slot is read from its tag,
//           that of any other slot is its unchanged sorted key).
// The result equals reorderDataAndFindCellStart (src/particles_kernel.cuh:
// 154-233) applied to the stable sort of the new keys, for any number of
// movers: every other cell boundary is the one the previous step wrote.
struct ShiftState {
  float2 *pos = nullptr, *vel = nullptr;
  float *theta = nullptr, *omega = nullptr, *radius = nullptr;
  uint32_t *id = nullptr, *type = nullptr, *loc = nullptr;
  uint32_t* skey = nullptr;  // sorted keys: the previous ones on entry, the new ones after apply
  uint32_t *cell_start = nullptr, *cell_end = nullptr;
};
struct ShiftScratch {
  float2 *pos = nullptr, *vel = nullptr;
  float *theta = nullptr, *omega = nullptr, *radius = nullptr;
  uint32_t *id = nullptr, *type = nullptr, *key = nullptr;
  uint32_t *dst = nullptr, *src = nullptr;  // new and old slot of each listed particle
  uint32_t* emptied = nullptr;              // cells left without particles
  uint32_t* count = nullptr;                // [0] listed particles, [1] emptied cells, [2] step epoch
  // per slot: (epoch << 32 | new key) of the last step that listed it, so
  // that the new key of any slot is known while the state is written back
  uint64_t* tag = nullptr;
  uint32_t* host_listed = nullptr;          // device address of a host-mapped mirror of count[0] (or null)
  uint32_t capacity = 0;                    // n
};
hipError_t launch_resort_shift_plan(hipStream_t s, const uint32_t* key, const ShiftState& st, const ShiftScratch& sh,
                                    uint32_t num_cells, uint32_t key_bits, const ResortScratch& r, uint32_t n);
hipError_t launch_resort_shift_apply(hipStream_t s, const ShiftState& st, const ShiftScratch& sh, uint32_t num_cells,
                                     uint32_t n);
// The sorted index of the last in-place re-sort: identity except at the
// listed slots (index[new] = old).  identity_done: index already holds the
// identity (only the listed slots are written).
hipError_t launch_shift_index(hipStream_t s, const ShiftScratch& sh, uint32_t* index, uint32_t n,
                              bool identity_done = false);
hipError_t launch_reorder(hipStream_t s, int block, uint32_t* cell_start, uint32_t* cell_end,
                          float2* s_pos, float2* s_vel, float* s_theta, float* s_omega,
                          uint32_t* s_id, uint32_t* s_loc, uint32_t* s_type, float* s_radius,
                          const uint32_t* hash, const uint32_t* index, const float2* o_pos,
                          const float2* o_vel, const float* o_theta, const float* o_omega,
                          const uint32_t* o_id, const uint32_t* o_type, const float* o_radius,
                          uint32_t n, uint32_t num_cells, bool full = true, int per = 1);
// launch_collide's mode: the candidate mode (0 one at a time, 1 one ahead, 2
// the first of every row at once, 3 contact lists), plus kWideMode to force 64-bit addressing
// (the plain mode), which states over 2^29 particles or 2^30 cells get anyway.
// quiet (ceil(n / 64) words, or null): collide sets quiet[g] = 1 for each run
// g of 64 slots where no force acted, and then stores neither acc nor alpha
// there (both are +0); launch_integrate_to given the same array reads +0
// for those runs instead of loading them.
constexpr int kWideMode = 4;
hipError_t launch_collide(const StepConst& c, hipStream_t s, int block, const float2* pos,
                          const float2* vel, float2* acc, const float* omega, float* alpha,
                          const uint32_t* type, const float* radius, const uint32_t* cell_start,
                          const uint32_t* cell_end, float* pressure, uint32_t n, int mode = 0,
                          const struct ControlOut* ctl = nullptr, const uint32_t* gat = nullptr,
                          uint32_t* quiet = nullptr);
// The next step's hash, fused into integrate: bit i of `bits` set when the
// cell key of the new position differs from the key the particle is sorted
// by (the hash of its position before integrate), and then key[i] = the new
// key (not written for the others: their key stays the sorted key);
// tile_count[i / kResortTile] += movers (the counts must be zero before).
struct NextKeys {
  uint32_t* key = nullptr;
  uint32_t* bits = nullptr;
  uint32_t* tile_count = nullptr;
};
hipError_t launch_integrate(const StepConst& c, hipStream_t s, int block, float2* pos, float2* vel,
                            const float2* acc, float* theta, float* omega, const float* alpha,
                            const uint32_t* type, const float* radius, float* pressure, uint32_t n,
                            const NextKeys* next = nullptr, int per = 1);
// integrate from one set of state arrays into another (may be the same).
hipError_t launch_integrate_to(const StepConst& c, hipStream_t s, int block, const float2* pos_in,
                               const float2* vel_in, const float* theta_in, const float* omega_in, float2* pos_out,
                               float2* vel_out, float* theta_out, float* omega_out, const float2* acc,
                               const float* alpha, const uint32_t* type, const float* radius, float* pressure,
                               uint32_t n, const NextKeys* next = nullptr, const uint32_t* gat = nullptr,
                               int per = 1, const uint32_t* quiet = nullptr);
// collide + integrate + the next step's hash in one kernel: forces from the
// sorted state *_in, acc/alpha/pressure as collide writes them, the
// integrated state to *_out (a different buffer).
hipError_t launch_collide_integrate(const StepConst& c, hipStream_t s, int block, const float2* pos_in,
                                    const float2* vel_in, const float* omega_in, const float* theta_in,
                                    const uint32_t* type, const float* radius, const uint32_t* cell_start,
                                    const uint32_t* cell_end, float2* acc, float* alpha, float* pressure,
                                    float2* pos_out, float2* vel_out, float* omega_out, float* theta_out,
                                    const NextKeys& next, uint32_t n, int mode = 0,
                                    const struct ControlOut* ctl = nullptr);
// The controlled particle's reaction: per-block tree sums of the three
// per-particle reaction vectors into bx/by/bm (ceil(n/kReactBlock) each),
// their ordered sum into reaction[3], then (move) its kinematic move.
hipError_t launch_control_finish(hipStream_t s, const struct ControlOut& ctl, float* bx, float* by, float* bm,
                                 uint32_t n, float* reaction, dem_controlled* state, bool move);

// Radix sort of (key, value) pairs, stable, ascending, on the low key_bits.
struct SortScratch {
  uint32_t* keys_alt = nullptr;
  uint32_t* vals_alt = nullptr;
  uint32_t* keys_alt2 = nullptr;     // optional second pair of arrays (see launch_sort_to)
  uint32_t* vals_alt2 = nullptr;
  uint32_t* block_counts = nullptr;  // 256 * max_tiles
  uint32_t* row_totals = nullptr;    // 256
  uint32_t capacity = 0;             // max keys
};
// Keys per tile of the radix sort (its per-tile digit counts are 256 words).
constexpr uint32_t kSortTile = 8192;
int sort_passes(uint32_t key_bits);
uint32_t sort_tiles(uint32_t n);
size_t sort_block_count_elems(uint32_t n);
// Stable sort of (keys_in, vals_in) by the low key_bits of the keys into
// (keys_out, vals_out), which may be the input arrays themselves but no
// scratch array.  The input arrays are overwritten (the passes alternate
// between them and the scratch arrays when the scratch has no second pair).
// The last pass writes the output directly, except when the output is that
// pass's input: an in-place sort with an odd number of passes and no second
// pair (or one pass), which ends with a copy.
hipError_t launch_sort_to(hipStream_t s, uint32_t* keys_in, uint32_t* vals_in, uint32_t* keys_out,
                          uint32_t* vals_out, uint32_t n, uint32_t key_bits, const SortScratch& scratch);

}  // namespace dem
