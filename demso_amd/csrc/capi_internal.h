// capi_internal.h -- host-side structures behind the C ABI (context, owned
// simulation) and the helpers shared by capi.hip and adjoint.hip.
#pragma once

#include <string>
#include <vector>

#include "dem_internal.h"

struct dem_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool params_set = false;
  dem::StepConst c{};
  dem::Material* d_mat = nullptr;
  uint32_t mat_cap = 0;
  float4* d_pair = nullptr;  // the contact law's per-pair parameters (StepConst::pair)
  uint32_t pair_cap = 0;
  double* d_dpair = nullptr;  // their derivatives for the adjoint (StepConst::dpair)
  uint32_t dpair_cap = 0;
  dem::Instance* d_inst = nullptr;
  uint32_t inst_cap = 0;
  std::vector<dem::Material> h_mat;
  std::vector<dem::Instance> h_inst;
  uint32_t total_cells = 0;  // every key produced by hashing is < total_cells
  uint32_t key_bits = 32;
  dem::SortScratch sort;
  uint32_t* d_copy_tmp = nullptr;
  int bs_hash = 256, bs_reorder = 256, bs_collide = 128, bs_integrate = 256;
  // particles a lane of hash, reorder and integrate (1, 2 or 4; chosen by
  // dem_sim_autotune with the block sizes)
  int per_hash = 1, per_reorder = 1, per_integrate = 1;
  bool fuse = false;      // owned simulation: collide + integrate as one kernel (chosen by dem_sim_autotune)
  int cmode = 0;  // collide's candidate loads: 0 one at a time, 1 one ahead, 2 first of every row
                 // at once (chosen by dem_sim_autotune)
  // owned simulation: re-sorted steps reorder the whole state into the other
  // buffer (the layout of radix-sorted steps) instead of re-sorting it in place
  bool copy = false;
  bool wide = false;  // DEM_STEP_WIDE: collide with 64-bit addresses
  uint64_t generation = 1;  // bumped by every parameter change
  // Controlled particle of the owned simulations: host copy of the last
  // state set, and the device state (moved by every step) followed by the
  // last step's reaction (3 floats).  d_stage_control is the same pair for
  // dem_collide_controlled, so that stage calls leave the owned state alone.
  dem_controlled h_control{};
  dem_controlled* d_control = nullptr;
  float* d_reaction = nullptr;
  dem_controlled* d_stage_control = nullptr;
  float* d_stage_reaction = nullptr;
  // particle type vector of the scene generators (grown as needed)
  uint32_t* d_typevec = nullptr;
  uint32_t typevec_cap = 0;
  // images: particle colours (r, g, b per material row), the per-type
  // render table, the pixel owner map and a one-word result (grown as needed)
  std::vector<float> h_color;
  uint32_t* d_render_types = nullptr;
  size_t render_types_cap = 0;
  uint32_t* d_owner = nullptr;
  size_t owner_cap = 0;
  uint32_t* d_scalar = nullptr;
  size_t scalar_cap = 0;
};

// Every field of the owned simulation's state lives in buffer `cur` (0 or 1)
// between steps; a step writes the next state into buffer cur ^ 1.  The
// dynamic fields (pos, vel, theta, omega) have a third buffer, kScratch, for
// the fully sorted copy that the fused collide + integrate and the adjoint's
// recorded steps read.
constexpr int kScratch = 2;

struct dem_sim {
  dem_context* ctx = nullptr;
  uint32_t n = 0, cells = 0;
  bool rad = false;
  float2 *pos[3] = {}, *vel[3] = {};
  float *theta[3] = {}, *omega[3] = {}, *radius[2] = {};
  uint32_t *id[2] = {}, *type[2] = {}, *loc[2] = {};
  float2* acc = nullptr;
  float *alpha = nullptr, *pressure = nullptr;
  uint32_t *key = nullptr, *idx = nullptr;
  uint32_t *sorted_key = nullptr, *sorted_idx = nullptr;
  // cellStart of the state in buffer b is cs[b] (a step reads the previous
  // one for the re-sort while it clears and rewrites the next); cell_start
  // names cs[cur].  cellEnd is one array (only read where cellStart is set).
  uint32_t* cs[2] = {};
  uint32_t *cell_start = nullptr, *cell_end = nullptr;
  dem::SortScratch sort;
  // Re-sort (dem::launch_resort): skey[b] holds the sorted keys of the last
  // step that wrote state buffer b, sidx the sorted index of the last step.
  uint32_t *skey[2] = {}, *sidx = nullptr;
  dem::ResortScratch resort;
  // In-place re-sort (dem::launch_resort_shift_*): the list and staging of
  // the particles that change slot; pos/vel/theta/omega stage in kScratch.
  // After an in-place step (perm_list) the sorted index is the identity
  // except at the listed slots, and sidx is rebuilt from the list on demand.
  dem::ShiftScratch shift;
  bool perm_list = false;
  // In-place steps: collide marks the runs of 64 slots where no force acted
  // (quiet, ceil(n / 64) words) and leaves their acc and alpha unwritten
  // (+0 by definition, which integrate uses); acc_lazy says the arrays still
  // need those zeros before anything outside the step reads them
  // (sim_acc_complete).
  uint32_t* quiet = nullptr;
  bool acc_lazy = false;
  // host-mapped: [0] the mover count of the last re-sort, [1] the particles
  // the last in-place re-sort listed (slot or key changed)
  uint32_t* h_count = nullptr;
  // sort_mode 0 after a re-sort judged too costly: radix-sorted steps since
  // then, and how many to take before the re-sort is tried again
  uint32_t radix_run = 0, radix_wait = 0;
  bool resort_ok = false;       // state in the order of skey[cur]; cell arrays describe skey[cur]
  int sort_mode = 0;            // 0 auto, 1 radix sort every step, 2 re-sort whenever resort_ok
  int cur = 0;
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t exec[2] = {nullptr, nullptr};      // two radix-sorted steps, by starting buffer
  hipGraphExec_t exec_inc[2] = {nullptr, nullptr};  // two re-sorted steps, by starting buffer
  uint64_t graph_gen = 0;
  // sort_mode 0: an event recorded after every kLookGroup-th replay (or
  // eager step); before it starts a group, the host waits on the event of
  // the group two back, so at most 2 kLookGroup are queued and the counts it
  // picks from trail by less than that
  static constexpr int kLookGroup = 4;
  hipEvent_t look[2] = {nullptr, nullptr};
  // reaction of each sorted particle on the controlled particle, and its
  // per-block sums (ceil(n / kReactBlock) each)
  float *react_x = nullptr, *react_y = nullptr, *react_m = nullptr;
  float *react_bx = nullptr, *react_by = nullptr, *react_bm = nullptr;
  std::vector<void*> allocs;
};

namespace dem {

int fail(dem_context* ctx, int code, const std::string& msg);
int hip_fail(dem_context* ctx, hipError_t e, const char* what);
int check_ctx(dem_context* ctx);
int ensure_device(dem_context* ctx);
bool multi(const dem_context* ctx);
int sim_check(dem_sim* m);
// One step (src/main.cu:431-507) on stream s; ev (7 events) or NULL.
// resort: sort by dem::launch_resort (needs m->resort_ok) instead of the
// radix sort; without a hook and with the split collide/integrate layout the
// state is then re-sorted in place (unless ctx->copy).
// hook: given the step's sorted pre-integrate state, e.g. to store it (the
// adjoint): dyn is the buffer holding the sorted pos/vel/omega (kScratch in
// the full layout, where the hook runs after integrate, or the state buffer
// itself in place, where it runs between collide and integrate), stat the
// buffer of the sorted radius/type; list: the step's sorted index is the
// identity except at the slots listed in m->shift (in place), else it is
// m->sorted_idx.
struct StepHook {
  hipError_t (*fn)(void* user, dem_sim* m, hipStream_t s, int dyn, int stat, bool list) = nullptr;
  void* user = nullptr;
};
hipError_t sim_step_on(dem_sim* m, hipStream_t s, hipEvent_t* ev, bool resort, const StepHook* hook = nullptr);
// Sort method for the next step of m (auto mode: re-sort while the last
// observed mover count is small).
bool sim_use_resort(dem_sim* m);
// Writes the +0 accelerations of the quiet runs of the last in-place step
// into acc and alpha (nothing if they are complete).
hipError_t sim_acc_complete(dem_sim* m, hipStream_t s);
// Whether a re-sorted step of m runs in place (see sim_step_on).
bool sim_inplace(const dem_sim* m, const StepHook* hook = nullptr);

}  // namespace dem

#define DEM_HIP(ctx, call, what)                              \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return dem::hip_fail(ctx, e_, what); \
  } while (0)
