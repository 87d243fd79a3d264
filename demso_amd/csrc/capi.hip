// capi.hip -- the extern "C" boundary of include/demso.h.
//
// Host-side replacements for DemSo's launch wrappers (src/functions.cu) and
// step loop (src/main.cu:371-565): argument checking, parameter tables in
// device memory instead of __constant__ symbols, scratch management, the
// owned double-buffered simulation, hipGraph capture/replay, HIP-event stage
// timing and block-size autotuning.  No call here exits the process.
#include "capi_internal.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <string>
#include <vector>

using dem::Instance;
using dem::Material;
using dem::SortScratch;
using dem::StepConst;

namespace {
thread_local std::string g_create_error;
}  // namespace

namespace dem {

int fail(dem_context* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  else g_create_error = msg;
  return code;
}

int hip_fail(dem_context* ctx, hipError_t e, const char* what) {
  return fail(ctx, e == hipErrorOutOfMemory ? DEM_ERR_ALLOC : DEM_ERR_LAUNCH,
              std::string(what) + ": " + hipGetErrorString(e));
}

bool is_pow2(uint32_t v) { return v && !(v & (v - 1u)); }

uint32_t bits_for(uint32_t total_cells) {
  if (total_cells <= 1) return 0;
  uint32_t v = total_cells - 1u, b = 0;
  while (v) { ++b; v >>= 1; }
  return b;
}

int check_ctx(dem_context* ctx) {
  if (!ctx) return DEM_ERR_INVALID;
  if (!ctx->params_set) return fail(ctx, DEM_ERR_STATE, "parameters not set (call dem_set_params)");
  return DEM_OK;
}

int ensure_device(dem_context* ctx) {
  DEM_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  return DEM_OK;
}

template <class T>
int grow(dem_context* ctx, T** p, uint32_t* cap, uint32_t need) {
  if (*cap >= need && *p) return DEM_OK;
  if (*p) { (void)hipFree(*p); *p = nullptr; *cap = 0; }
  DEM_HIP(ctx, hipMalloc((void**)p, sizeof(T) * (size_t)std::max(need, 1u)), "hipMalloc");
  *cap = need;
  return DEM_OK;
}

void free_sort(SortScratch& s) {
  for (uint32_t* p : {s.keys_alt, s.vals_alt, s.keys_alt2, s.vals_alt2, s.block_counts, s.row_totals})
    if (p) (void)hipFree(p);
  s = SortScratch{};
}

// second: also the second pair of arrays, which spares the in-place sort
// (dem_sort_particles) its copy after an odd number of passes.
int alloc_sort(dem_context* ctx, SortScratch& s, uint32_t n, bool second) {
  if (s.capacity >= n && s.keys_alt && (!second || s.keys_alt2)) return DEM_OK;
  free_sort(s);
  const size_t cnt = dem::sort_block_count_elems(std::max(n, 1u));
  const size_t words = std::max(n, 1u);
  DEM_HIP(ctx, hipMalloc((void**)&s.keys_alt, sizeof(uint32_t) * words), "hipMalloc(sort keys)");
  DEM_HIP(ctx, hipMalloc((void**)&s.vals_alt, sizeof(uint32_t) * words), "hipMalloc(sort vals)");
  if (second) {
    DEM_HIP(ctx, hipMalloc((void**)&s.keys_alt2, sizeof(uint32_t) * words), "hipMalloc(sort keys)");
    DEM_HIP(ctx, hipMalloc((void**)&s.vals_alt2, sizeof(uint32_t) * words), "hipMalloc(sort vals)");
  }
  DEM_HIP(ctx, hipMalloc((void**)&s.block_counts, sizeof(uint32_t) * cnt), "hipMalloc(sort counts)");
  DEM_HIP(ctx, hipMalloc((void**)&s.row_totals, sizeof(uint32_t) * 256), "hipMalloc(sort totals)");
  s.capacity = n;
  return DEM_OK;
}

// seriesAssociation (src/particles_kernel.cuh:36-41) in float, as the
// kernels compute it.
float series_host(float a, float b) {
  if (a == 0.0f || b == 0.0f) return 0.0f;
  return (a * b) / (a + b);
}

// d(a b/(a+b))/da = b^2/(a+b)^2 and d/db = a^2/(a+b)^2 in double, 0 where
// seriesAssociation is 0 (the adjoint's formula).
void series_grad_host(double a, double b, double& da, double& db) {
  if (a == 0.0 || b == 0.0) {
    da = db = 0.0;
    return;
  }
  const double inv = 1.0 / ((a + b) * (a + b));
  da = b * b * inv;
  db = a * a * inv;
}

int upload_tables(dem_context* ctx) {
  const uint32_t nt = (uint32_t)ctx->h_mat.size(), ni = (uint32_t)ctx->h_inst.size();
  ctx->c.cells = ctx->total_cells;
  if (int r = grow(ctx, &ctx->d_mat, &ctx->mat_cap, nt)) return r;
  if (int r = grow(ctx, &ctx->d_inst, &ctx->inst_cap, ni)) return r;
  if (nt > dem::kMaxTypes) return fail(ctx, DEM_ERR_INVALID, "at most 4096 material rows per context");
  // the contact law's pair parameters for every ordered pair of materials
  std::vector<float4> pair((size_t)nt * nt);
  for (uint32_t a = 0; a < nt; ++a)
    for (uint32_t b = 0; b < nt; ++b) {
      const Material &A = ctx->h_mat[a], &B = ctx->h_mat[b];
      pair[(size_t)a * nt + b] = make_float4(series_host(A.kn, B.kn), series_host(A.kt, B.kt),
                                             series_host(A.cn, B.cn), (A.mu + B.mu) / 2);
    }
  if (int r = grow(ctx, &ctx->d_pair, &ctx->pair_cap, (uint32_t)pair.size())) return r;
  DEM_HIP(ctx, hipMemcpyAsync(ctx->d_pair, pair.data(), sizeof(float4) * pair.size(), hipMemcpyHostToDevice,
                              ctx->stream),
          "upload material pairs");
  ctx->c.pair = ctx->d_pair;
  ctx->c.dpair = nullptr;
  if (nt <= dem::kDPairMaxTypes) {
    // d(series)/d(a), d(series)/d(b) of kn, kt, cn for the adjoint, in double
    std::vector<double> dp(6 * (size_t)nt * nt);
    for (uint32_t a = 0; a < nt; ++a)
      for (uint32_t b = 0; b < nt; ++b) {
        const Material &A = ctx->h_mat[a], &B = ctx->h_mat[b];
        double* d = dp.data() + 6 * ((size_t)a * nt + b);
        series_grad_host(A.kn, B.kn, d[0], d[1]);
        series_grad_host(A.kt, B.kt, d[2], d[3]);
        series_grad_host(A.cn, B.cn, d[4], d[5]);
      }
    if (int r = grow(ctx, &ctx->d_dpair, &ctx->dpair_cap, (uint32_t)dp.size())) return r;
    DEM_HIP(ctx, hipMemcpyAsync(ctx->d_dpair, dp.data(), sizeof(double) * dp.size(), hipMemcpyHostToDevice,
                                ctx->stream),
            "upload material pair derivatives");
    ctx->c.dpair = ctx->d_dpair;
  }
  DEM_HIP(ctx, hipMemcpyAsync(ctx->d_mat, ctx->h_mat.data(), sizeof(Material) * nt, hipMemcpyHostToDevice, ctx->stream),
          "upload materials");
  DEM_HIP(ctx, hipMemcpyAsync(ctx->d_inst, ctx->h_inst.data(), sizeof(Instance) * ni, hipMemcpyHostToDevice, ctx->stream),
          "upload instances");
  DEM_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  ctx->c.mat = ctx->d_mat;
  ctx->c.inst = ctx->d_inst;
  ctx->c.num_types = nt;
  ctx->c.num_instances = ni;
  ctx->c.inst0 = ctx->h_inst[0];
  ctx->generation++;
  return DEM_OK;
}

int validate_instances(dem_context* ctx, const std::vector<Instance>& inst, uint32_t flags,
                       uint32_t* total_cells) {
  uint64_t top = 0;
  for (size_t k = 0; k < inst.size(); ++k) {
    const Instance& in = inst[k];
    if (in.grid[0] == 0 || in.grid[1] == 0) return fail(ctx, DEM_ERR_INVALID, "grid size must be >= 1");
    if (!(flags & DEM_GRID_CLAMP) && !(is_pow2(in.grid[0]) && is_pow2(in.grid[1])))
      return fail(ctx, DEM_ERR_INVALID, "wrapping grids need power-of-two sizes (set DEM_GRID_CLAMP)");
    if (!(in.extent[0] > 0.0f) || !(in.extent[1] > 0.0f) || !isfinite(in.extent[0]) || !isfinite(in.extent[1]))
      return fail(ctx, DEM_ERR_INVALID, "grid extent must be finite and > 0");
    if (!isfinite(in.box[0]) || !isfinite(in.box[1])) return fail(ctx, DEM_ERR_INVALID, "box must be finite");
    const uint64_t end = (uint64_t)in.cell_base + (uint64_t)in.grid[0] * in.grid[1];
    if (end >= 0xFFFFFFFFull) return fail(ctx, DEM_ERR_INVALID, "cell keys exceed 32 bits");
    top = std::max(top, end);
  }
  *total_cells = (uint32_t)top;
  return DEM_OK;
}

Instance make_instance(const float box[2], const float extent[2], const uint32_t grid[2], uint32_t base) {
  Instance in{};
  in.box[0] = box[0];
  in.box[1] = box[1];
  in.extent[0] = (extent[0] == 0.0f && extent[1] == 0.0f) ? box[0] : extent[0];
  in.extent[1] = (extent[0] == 0.0f && extent[1] == 0.0f) ? box[1] : extent[1];
  in.grid[0] = grid[0];
  in.grid[1] = grid[1];
  in.cell_base = base;
  return in;
}

int check_cells(dem_context* ctx, uint32_t num_cells) {
  if (num_cells < ctx->total_cells)
    return fail(ctx, DEM_ERR_INVALID, "numCells is smaller than the configured grid");
  return DEM_OK;
}

bool multi(const dem_context* ctx) { return ctx->c.num_instances > 1; }

// The context's collide mode for launch_collide.
int collide_mode(const dem_context* ctx) { return ctx->cmode | (ctx->wide ? kWideMode : 0); }

}  // namespace dem

using namespace dem;

namespace {

// Particle colours (colorR/G/B of ParticleProperties), kept for images.
std::vector<float> colors_of(const dem_particle_props* types, uint32_t num_types) {
  std::vector<float> c(3 * (size_t)num_types);
  for (uint32_t t = 0; t < num_types; ++t) {
    c[3 * t] = types[t].color_r;
    c[3 * t + 1] = types[t].color_g;
    c[3 * t + 2] = types[t].color_b;
  }
  return c;
}

}  // namespace

extern "C" {

int dem_abi_version(void) { return DEMSO_ABI_VERSION; }

const char* dem_last_error(const dem_context* ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

namespace {
void resolve_scratch(int device);
}

int dem_create(dem_context** out, int device, void* hip_stream) {
  if (!out) return fail(nullptr, DEM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) return fail(nullptr, DEM_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(nullptr, DEM_ERR_NO_DEVICE, "device index out of range");
  auto* ctx = new (std::nothrow) dem_context();
  if (!ctx) return fail(nullptr, DEM_ERR_ALLOC, "host allocation failed");
  ctx->device = device;
  ctx->stream = (hipStream_t)hip_stream;
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return fail(nullptr, DEM_ERR_NO_DEVICE, "hipSetDevice failed");
  }
  // the no-allocation sort's scratch address on this device (its error, if
  // any, is reported by dem_sort_particles_noalloc)
  resolve_scratch(device);
  *out = ctx;
  return DEM_OK;
}

int dem_destroy(dem_context* ctx) {
  if (!ctx) return DEM_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_mat) (void)hipFree(ctx->d_mat);
  if (ctx->d_pair) (void)hipFree(ctx->d_pair);
  if (ctx->d_dpair) (void)hipFree(ctx->d_dpair);
  if (ctx->d_inst) (void)hipFree(ctx->d_inst);
  if (ctx->d_copy_tmp) (void)hipFree(ctx->d_copy_tmp);
  if (ctx->d_control) (void)hipFree(ctx->d_control);
  if (ctx->d_stage_control) (void)hipFree(ctx->d_stage_control);
  if (ctx->d_typevec) (void)hipFree(ctx->d_typevec);
  if (ctx->d_render_types) (void)hipFree(ctx->d_render_types);
  if (ctx->d_owner) (void)hipFree(ctx->d_owner);
  if (ctx->d_scalar) (void)hipFree(ctx->d_scalar);
  free_sort(ctx->sort);
  delete ctx;
  return DEM_OK;
}

int dem_set_stream(dem_context* ctx, void* hip_stream) {
  if (!ctx) return DEM_ERR_INVALID;
  ctx->stream = (hipStream_t)hip_stream;
  ctx->generation++;
  return DEM_OK;
}

int dem_synchronize(dem_context* ctx) {
  if (!ctx) return DEM_ERR_INVALID;
  if (int r = ensure_device(ctx)) return r;
  DEM_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return DEM_OK;
}

static int convert_materials(dem_context* ctx, const dem_particle_props* types, uint32_t num_types,
                      uint32_t num_instances, std::vector<Material>& mat, bool& attract) {
  if (!types) return fail(ctx, DEM_ERR_INVALID, "types is NULL");
  if (num_types == 0 || num_types > DEM_MAX_TYPES) return fail(ctx, DEM_ERR_INVALID, "num_types out of range");
  mat.assign(num_types, Material{});
  attract = false;
  for (uint32_t t = 0; t < num_types; ++t) {
    const dem_particle_props& p = types[t];
    if (p.instance >= num_instances) return fail(ctx, DEM_ERR_INVALID, "a material names a missing instance");
    Material m{};
    m.radius = p.radius; m.mass = p.mass; m.inertia = p.inertia;
    m.kn = p.normal_stiffness; m.kt = p.shear_stiffness; m.cn = p.normal_damping;
    m.bdamp = p.boundary_damping; m.mu = p.friction_coefficient; m.attract = p.attract_coefficient;
    m.areal = p.areal_density; m.inst = p.instance;
    attract = attract || (p.attract_coefficient != 0.0f);
    mat[t] = m;
  }
  return DEM_OK;
}

int dem_set_params(dem_context* ctx, const dem_system_props* sys, const dem_particle_props* types,
                   uint32_t num_types) {
  if (!ctx) return DEM_ERR_INVALID;
  if (!sys) return fail(ctx, DEM_ERR_INVALID, "sys is NULL");
  if (!(sys->time_step > 0.0f) || !isfinite(sys->time_step)) return fail(ctx, DEM_ERR_INVALID, "time_step must be > 0");
  if ((uint64_t)sys->grid_size[0] * sys->grid_size[1] != sys->num_cells)
    return fail(ctx, DEM_ERR_INVALID, "num_cells != grid_size[0]*grid_size[1]");
  if (sys->flags & ~(DEM_GRID_CLAMP | DEM_OPEN_TOP)) return fail(ctx, DEM_ERR_INVALID, "unknown flags");
  std::vector<Material> mat;
  bool attract = false;
  if (int r = convert_materials(ctx, types, num_types, 1u, mat, attract)) return r;
  if (int r = ensure_device(ctx)) return r;
  std::vector<Instance> inst(1, make_instance(sys->cube_dimension, sys->grid_extent, sys->grid_size, 0));
  uint32_t total = 0;
  if (int r = validate_instances(ctx, inst, sys->flags, &total)) return r;
  ctx->h_mat = mat;
  ctx->h_color = colors_of(types, num_types);
  ctx->h_inst = inst;
  ctx->c.dt = sys->time_step;
  ctx->c.gx = sys->gravity[0];
  ctx->c.gy = sys->gravity[1];
  ctx->c.bkn = sys->boundary_normal_stiffness;
  ctx->c.bkt = sys->boundary_shear_stiffness;
  ctx->c.flags = sys->flags;
  ctx->c.attract = attract ? 1u : 0u;
  ctx->total_cells = total;
  ctx->key_bits = bits_for(total);
  if (int r = upload_tables(ctx)) return r;
  ctx->params_set = true;
  return DEM_OK;
}

int dem_set_materials(dem_context* ctx, const dem_particle_props* types, uint32_t num_types) {
  if (int r = check_ctx(ctx)) return r;
  std::vector<Material> mat;
  bool attract = false;
  if (int r = convert_materials(ctx, types, num_types, (uint32_t)ctx->h_inst.size(), mat, attract)) return r;
  if (int r = ensure_device(ctx)) return r;
  ctx->h_mat = mat;
  ctx->h_color = colors_of(types, num_types);
  ctx->c.attract = attract ? 1u : 0u;
  return upload_tables(ctx);
}

int dem_set_instances(dem_context* ctx, const dem_instance* inst, uint32_t num_instances) {
  if (int r = check_ctx(ctx)) return r;
  if (!inst || num_instances == 0) return fail(ctx, DEM_ERR_INVALID, "no instances");
  if (int r = ensure_device(ctx)) return r;
  std::vector<Instance> v(num_instances);
  for (uint32_t k = 0; k < num_instances; ++k)
    v[k] = make_instance(inst[k].cube_dimension, inst[k].grid_extent, inst[k].grid_size, inst[k].cell_base);
  uint32_t total = 0;
  if (int r = validate_instances(ctx, v, ctx->c.flags, &total)) return r;
  for (const Material& m : ctx->h_mat)
    if (m.inst >= num_instances) return fail(ctx, DEM_ERR_INVALID, "a material names a missing instance");
  ctx->h_inst = v;
  ctx->total_cells = total;
  ctx->key_bits = bits_for(total);
  return upload_tables(ctx);
}

int dem_set_gravity(dem_context* ctx, float gx, float gy) {
  if (int r = check_ctx(ctx)) return r;
  ctx->c.gx = gx;
  ctx->c.gy = gy;
  ctx->generation++;
  return DEM_OK;
}

int dem_reserve(dem_context* ctx, uint32_t max_particles) {
  if (!ctx) return DEM_ERR_INVALID;
  if (int r = ensure_device(ctx)) return r;
  return alloc_sort(ctx, ctx->sort, max_particles, true);
}

// ---- stages, reference signatures ------------------------------------------

int dem_calc_hash(dem_context* ctx, const float* pos, uint32_t* gridParticleIndex,
                  uint32_t* gridParticleHash, uint32_t numParticles) {
  return dem_calc_hash_ex(ctx, pos, nullptr, gridParticleIndex, gridParticleHash, numParticles);
}

int dem_calc_hash_ex(dem_context* ctx, const float* pos, const uint32_t* type,
                     uint32_t* gridParticleIndex, uint32_t* gridParticleHash, uint32_t numParticles) {
  if (int r = check_ctx(ctx)) return r;
  if (numParticles == 0) return DEM_OK;
  if (!pos || !gridParticleIndex || !gridParticleHash) return fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (multi(ctx) && !type) return fail(ctx, DEM_ERR_INVALID, "an ensemble needs the type array to hash");
  if (int r = ensure_device(ctx)) return r;
  DEM_HIP(ctx, dem::launch_hash(ctx->c, ctx->stream, ctx->bs_hash, (const float2*)pos, type, gridParticleHash,
                                gridParticleIndex, numParticles, ctx->per_hash),
          "hash kernel");
  return DEM_OK;
}

namespace {

// Sorts with the given scratch and leaves the result in the caller's arrays.
int sort_into_place(dem_context* ctx, uint32_t* keys, uint32_t* vals, uint32_t n, const dem::SortScratch& scratch) {
  // Keys come from hashing, so they are < total_cells; without parameters
  // the full 32 bits are sorted.
  const uint32_t bits = ctx->params_set ? ctx->key_bits : 32u;
  DEM_HIP(ctx, dem::launch_sort_to(ctx->stream, keys, vals, keys, vals, n, bits, scratch), "sort kernels");
  return DEM_OK;
}

// ---- scratch of dem_sort_particles_noalloc ----------------------------------
// A fixed array per GPU: the sort's second key and value arrays
// (kNoAllocKeys words each), its per-tile digit counts (256 per tile of
// kSortTile keys) and its 256 digit totals: 32.5 MiB + 1 KiB, held by every GPU
// the library runs on until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address); it uses the per-device address hipGetSymbolAddress returns.
constexpr uint32_t kNoAllocKeys = DEM_SORT_NOALLOC_MAX;
constexpr size_t kNoAllocCounts = 256 * (((size_t)kNoAllocKeys + dem::kSortTile - 1) / dem::kSortTile);
constexpr size_t kNoAllocWords = 2 * (size_t)kNoAllocKeys + kNoAllocCounts + 256;
static_assert(kNoAllocKeys % 64 == 0, "sub-arrays stay 256-byte aligned");

__device__ __align__(256) uint32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// Once per device, with that device current: the device address of this
// GPU's copy of `scratch`.  Called from dem_create.
void resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e == hipSuccess && dem::sort_block_count_elems(kNoAllocKeys) != kNoAllocCounts) e = hipErrorInvalidValue;
  g_scratch_ptr[device] = e == hipSuccess ? static_cast<uint32_t*>(ptr) : nullptr;
  g_scratch_err[device] = e;
  g_scratch_done[device] = true;
}

}  // namespace

int dem_sort_particles(dem_context* ctx, uint32_t* gridParticleHash, uint32_t* gridParticleIndex,
                       uint32_t numParticles) {
  if (!ctx) return DEM_ERR_INVALID;
  if (numParticles <= 1) return DEM_OK;
  if (!gridParticleHash || !gridParticleIndex) return fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (int r = ensure_device(ctx)) return r;
  if (int r = alloc_sort(ctx, ctx->sort, numParticles, true)) return r;
  return sort_into_place(ctx, gridParticleHash, gridParticleIndex, numParticles, ctx->sort);
}

int dem_sort_particles_noalloc(dem_context* ctx, uint32_t* gridParticleHash, uint32_t* gridParticleIndex,
                               uint32_t numParticles) {
  if (!ctx) return DEM_ERR_INVALID;
  if (numParticles > kNoAllocKeys)
    return fail(ctx, DEM_ERR_INVALID,
                "dem_sort_particles_noalloc sorts at most DEM_SORT_NOALLOC_MAX (4194304) keys; "
                "use dem_sort_particles");
  if (numParticles <= 1) return DEM_OK;
  if (!gridParticleHash || !gridParticleIndex) return fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (int r = ensure_device(ctx)) return r;
  const int dev = ctx->device;
  uint32_t* base = nullptr;
  hipError_t err = hipErrorInvalidDevice;
  if (dev >= 0 && dev < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = g_scratch_ptr[dev];
    err = g_scratch_done[dev] ? g_scratch_err[dev] : hipErrorNotInitialized;
  }
  if (!base) return hip_fail(ctx, err, "scratch address of dem_sort_particles_noalloc");
  dem::SortScratch sc;
  sc.keys_alt = base;
  sc.vals_alt = base + kNoAllocKeys;
  sc.block_counts = base + 2 * (size_t)kNoAllocKeys;
  sc.row_totals = sc.block_counts + kNoAllocCounts;
  sc.capacity = kNoAllocKeys;
  // every scratch word a pass reads was written earlier in that pass (digit
  // counts by the upsweep, totals by the row scan, keys and values by the
  // previous pass's downsweep), so nothing needs clearing between calls
  return sort_into_place(ctx, gridParticleHash, gridParticleIndex, numParticles, sc);
}

int dem_reorder_data_and_find_cell_start(
    dem_context* ctx, uint32_t* cellStart, uint32_t* cellEnd, float* sortedPos, float* sortedVel,
    float* sortedTheta, float* sortedOmega, uint32_t* sortedID, uint32_t* sortedLoc,
    uint32_t* sortedType, const uint32_t* gridParticleHash, const uint32_t* gridParticleIndex,
    const float* oldPos, const float* oldVel, const float* oldTheta, const float* oldOmega,
    const uint32_t* oldID, const uint32_t* oldType, uint32_t numParticles, uint32_t numCells) {
  return dem_reorder_ex(ctx, cellStart, cellEnd, sortedPos, sortedVel, sortedTheta, sortedOmega, sortedID,
                        sortedLoc, sortedType, nullptr, gridParticleHash, gridParticleIndex, oldPos, oldVel,
                        oldTheta, oldOmega, oldID, oldType, nullptr, numParticles, numCells);
}

int dem_reorder_ex(dem_context* ctx, uint32_t* cellStart, uint32_t* cellEnd, float* sortedPos,
                   float* sortedVel, float* sortedTheta, float* sortedOmega, uint32_t* sortedID,
                   uint32_t* sortedLoc, uint32_t* sortedType, float* sortedRadius,
                   const uint32_t* gridParticleHash, const uint32_t* gridParticleIndex,
                   const float* oldPos, const float* oldVel, const float* oldTheta,
                   const float* oldOmega, const uint32_t* oldID, const uint32_t* oldType,
                   const float* oldRadius, uint32_t numParticles, uint32_t numCells) {
  if (!ctx) return DEM_ERR_INVALID;
  if (!cellStart || !cellEnd) return fail(ctx, DEM_ERR_INVALID, "NULL cell array");
  if ((sortedRadius == nullptr) != (oldRadius == nullptr))
    return fail(ctx, DEM_ERR_INVALID, "sortedRadius and oldRadius must both be given or both NULL");
  if (numParticles && (!sortedPos || !sortedVel || !sortedTheta || !sortedOmega || !sortedID || !sortedLoc ||
                       !sortedType || !gridParticleHash || !gridParticleIndex || !oldPos || !oldVel || !oldTheta ||
                       !oldOmega || !oldID || !oldType))
    return fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (int r = ensure_device(ctx)) return r;
  DEM_HIP(ctx, dem::launch_fill(ctx->stream, cellStart, dem::kEmpty, numCells), "cell fill kernel");
  DEM_HIP(ctx, dem::launch_reorder(ctx->stream, ctx->bs_reorder, cellStart, cellEnd, (float2*)sortedPos,
                                   (float2*)sortedVel, sortedTheta, sortedOmega, sortedID, sortedLoc, sortedType,
                                   sortedRadius, gridParticleHash, gridParticleIndex, (const float2*)oldPos,
                                   (const float2*)oldVel, oldTheta, oldOmega, oldID, oldType, oldRadius,
                                   numParticles, numCells, true, ctx->per_reorder),
          "reorder kernel");
  return DEM_OK;
}

int dem_collide(dem_context* ctx, const float* pos, const float* vel, float* acc, const float* omega,
                float* alpha, const uint32_t* type, const uint32_t* cellStart, const uint32_t* cellEnd,
                uint32_t numParticles, uint32_t numCells, float* pressure) {
  return dem_collide_ex(ctx, pos, vel, acc, omega, alpha, type, nullptr, cellStart, cellEnd, numParticles,
                        numCells, pressure);
}

int dem_collide_ex(dem_context* ctx, const float* pos, const float* vel, float* acc, const float* omega,
                   float* alpha, const uint32_t* type, const float* radius, const uint32_t* cellStart,
                   const uint32_t* cellEnd, uint32_t numParticles, uint32_t numCells, float* pressure) {
  if (int r = check_ctx(ctx)) return r;
  if (int r = check_cells(ctx, numCells)) return r;
  if (numParticles == 0) return DEM_OK;
  if (!pos || !vel || !acc || !omega || !alpha || !type || !cellStart || !cellEnd || !pressure)
    return fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (int r = ensure_device(ctx)) return r;
  DEM_HIP(ctx, dem::launch_collide(ctx->c, ctx->stream, ctx->bs_collide, (const float2*)pos, (const float2*)vel,
                                   (float2*)acc, omega, alpha, type, radius, cellStart, cellEnd, pressure,
                                   numParticles, collide_mode(ctx)),
          "collide kernel");
  return DEM_OK;
}

int dem_integrate_system(dem_context* ctx, float* pos, float* vel, const float* acc, float* theta, float* omega,
                         const float* alpha, const uint32_t* type, float* pressure, uint32_t numParticles) {
  return dem_integrate_ex(ctx, pos, vel, acc, theta, omega, alpha, type, nullptr, pressure, numParticles);
}

int dem_integrate_ex(dem_context* ctx, float* pos, float* vel, const float* acc, float* theta, float* omega,
                     const float* alpha, const uint32_t* type, const float* radius, float* pressure,
                     uint32_t numParticles) {
  if (int r = check_ctx(ctx)) return r;
  if (numParticles == 0) return DEM_OK;
  if (!pos || !vel || !acc || !theta || !omega || !alpha || !type || !pressure)
    return fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (int r = ensure_device(ctx)) return r;
  DEM_HIP(ctx, dem::launch_integrate(ctx->c, ctx->stream, ctx->bs_integrate, (float2*)pos, (float2*)vel,
                                     (const float2*)acc, theta, omega, alpha, type, radius, pressure,
                                     numParticles, nullptr, ctx->per_integrate),
          "integrate kernel");
  return DEM_OK;
}

}  // extern "C"

// ---- owned simulation ------------------------------------------------------

namespace dem {

template <class T>
int sim_alloc(dem_sim* m, T** p, size_t count) {
  hipError_t e = hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1));
  if (e != hipSuccess) return hip_fail(m->ctx, e, "hipMalloc(sim state)");
  m->allocs.push_back((void*)*p);
  return DEM_OK;
}

void sim_free(dem_sim* m) {
  for (int k = 0; k < 2; ++k) {
    if (m->exec[k]) (void)hipGraphExecDestroy(m->exec[k]);
    if (m->exec_inc[k]) (void)hipGraphExecDestroy(m->exec_inc[k]);
  }
  if (m->h_count) (void)hipHostFree(m->h_count);
  for (hipEvent_t& e : m->look)
    if (e) (void)hipEventDestroy(e);
  if (m->cap_stream) (void)hipStreamDestroy(m->cap_stream);
  for (void* p : m->allocs) (void)hipFree(p);
  free_sort(m->sort);
  delete m;
}

// Stage sequence of one step (src/main.cu:431-507) on stream s; `ev` (8
// events or NULL) brackets each stage for timing.  The sorted keys go to
// skey[t] and the sorted index to sidx with either sort, so the next step
// can re-sort from them.
namespace {

dem::ShiftState shift_state(dem_sim* m) {
  const int b = m->cur;
  dem::ShiftState st;
  st.pos = m->pos[b];
  st.vel = m->vel[b];
  st.theta = m->theta[b];
  st.omega = m->omega[b];
  st.radius = m->rad ? m->radius[b] : nullptr;
  st.id = m->id[b];
  st.type = m->type[b];
  st.loc = m->loc[b];
  st.skey = m->skey[b];
  st.cell_start = m->cs[b];
  st.cell_end = m->cell_end;
  return st;
}

// A re-sorted step that keeps the state in its buffer: the particles whose
// slot or key changes are moved where they belong, cellStart/cellEnd are
// fixed around them, and collide + integrate run on the sorted state in
// place (integrate computes the next step's keys and mover bits as in the
// other layout).
hipError_t sim_step_inplace(dem_sim* m, hipStream_t s, hipEvent_t* ev, const StepHook* hook) {
  dem_context* ctx = m->ctx;
  const StepConst& c = ctx->c;
  const int b = m->cur;
  const uint32_t n = m->n;
  const dem::ShiftState st = shift_state(m);
  hipError_t e;
  if (ev) (void)hipEventRecord(ev[0], s);
  if (ev) (void)hipEventRecord(ev[1], s);
  e = dem::launch_resort_shift_plan(s, m->key, st, m->shift, m->cells, ctx->key_bits, m->resort, n);
  if (e != hipSuccess) return e;
  if (ev) (void)hipEventRecord(ev[2], s);
  if (ev) (void)hipEventRecord(ev[3], s);
  e = dem::launch_resort_shift_apply(s, st, m->shift, m->cells, n);
  if (e != hipSuccess) return e;
  if (ev) (void)hipEventRecord(ev[4], s);
  ControlOut co;
  const bool control = ctx->h_control.enabled != 0;
  if (control) {
    co.state = ctx->d_control;
    co.react_x = m->react_x;
    co.react_y = m->react_y;
    co.react_m = m->react_m;
  }
  const float* radius = m->rad ? m->radius[b] : nullptr;
  // collide leaves acc and alpha of its quiet runs unwritten; integrate
  // reads the runs' flags
  e = dem::launch_collide(c, s, ctx->bs_collide, m->pos[b], m->vel[b], m->acc, m->omega[b], m->alpha, m->type[b],
                          radius, m->cs[b], m->cell_end, m->pressure, n, collide_mode(ctx), control ? &co : nullptr,
                          nullptr, m->quiet);
  if (e != hipSuccess) return e;
  m->acc_lazy = true;
  if (control) {
    e = dem::launch_control_finish(s, co, m->react_bx, m->react_by, m->react_bm, n, ctx->d_reaction, ctx->d_control,
                                   true);
    if (e != hipSuccess) return e;
  }
  m->sorted_key = m->skey[b];
  if (hook && hook->fn) {
    // the sorted pre-integrate state is buffer b itself
    e = hook->fn(hook->user, m, s, b, b, true);
    if (e != hipSuccess) return e;
  }
  if (ev) (void)hipEventRecord(ev[5], s);
  dem::NextKeys nk;
  nk.key = m->key;
  nk.bits = m->resort.bits;
  nk.tile_count = m->resort.tile_count;
  e = dem::launch_integrate_to(c, s, ctx->bs_integrate, m->pos[b], m->vel[b], m->theta[b], m->omega[b], m->pos[b],
                               m->vel[b], m->theta[b], m->omega[b], m->acc, m->alpha, m->type[b], radius,
                               m->pressure, n, &nk, nullptr, ctx->per_integrate, m->quiet);
  if (e != hipSuccess) return e;
  if (ev) (void)hipEventRecord(ev[6], s);
  m->sorted_key = m->skey[b];
  m->sorted_idx = m->sidx;
  m->cell_start = m->cs[b];
  m->perm_list = true;
  return hipSuccess;
}

}  // namespace

hipError_t sim_step_on(dem_sim* m, hipStream_t s, hipEvent_t* ev, bool resort, const StepHook* hook) {
  dem_context* ctx = m->ctx;
  const StepConst& c = ctx->c;
  // The state is in buffer o and the step writes buffer t.  Two layouts:
  //  * gather (the default): reorder sorts only the positions (into t) and
  //    the per-particle constants; collide and integrate read velocity,
  //    spin and angle from buffer o through the sorted index, and integrate
  //    writes the new state into t (positions in place);
  //  * full (fused collide + integrate, or a hook that records the sorted
  //    pre-integrate state): reorder sorts every field into kScratch, from
  //    which collide + integrate write buffer t.
  if (resort && sim_inplace(m, hook)) return sim_step_inplace(m, s, ev, hook);
  const int o = m->cur, t = o ^ 1;
  const bool gather = !ctx->fuse && !(hook && hook->fn);
  const int S = gather ? t : kScratch;
  const uint32_t n = m->n;
  const uint32_t* type_for_hash = multi(ctx) ? m->type[o] : nullptr;
  hipError_t e;
  if (ev) (void)hipEventRecord(ev[0], s);
  if (resort) {
    // the keys, mover bits and counts were written by the previous step's
    // collide + integrate (the hash stage is fused into it)
    if (ev) (void)hipEventRecord(ev[1], s);
    // the placement kernel also clears this step's cellStart (cs[t]) while
    // its movers look up the previous step's (cs[o])
    e = dem::launch_resort(s, m->key, m->skey[o], m->cs[o], m->cell_end, m->cells, true, ctx->key_bits,
                           m->skey[t], m->sidx, m->resort, n, m->cs[t], m->cells);
    if (e != hipSuccess) return e;
  } else {
    e = dem::launch_hash(c, s, ctx->bs_hash, m->pos[o], type_for_hash, m->key, m->idx, n, ctx->per_hash);
    if (e != hipSuccess) return e;
    // this step's collide + integrate accumulates the next step's mover counts
    const uint32_t tiles = (n + dem::kResortTile - 1u) / dem::kResortTile;
    e = dem::launch_fill(s, m->resort.tile_count, 0u, tiles);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[1], s);
    // the last pass writes the sorted keys and index where the step keeps them
    e = dem::launch_sort_to(s, m->key, m->idx, m->skey[t], m->sidx, n, ctx->key_bits, m->sort);
    if (e != hipSuccess) return e;
  }
  m->sorted_key = m->skey[t];
  m->sorted_idx = m->sidx;
  if (ev) (void)hipEventRecord(ev[2], s);
  if (!resort) {
    e = dem::launch_fill(s, m->cs[t], dem::kEmpty, m->cells);
    if (e != hipSuccess) return e;
  }
  if (ev) (void)hipEventRecord(ev[3], s);
  e = dem::launch_reorder(s, ctx->bs_reorder, m->cs[t], m->cell_end, m->pos[S], gather ? nullptr : m->vel[S],
                          gather ? nullptr : m->theta[S], gather ? nullptr : m->omega[S], m->id[t], m->loc[t],
                          m->type[t], m->rad ? m->radius[t] : nullptr, m->sorted_key, m->sorted_idx, m->pos[o],
                          m->vel[o], m->theta[o], m->omega[o], m->id[o], m->type[o],
                          m->rad ? m->radius[o] : nullptr, n, m->cells, !gather, ctx->per_reorder);
  if (e != hipSuccess) return e;
  if (ev) (void)hipEventRecord(ev[4], s);
  dem::NextKeys nk;
  nk.key = m->key;
  nk.bits = m->resort.bits;
  nk.tile_count = m->resort.tile_count;
  const float* S_radius = m->rad ? m->radius[t] : nullptr;
  // the controlled particle: pair term in collide, then its reaction and
  // move (src/main.cu:466-514); the move does not feed this step's integrate
  ControlOut co;
  const bool control = ctx->h_control.enabled != 0;
  if (control) {
    co.state = ctx->d_control;
    co.react_x = m->react_x;
    co.react_y = m->react_y;
    co.react_m = m->react_m;
  }
  if (!gather && ctx->fuse) {
    e = dem::launch_collide_integrate(c, s, ctx->bs_collide, m->pos[S], m->vel[S], m->omega[S], m->theta[S],
                                      m->type[t], S_radius, m->cs[t], m->cell_end, m->acc, m->alpha, m->pressure,
                                      m->pos[t], m->vel[t], m->omega[t], m->theta[t], nk, n, ctx->cmode,
                                      control ? &co : nullptr);
    if (e != hipSuccess) return e;
    if (control) {
      e = dem::launch_control_finish(s, co, m->react_bx, m->react_by, m->react_bm, n, ctx->d_reaction, ctx->d_control,
                                     true);
      if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[5], s);
  } else {
    // with gather, velocity and spin come from buffer o through the index
    const int D = gather ? o : S;
    const uint32_t* gat = gather ? m->sorted_idx : nullptr;
    e = dem::launch_collide(c, s, ctx->bs_collide, m->pos[S], m->vel[D], m->acc, m->omega[D], m->alpha, m->type[t],
                            S_radius, m->cs[t], m->cell_end, m->pressure, n, collide_mode(ctx),
                            control ? &co : nullptr, gat);
    if (e != hipSuccess) return e;
    if (control) {
      e = dem::launch_control_finish(s, co, m->react_bx, m->react_by, m->react_bm, n, ctx->d_reaction, ctx->d_control,
                                     true);
      if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[5], s);
    // integrate into buffer t (with gather the positions in place, and
    // velocity, angle and spin through the index from buffer o)
    e = dem::launch_integrate_to(c, s, ctx->bs_integrate, m->pos[S], m->vel[D], m->theta[D], m->omega[D], m->pos[t],
                                 m->vel[t], m->theta[t], m->omega[t], m->acc, m->alpha, m->type[t], S_radius,
                                 m->pressure, n, &nk, gat, ctx->per_integrate);
    if (e != hipSuccess) return e;
  }
  if (hook && hook->fn) {
    // the full layout: the sorted pre-integrate state is still in kScratch / t
    e = hook->fn(hook->user, m, s, kScratch, t, false);
    if (e != hipSuccess) return e;
  }
  if (ev) (void)hipEventRecord(ev[6], s);
  m->cur = t;
  m->cell_start = m->cs[t];
  m->perm_list = false;
  m->acc_lazy = false;  // these layouts write every acc and alpha
  return hipSuccess;
}

bool sim_inplace(const dem_sim* m, const StepHook*) { return !m->ctx->copy && !m->ctx->fuse; }

namespace {
// acc = alpha = +0 in the runs of 64 slots that collide marked quiet.
__global__ void k_quiet_zero(const uint32_t* __restrict__ quiet, float2* __restrict__ acc, float* __restrict__ alpha,
                             uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || quiet[i / 64u] == 0u) return;
  acc[i] = make_float2(0.0f, 0.0f);
  alpha[i] = 0.0f;
}
}  // namespace

hipError_t sim_acc_complete(dem_sim* m, hipStream_t s) {
  if (!m->acc_lazy || m->n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_quiet_zero, dim3((m->n + 255u) / 256u), dim3(256), 0, s, m->quiet, m->acc, m->alpha, m->n);
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) m->acc_lazy = false;
  return e;
}

// In place, the re-sort stays cheaper than the radix sort while it lists at
// most n / kListedFrac particles.
constexpr uint32_t kListedFrac = 4;

bool sim_use_resort(dem_sim* m) {
  if (!m->resort_ok || m->sort_mode == 1) return false;
  if (m->sort_mode == 2) return true;
  // The radix sort and full reorder cost about the same every step.  The
  // re-sort's movers step is one workgroup, fast up to a few thousand
  // movers; in place, every particle it lists (its slot changed: each mover
  // shifts the particles between its old and new slot) has its state staged
  // and written back.  Past either limit the steps radix sort, and the
  // re-sort is tried again after 1, 2, 4 ... 64 radix-sorted steps (its
  // counts are only measured by a re-sort), or after 8, 16 ... 64 when the
  // counts were at twice a limit or more.  The counts are the latest the
  // device has written, which may trail the steps already enqueued: graph
  // replays are queued faster than they run, so a trial's counts reach the
  // host only after the choices that follow it are made, and a trial that
  // lists most of the state costs about two radix-sorted steps (the collapse
  // of a dense pack lists 2-3M of 4.2M discs a step for a thousand steps).
  const volatile uint32_t* h = m->h_count;
  const uint32_t movers = h[0], listed = h[1];
  const uint32_t mover_limit = std::max(4096u, m->n / 64u), listed_limit = m->n / kListedFrac;
  const bool inplace = sim_inplace(m);
  const bool cheap = movers <= mover_limit && (!inplace || listed <= listed_limit);
  if (cheap) {
    m->radix_run = m->radix_wait = 0;
    return true;
  }
  if (m->radix_run < m->radix_wait) {
    ++m->radix_run;
    return false;
  }
  m->radix_run = 0;
  if (m->radix_wait == 0) {
    const bool far = movers / 2u >= mover_limit || (inplace && listed / 2u >= listed_limit);
    m->radix_wait = far ? 8u : 1u;
    if (far) return false;
  } else {
    m->radix_wait = std::min(2u * m->radix_wait, 64u);
  }
  return true;
}

// The sim's cell arrays were sized at creation; a later parameter change
// must not make hashing produce keys beyond them.
int sim_check(dem_sim* m) {
  if (m->ctx->total_cells > m->cells)
    return fail(m->ctx, DEM_ERR_STATE, "the grid grew beyond the sim's cell arrays; create a new sim");
  return ensure_device(m->ctx);
}

int sim_capture(dem_sim* m) {
  dem_context* ctx = m->ctx;
  if (!m->cap_stream) DEM_HIP(ctx, hipStreamCreateWithFlags(&m->cap_stream, hipStreamNonBlocking), "hipStreamCreate");
  for (int k = 0; k < 2; ++k) {
    if (m->exec[k]) (void)hipGraphExecDestroy(m->exec[k]);
    if (m->exec_inc[k]) (void)hipGraphExecDestroy(m->exec_inc[k]);
    m->exec[k] = m->exec_inc[k] = nullptr;
  }
  const int saved = m->cur;
  uint32_t* const saved_key = m->sorted_key;
  uint32_t* const saved_idx = m->sorted_idx;
  const bool saved_list = m->perm_list;
  for (int variant = 0; variant < 2; ++variant) {
    for (int start = 0; start < 2; ++start) {
      hipGraph_t g = nullptr;
      DEM_HIP(ctx, hipStreamBeginCapture(m->cap_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      m->cur = start;
      hipError_t e1 = sim_step_on(m, m->cap_stream, nullptr, variant == 1);
      hipError_t e2 = e1 == hipSuccess ? sim_step_on(m, m->cap_stream, nullptr, variant == 1) : e1;
      hipError_t e3 = hipStreamEndCapture(m->cap_stream, &g);
      m->cur = saved;
      m->cell_start = m->cs[saved];
      m->sorted_key = saved_key;
      m->sorted_idx = saved_idx;
      m->perm_list = saved_list;
      if (e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(ctx, e2, "capturing the step");
      }
      if (e3 != hipSuccess) return hip_fail(ctx, e3, "hipStreamEndCapture");
      hipGraphExec_t* slot = variant == 1 ? &m->exec_inc[start] : &m->exec[start];
      hipError_t e4 = hipGraphInstantiate(slot, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e4 != hipSuccess) return hip_fail(ctx, e4, "hipGraphInstantiate");
    }
  }
  m->graph_gen = ctx->generation;
  return DEM_OK;
}

}  // namespace dem

extern "C" {

int dem_sim_create(dem_context* ctx, uint32_t n, uint32_t num_cells, int with_radius, dem_sim** out) {
  if (!out) return fail(ctx, DEM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (int r = check_ctx(ctx)) return r;
  if (n == 0) return fail(ctx, DEM_ERR_INVALID, "n must be > 0");
  if (int r = check_cells(ctx, num_cells)) return r;
  if (int r = ensure_device(ctx)) return r;
  auto* m = new (std::nothrow) dem_sim();
  if (!m) return fail(ctx, DEM_ERR_ALLOC, "host allocation failed");
  m->ctx = ctx;
  m->n = n;
  m->cells = num_cells;
  m->rad = with_radius != 0;
  int r = DEM_OK;
  for (int b = 0; b < 2 && r == DEM_OK; ++b) {
    if (!r) r = sim_alloc(m, &m->pos[b], n);
    if (!r) r = sim_alloc(m, &m->vel[b], n);
    if (!r) r = sim_alloc(m, &m->theta[b], n);
    if (!r) r = sim_alloc(m, &m->omega[b], n);
    if (!r) r = sim_alloc(m, &m->id[b], n);
    if (!r) r = sim_alloc(m, &m->type[b], n);
    if (!r) r = sim_alloc(m, &m->loc[b], n);
    if (!r && m->rad) r = sim_alloc(m, &m->radius[b], n);
  }
  if (!r) r = sim_alloc(m, &m->pos[kScratch], n);
  if (!r) r = sim_alloc(m, &m->vel[kScratch], n);
  if (!r) r = sim_alloc(m, &m->theta[kScratch], n);
  if (!r) r = sim_alloc(m, &m->omega[kScratch], n);
  if (!r) r = sim_alloc(m, &m->acc, n);
  if (!r) r = sim_alloc(m, &m->alpha, n);
  if (!r) r = sim_alloc(m, &m->quiet, (n + 63u) / 64u);
  if (!r) r = sim_alloc(m, &m->pressure, n);
  if (!r) r = sim_alloc(m, &m->react_x, n);
  if (!r) r = sim_alloc(m, &m->react_y, n);
  if (!r) r = sim_alloc(m, &m->react_m, n);
  const uint32_t react_blocks = (n + dem::kReactBlock - 1u) / dem::kReactBlock;
  if (!r) r = sim_alloc(m, &m->react_bx, react_blocks);
  if (!r) r = sim_alloc(m, &m->react_by, react_blocks);
  if (!r) r = sim_alloc(m, &m->react_bm, react_blocks);
  if (!r) r = sim_alloc(m, &m->key, n);
  if (!r) r = sim_alloc(m, &m->idx, n);
  if (!r) r = sim_alloc(m, &m->cs[0], num_cells);
  if (!r) r = sim_alloc(m, &m->cs[1], num_cells);
  if (!r) r = sim_alloc(m, &m->cell_end, num_cells);
  if (!r) r = alloc_sort(ctx, m->sort, n, false);
  const uint32_t words = (n + 31u) / 32u, tiles = (n + dem::kResortTile - 1u) / dem::kResortTile;
  for (int b = 0; b < 2 && r == DEM_OK; ++b) {
    if (!r) r = sim_alloc(m, &m->skey[b], n);
    if (!r) r = sim_alloc(m, &m->resort.mkey[b], n);
    if (!r) r = sim_alloc(m, &m->resort.mpos[b], n);
  }
  if (!r) r = sim_alloc(m, &m->sidx, n);
  if (!r) r = sim_alloc(m, &m->resort.bits, words);
  if (!r) r = sim_alloc(m, &m->resort.word_prefix, words);
  if (!r) r = sim_alloc(m, &m->resort.tile_count, tiles);
  if (!r) r = sim_alloc(m, &m->resort.tile_off, tiles + 1);
  if (!r) r = sim_alloc(m, &m->resort.tile_list, tiles);
  if (!r) r = sim_alloc(m, &m->resort.info, 2);
  // in-place re-sort: list, staging (dynamic fields in kScratch) and counts
  if (!r) r = sim_alloc(m, &m->shift.key, n);
  if (!r) r = sim_alloc(m, &m->shift.id, n);
  if (!r) r = sim_alloc(m, &m->shift.type, n);
  if (!r) r = sim_alloc(m, &m->shift.dst, n);
  if (!r) r = sim_alloc(m, &m->shift.src, n);
  if (!r) r = sim_alloc(m, &m->shift.emptied, n);
  if (!r && m->rad) r = sim_alloc(m, &m->shift.radius, n);
  if (!r) r = sim_alloc(m, &m->shift.count, 4);
  if (!r) r = sim_alloc(m, &m->shift.tag, n);
  m->shift.pos = m->pos[kScratch];
  m->shift.vel = m->vel[kScratch];
  m->shift.theta = m->theta[kScratch];
  m->shift.omega = m->omega[kScratch];
  m->shift.capacity = n;
  if (!r) {
    if (hipHostMalloc((void**)&m->h_count, 2 * sizeof(uint32_t), hipHostMallocMapped) != hipSuccess) {
      m->h_count = nullptr;
      r = fail(ctx, DEM_ERR_ALLOC, "hipHostMalloc(re-sort counts)");
    } else {
      m->h_count[0] = m->h_count[1] = 0;
      void* dp = nullptr;
      if (hipHostGetDevicePointer(&dp, m->h_count, 0) != hipSuccess) r = fail(ctx, DEM_ERR_ALLOC, "hipHostGetDevicePointer");
      m->resort.host_count = (uint32_t*)dp;
      m->shift.host_listed = dp ? (uint32_t*)dp + 1 : nullptr;
    }
  }
  m->resort.capacity = n;
  if (r) {
    sim_free(m);
    return r;
  }
  // cellEnd starts at 0 as in allocateVectors (src/functions.cu:140).
  hipError_t e = dem::launch_fill(ctx->stream, m->cell_end, 0u, num_cells);
  if (e == hipSuccess) e = dem::launch_fill(ctx->stream, m->shift.count, 0u, 4);
  if (e == hipSuccess) e = dem::launch_fill(ctx->stream, (uint32_t*)m->shift.tag, 0u, 2 * n);
  if (e == hipSuccess) e = dem::launch_fill(ctx->stream, m->cs[0], dem::kEmpty, num_cells);
  if (e == hipSuccess) e = dem::launch_fill(ctx->stream, m->cs[1], dem::kEmpty, num_cells);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    sim_free(m);
    return hip_fail(ctx, e, "sim init");
  }
  m->sorted_key = m->key;
  m->cell_start = m->cs[m->cur];
  m->sorted_idx = m->idx;
  *out = m;
  return DEM_OK;
}

int dem_sim_destroy(dem_sim* m) {
  if (!m) return DEM_ERR_INVALID;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  sim_free(m);
  return DEM_OK;
}

int dem_sim_upload(dem_sim* m, const float* pos, const float* vel, const float* theta, const float* omega,
                   const uint32_t* id, const uint32_t* type, const uint32_t* loc, const float* radius) {
  if (!m) return DEM_ERR_INVALID;
  dem_context* ctx = m->ctx;
  if (!pos || !vel || !theta || !omega || !id || !type || !loc || (m->rad && !radius))
    return fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (int r = ensure_device(ctx)) return r;
  const int c = m->cur;
  const size_t n = m->n;
  hipStream_t s = ctx->stream;
  DEM_HIP(ctx, hipMemcpyAsync(m->pos[c], pos, 8 * n, hipMemcpyDefault, s), "upload");
  DEM_HIP(ctx, hipMemcpyAsync(m->vel[c], vel, 8 * n, hipMemcpyDefault, s), "upload");
  DEM_HIP(ctx, hipMemcpyAsync(m->theta[c], theta, 4 * n, hipMemcpyDefault, s), "upload");
  DEM_HIP(ctx, hipMemcpyAsync(m->omega[c], omega, 4 * n, hipMemcpyDefault, s), "upload");
  DEM_HIP(ctx, hipMemcpyAsync(m->id[c], id, 4 * n, hipMemcpyDefault, s), "upload");
  DEM_HIP(ctx, hipMemcpyAsync(m->type[c], type, 4 * n, hipMemcpyDefault, s), "upload");
  DEM_HIP(ctx, hipMemcpyAsync(m->loc[c], loc, 4 * n, hipMemcpyDefault, s), "upload");
  if (m->rad) DEM_HIP(ctx, hipMemcpyAsync(m->radius[c], radius, 4 * n, hipMemcpyDefault, s), "upload");
  DEM_HIP(ctx, hipStreamSynchronize(s), "upload sync");
  m->resort_ok = false;  // an uploaded state is in no known sorted order
  m->perm_list = false;
  m->acc_lazy = false;
  m->h_count[0] = m->h_count[1] = 0;
  m->radix_run = m->radix_wait = 0;
  return DEM_OK;
}

int dem_sim_download(dem_sim* m, float* pos, float* vel, float* theta, float* omega, uint32_t* id,
                     uint32_t* type, uint32_t* loc, float* radius, float* acc, float* alpha, float* pressure,
                     uint32_t* hash, uint32_t* index, uint32_t* cellStart, uint32_t* cellEnd) {
  if (!m) return DEM_ERR_INVALID;
  dem_context* ctx = m->ctx;
  if (int r = ensure_device(ctx)) return r;
  const int c = m->cur;
  const size_t n = m->n;
  hipStream_t s = ctx->stream;
  if (acc || alpha) DEM_HIP(ctx, sim_acc_complete(m, s), "accelerations");
  if (index && m->perm_list) {
    // the last step re-sorted in place: its index is the identity except
    // at the slots it listed
    DEM_HIP(ctx, dem::launch_shift_index(s, m->shift, m->sidx, m->n), "index");
    m->sorted_idx = m->sidx;
  }
  struct Item { void* dst; const void* src; size_t bytes; } items[] = {
      {pos, m->pos[c], 8 * n},      {vel, m->vel[c], 8 * n},       {theta, m->theta[c], 4 * n},
      {omega, m->omega[c], 4 * n},  {id, m->id[c], 4 * n},             {type, m->type[c], 4 * n},
      {loc, m->loc[c], 4 * n},          {radius, m->rad ? m->radius[c] : nullptr, 4 * n},
      {acc, m->acc, 8 * n},             {alpha, m->alpha, 4 * n},          {pressure, m->pressure, 4 * n},
      {hash, m->sorted_key, 4 * n},     {index, m->sorted_idx, 4 * n},
      {cellStart, m->cell_start, 4 * (size_t)m->cells}, {cellEnd, m->cell_end, 4 * (size_t)m->cells}};
  for (const Item& it : items)
    if (it.dst && it.src) DEM_HIP(ctx, hipMemcpyAsync(it.dst, it.src, it.bytes, hipMemcpyDeviceToHost, s), "download");
  DEM_HIP(ctx, hipStreamSynchronize(s), "download sync");
  return DEM_OK;
}

// One eager step with the sort method chosen for it.
int sim_step_auto(dem_sim* m) {
  const bool resort = sim_use_resort(m);
  DEM_HIP(m->ctx, sim_step_on(m, m->ctx->stream, nullptr, resort), "step");
  m->resort_ok = true;
  return DEM_OK;
}

// The queue-bounding events only mark progress: no system-scope fence (a
// cache writeback at every record; the counts the host reads live in
// host-mapped memory, and a late count only delays a choice that gives the
// same bits either way).
constexpr unsigned kLookFlags = hipEventDisableTiming | hipEventDisableSystemFence;
// Below this many particles a step takes tens of microseconds, and the
// events and waits cost more (about 9% on 2,000-60,000-disc columns) than
// a late choice does; the queue is bounded only for larger states.
constexpr uint32_t kLookMinParticles = 1u << 18;

int dem_sim_step(dem_sim* m, uint32_t steps) {
  if (!m) return DEM_ERR_INVALID;
  if (int r = sim_check(m)) return r;
  // as dem_sim_step_graph: under the automatic sort choice the steps are
  // queued in groups, at most two ahead, so the counts it reads stay recent
  constexpr uint32_t G = dem_sim::kLookGroup;
  const bool look = m->sort_mode == 0 && m->n >= kLookMinParticles && steps > 2u * G;
  if (look)
    for (hipEvent_t& e : m->look)
      if (!e) DEM_HIP(m->ctx, hipEventCreateWithFlags(&e, kLookFlags), "hipEventCreate");
  for (uint32_t k = 0; k < steps; ++k) {
    hipEvent_t& slot = m->look[(k / G) & 1u];
    if (look && k % G == 0 && k >= 2u * G) DEM_HIP(m->ctx, hipEventSynchronize(slot), "hipEventSynchronize");
    if (int r = sim_step_auto(m)) return r;
    if (look && k % G == G - 1u) DEM_HIP(m->ctx, hipEventRecord(slot, m->ctx->stream), "hipEventRecord");
  }
  return DEM_OK;
}

int dem_sim_step_graph(dem_sim* m, uint32_t steps) {
  if (!m) return DEM_ERR_INVALID;
  dem_context* ctx = m->ctx;
  if (int r = sim_check(m)) return r;
  if (steps & 1u) {
    if (int r = sim_step_auto(m)) return r;
    --steps;
  }
  if (steps == 0) return DEM_OK;
  if (!m->exec[0] || m->graph_gen != ctx->generation)
    if (int r = sim_capture(m)) return r;
  // The automatic sort choice reads counts that the replays write: with
  // every replay queued at once they would all be chosen from the counts of
  // before the call.  It queues the replays in groups of kLookGroup instead,
  // waiting before each group for the group two back to finish (the GPU
  // keeps a group queued meanwhile); one event a group.
  constexpr uint32_t G = dem_sim::kLookGroup;
  const bool look = m->sort_mode == 0 && m->n >= kLookMinParticles;
  if (look)
    for (hipEvent_t& e : m->look)
      if (!e) DEM_HIP(ctx, hipEventCreateWithFlags(&e, kLookFlags), "hipEventCreate");
  for (uint32_t k = 0, r = 0; k < steps; k += 2, ++r) {
    hipEvent_t& slot = m->look[(r / G) & 1u];
    if (look && r % G == 0 && r >= 2u * G) DEM_HIP(ctx, hipEventSynchronize(slot), "hipEventSynchronize");
    const bool resort = sim_use_resort(m);
    DEM_HIP(ctx, hipGraphLaunch(resort ? m->exec_inc[m->cur] : m->exec[m->cur], ctx->stream), "hipGraphLaunch");
    if (look && r % G == G - 1u) DEM_HIP(ctx, hipEventRecord(slot, ctx->stream), "hipEventRecord");
    m->resort_ok = true;
    m->perm_list = resort && sim_inplace(m);
    m->acc_lazy = m->perm_list;
  }
  // Each replay runs two steps, which leaves the state in the same buffer;
  // the sorted keys are in skey[cur] and the sorted index in sidx.
  m->sorted_key = m->skey[m->cur];
  m->sorted_idx = m->sidx;
  return DEM_OK;
}

int dem_set_step_variant(dem_context* ctx, int flags) {
  if (!ctx) return DEM_ERR_INVALID;
  if (flags & ~(DEM_STEP_FUSED | DEM_STEP_PREFETCH | DEM_STEP_BATCH | DEM_STEP_COPY | DEM_STEP_WIDE | DEM_STEP_LIST |
                DEM_STEP_PER_MASK))
    return fail(ctx, DEM_ERR_INVALID, "unknown step variant flags");
  const int lg_hash = (flags >> 8) & 3, lg_reorder = (flags >> 10) & 3, lg_integrate = (flags >> 12) & 3;
  if (lg_hash == 3 || lg_reorder == 3 || lg_integrate == 3)
    return fail(ctx, DEM_ERR_INVALID, "particles a lane must be 1, 2 or 4");
  const int modes = !!(flags & DEM_STEP_PREFETCH) + !!(flags & DEM_STEP_BATCH) + !!(flags & DEM_STEP_LIST);
  if (modes > 1)
    return fail(ctx, DEM_ERR_INVALID, "DEM_STEP_PREFETCH, DEM_STEP_BATCH and DEM_STEP_LIST exclude each other");
  ctx->fuse = (flags & DEM_STEP_FUSED) != 0;
  ctx->cmode = (flags & DEM_STEP_PREFETCH) ? 1 : ((flags & DEM_STEP_BATCH) ? 2 : ((flags & DEM_STEP_LIST) ? 3 : 0));
  ctx->copy = (flags & DEM_STEP_COPY) != 0;
  ctx->wide = (flags & DEM_STEP_WIDE) != 0;
  ctx->per_hash = 1 << lg_hash;
  ctx->per_reorder = 1 << lg_reorder;
  ctx->per_integrate = 1 << lg_integrate;
  ctx->generation++;  // captured graphs follow the choice
  return DEM_OK;
}

int dem_sim_set_sort_mode(dem_sim* m, int mode) {
  if (!m) return DEM_ERR_INVALID;
  if (mode < 0 || mode > 2) return fail(m->ctx, DEM_ERR_INVALID, "sort mode must be 0, 1 or 2");
  m->sort_mode = mode;
  return DEM_OK;
}

int dem_sim_mover_count(dem_sim* m, uint32_t* count) {
  if (!m || !count) return m ? fail(m->ctx, DEM_ERR_INVALID, "count is NULL") : DEM_ERR_INVALID;
  *count = *(volatile const uint32_t*)m->h_count;
  return DEM_OK;
}

int dem_sim_resort_counts(dem_sim* m, uint32_t* movers, uint32_t* listed) {
  if (!m || !movers || !listed) return m ? fail(m->ctx, DEM_ERR_INVALID, "NULL count") : DEM_ERR_INVALID;
  const volatile uint32_t* h = m->h_count;
  *movers = h[0];
  *listed = h[1];
  return DEM_OK;
}

int dem_sim_stage_times(dem_sim* m, uint32_t reps, float* ms) {
  if (!m || !ms || reps == 0) return m ? fail(m->ctx, DEM_ERR_INVALID, "bad arguments") : DEM_ERR_INVALID;
  dem_context* ctx = m->ctx;
  if (int r = sim_check(m)) return r;
  hipEvent_t ev[7];
  for (int k = 0; k < 7; ++k) DEM_HIP(ctx, hipEventCreate(&ev[k]), "hipEventCreate");
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  int rc = DEM_OK;
  for (uint32_t r = 0; r < reps && rc == DEM_OK; ++r) {
    hipError_t e = sim_step_on(m, ctx->stream, ev, sim_use_resort(m));
    if (e == hipSuccess) m->resort_ok = true;
    if (e == hipSuccess) e = hipEventSynchronize(ev[6]);
    if (e != hipSuccess) {
      rc = hip_fail(ctx, e, "timed step");
      break;
    }
    for (int k = 0; k < 6; ++k) {
      float t = 0.0f;
      (void)hipEventElapsedTime(&t, ev[k], ev[k + 1]);
      acc[k] += t;
    }
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, ev[0], ev[6]);
    acc[6] += t;
  }
  for (int k = 0; k < 7; ++k) {
    (void)hipEventDestroy(ev[k]);
    ms[k] = (float)(acc[k] / reps);
  }
  return rc;
}

namespace {
__global__ void k_get_by_id(const uint32_t* __restrict__ id, const float2* __restrict__ pos,
                            const float2* __restrict__ vel, float2* __restrict__ pos_out, float2* __restrict__ vel_out,
                            uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = id[i];
  if (d >= n) return;
  if (pos_out) pos_out[d] = pos[i];
  if (vel_out) vel_out[d] = vel[i];
}

__global__ void k_set_by_id(const uint32_t* __restrict__ id, const float2* __restrict__ pos_in,
                            const float2* __restrict__ vel_in, float2* __restrict__ pos, float2* __restrict__ vel,
                            uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = id[i];
  if (d >= n) return;
  if (pos_in) pos[i] = pos_in[d];
  if (vel_in) vel[i] = vel_in[d];
}
}  // namespace

int dem_sim_get_by_id(dem_sim* m, float* pos_by_id, float* vel_by_id) {
  if (!m) return DEM_ERR_INVALID;
  dem_context* ctx = m->ctx;
  if (int r = ensure_device(ctx)) return r;
  const int c = m->cur;
  hipLaunchKernelGGL(k_get_by_id, dim3((m->n + 255) / 256), dim3(256), 0, ctx->stream, m->id[c], m->pos[c],
                     m->vel[c],
                     (float2*)pos_by_id, (float2*)vel_by_id, m->n);
  DEM_HIP(ctx, hipGetLastError(), "get by id");
  return DEM_OK;
}

int dem_sim_set_by_id(dem_sim* m, const float* pos_by_id, const float* vel_by_id) {
  if (!m) return DEM_ERR_INVALID;
  dem_context* ctx = m->ctx;
  if (int r = ensure_device(ctx)) return r;
  const int c = m->cur;
  hipLaunchKernelGGL(k_set_by_id, dim3((m->n + 255) / 256), dim3(256), 0, ctx->stream, m->id[c],
                     (const float2*)pos_by_id, (const float2*)vel_by_id, m->pos[m->cur], m->vel[m->cur], m->n);
  DEM_HIP(ctx, hipGetLastError(), "set by id");
  // the order still follows skey[cur], but new positions may move many
  // particles to other cells: the next step sorts from scratch
  if (pos_by_id) m->resort_ok = false;
  return DEM_OK;
}

namespace {
// One output row per followed particle, 12 words: id, type, pos.x, pos.y,
// vel.x, vel.y, acc.x, acc.y, theta, omega, alpha, pressure (floats as their
// bit patterns).  A followed ID outside [0, n) gives a row of 0xFFFFFFFF.
__global__ void k_gather_rows(const uint32_t* __restrict__ follow, uint32_t count, const uint32_t* __restrict__ loc,
                              const float2* __restrict__ pos, const float2* __restrict__ vel,
                              const float2* __restrict__ acc, const float* __restrict__ theta,
                              const float* __restrict__ omega, const float* __restrict__ alpha,
                              const uint32_t* __restrict__ id, const uint32_t* __restrict__ type,
                              const float* __restrict__ pressure, uint32_t n, uint32_t* __restrict__ rows) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= count) return;
  uint32_t* r = rows + (size_t)kRowWords * f;
  const uint32_t d = follow[f];
  const uint32_t i = d < n ? loc[d] : n;
  if (i >= n) {
    for (int k = 0; k < kRowWords; ++k) r[k] = 0xFFFFFFFFu;
    return;
  }
  const float2 p = pos[i], v = vel[i], a = acc[i];
  r[0] = id[i];
  r[1] = type[i];
  r[2] = __float_as_uint(p.x);
  r[3] = __float_as_uint(p.y);
  r[4] = __float_as_uint(v.x);
  r[5] = __float_as_uint(v.y);
  r[6] = __float_as_uint(a.x);
  r[7] = __float_as_uint(a.y);
  r[8] = __float_as_uint(theta[i]);
  r[9] = __float_as_uint(omega[i]);
  r[10] = __float_as_uint(alpha[i]);
  r[11] = __float_as_uint(pressure[i]);
}
}  // namespace

int dem_gather_rows(dem_context* ctx, const uint32_t* follow_ids, uint32_t count, const uint32_t* loc,
                    const float* pos, const float* vel, const float* acc, const float* theta, const float* omega,
                    const float* alpha, const uint32_t* id, const uint32_t* type, const float* pressure,
                    uint32_t numParticles, uint32_t* rows) {
  if (int r = check_ctx(ctx)) return r;
  if (count == 0) return DEM_OK;
  if (!follow_ids || !loc || !pos || !vel || !acc || !theta || !omega || !alpha || !id || !type || !pressure || !rows)
    return fail(ctx, DEM_ERR_INVALID, "dem_gather_rows: null pointer");
  if (int r = ensure_device(ctx)) return r;
  hipLaunchKernelGGL(k_gather_rows, dim3((count + 63) / 64), dim3(64), 0, ctx->stream, follow_ids, count, loc,
                     (const float2*)pos, (const float2*)vel, (const float2*)acc, theta, omega, alpha, id, type,
                     pressure, numParticles, rows);
  DEM_HIP(ctx, hipGetLastError(), "gather rows");
  return DEM_OK;
}

int dem_sim_gather_rows(dem_sim* m, const uint32_t* follow_ids, uint32_t count, uint32_t* rows) {
  if (!m) return DEM_ERR_INVALID;
  if (int r = ensure_device(m->ctx)) return r;
  DEM_HIP(m->ctx, sim_acc_complete(m, m->ctx->stream), "accelerations");
  const int c = m->cur;
  return dem_gather_rows(m->ctx, follow_ids, count, m->loc[c], (const float*)m->pos[c],
                         (const float*)m->vel[c], (const float*)m->acc, m->theta[c], m->omega[c], m->alpha,
                         m->id[c], m->type[c], m->pressure,
                         m->n, rows);
}

int dem_set_controlled(dem_context* ctx, const dem_controlled* state) {
  if (int r = check_ctx(ctx)) return r;
  if (!state) return fail(ctx, DEM_ERR_INVALID, "state is NULL");
  if (state->enabled) {
    if (multi(ctx)) return fail(ctx, DEM_ERR_INVALID, "the controlled particle needs a single instance");
    if (state->type >= ctx->c.num_types) return fail(ctx, DEM_ERR_INVALID, "controlled particle type out of range");
  }
  if (int r = ensure_device(ctx)) return r;
  if (!ctx->d_control) {
    // the state, then the 3-float reaction, in one allocation
    DEM_HIP(ctx, hipMalloc((void**)&ctx->d_control, sizeof(dem_controlled) + 4 * sizeof(float)),
            "hipMalloc(controlled)");
    ctx->d_reaction = (float*)(ctx->d_control + 1);
  }
  const float zero[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  DEM_HIP(ctx, hipMemcpyAsync(ctx->d_control, state, sizeof(dem_controlled), hipMemcpyHostToDevice, ctx->stream),
          "controlled upload");
  DEM_HIP(ctx, hipMemcpyAsync(ctx->d_reaction, zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream),
          "controlled upload");
  DEM_HIP(ctx, hipStreamSynchronize(ctx->stream), "controlled upload");
  if ((ctx->h_control.enabled != 0) != (state->enabled != 0)) ++ctx->generation;  // step kernels change
  ctx->h_control = *state;
  return DEM_OK;
}

int dem_get_controlled(dem_context* ctx, dem_controlled* state, float reaction[3]) {
  if (int r = check_ctx(ctx)) return r;
  if (int r = ensure_device(ctx)) return r;
  DEM_HIP(ctx, hipStreamSynchronize(ctx->stream), "controlled download");
  if (!ctx->d_control) {
    if (state) *state = ctx->h_control;
    if (reaction) reaction[0] = reaction[1] = reaction[2] = 0.0f;
    return DEM_OK;
  }
  if (state) DEM_HIP(ctx, hipMemcpy(state, ctx->d_control, sizeof(dem_controlled), hipMemcpyDeviceToHost),
                     "controlled download");
  if (reaction) DEM_HIP(ctx, hipMemcpy(reaction, ctx->d_reaction, 3 * sizeof(float), hipMemcpyDeviceToHost),
                        "controlled download");
  return DEM_OK;
}

int dem_collide_controlled(dem_context* ctx, const float* pos, const float* vel, float* acc, const float* omega,
                           float* alpha, const uint32_t* type, const uint32_t* cellStart, const uint32_t* cellEnd,
                           uint32_t numParticles, uint32_t numCells, const float control_state[6],
                           uint32_t control_type, float* forceVecX, float* blockX, float* forceVecY, float* blockY,
                           float* momentVec, float* blockM, float* ctrlF, float* ctrlM, float* pressure) {
  if (int r = check_ctx(ctx)) return r;
  if (int r = check_cells(ctx, numCells)) return r;
  if (!control_state || !ctrlF || !ctrlM) return fail(ctx, DEM_ERR_INVALID, "NULL control argument");
  if (multi(ctx)) return fail(ctx, DEM_ERR_INVALID, "the controlled particle needs a single instance");
  if (control_type >= ctx->c.num_types) return fail(ctx, DEM_ERR_INVALID, "control_type out of range");
  if (numParticles && (!pos || !vel || !acc || !omega || !alpha || !type || !cellStart || !cellEnd || !pressure ||
                       !forceVecX || !forceVecY || !momentVec || !blockX || !blockY || !blockM))
    return fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (int r = ensure_device(ctx)) return r;
  if (!ctx->d_stage_control) {
    DEM_HIP(ctx, hipMalloc((void**)&ctx->d_stage_control, sizeof(dem_controlled) + 4 * sizeof(float)),
            "hipMalloc(controlled)");
    ctx->d_stage_reaction = (float*)(ctx->d_stage_control + 1);
  }
  dem_controlled st{};
  st.pos[0] = control_state[0];
  st.pos[1] = control_state[1];
  st.vel[0] = control_state[2];
  st.vel[1] = control_state[3];
  st.theta = control_state[4];
  st.omega = control_state[5];
  st.type = control_type;
  st.enabled = 1;
  DEM_HIP(ctx, hipMemcpyAsync(ctx->d_stage_control, &st, sizeof(st), hipMemcpyHostToDevice, ctx->stream),
          "controlled upload");
  ControlOut co;
  co.state = ctx->d_stage_control;
  co.react_x = forceVecX;
  co.react_y = forceVecY;
  co.react_m = momentVec;
  DEM_HIP(ctx, dem::launch_collide(ctx->c, ctx->stream, ctx->bs_collide, (const float2*)pos, (const float2*)vel,
                                   (float2*)acc, omega, alpha, type, nullptr, cellStart, cellEnd, pressure,
                                   numParticles, collide_mode(ctx) & kWideMode, &co),
          "collide kernel");
  DEM_HIP(ctx, dem::launch_control_finish(ctx->stream, co, blockX, blockY, blockM, numParticles,
                                          ctx->d_stage_reaction, ctx->d_stage_control, false),
          "reaction kernel");
  float r[3];
  DEM_HIP(ctx, hipMemcpyAsync(r, ctx->d_stage_reaction, sizeof(r), hipMemcpyDeviceToHost, ctx->stream),
          "reaction download");
  DEM_HIP(ctx, hipStreamSynchronize(ctx->stream), "reaction download");
  ctrlF[0] = r[0];
  ctrlF[1] = r[1];
  ctrlM[0] = r[2];
  return DEM_OK;
}

int dem_set_contact_params(dem_context* ctx, const double* p, uint32_t num_types) {
  if (int r = check_ctx(ctx)) return r;
  if (!p || num_types != ctx->h_mat.size())
    return fail(ctx, DEM_ERR_INVALID, "need one (kn, kt, cn, mu) row per material");
  if (int r = ensure_device(ctx)) return r;
  for (uint32_t t = 0; t < num_types; ++t) {
    ctx->h_mat[t].kn = (float)p[4 * t + 0];
    ctx->h_mat[t].kt = (float)p[4 * t + 1];
    ctx->h_mat[t].cn = (float)p[4 * t + 2];
    ctx->h_mat[t].mu = (float)p[4 * t + 3];
  }
  return upload_tables(ctx);
}

int dem_sim_sort_passes(dem_sim* m) {
  if (!m) return DEM_ERR_INVALID;
  return dem::sort_passes(m->ctx->key_bits);
}

int dem_sim_info(dem_sim* m, uint32_t* n, int* device) {
  if (!m) return DEM_ERR_INVALID;
  if (n) *n = m->n;
  if (device) *device = m->ctx->device;
  return DEM_OK;
}

int dem_sim_autotune(dem_sim* m, uint32_t* block) {
  if (!m) return DEM_ERR_INVALID;
  dem_context* ctx = m->ctx;
  if (int r = sim_check(m)) return r;
  hipStream_t s = ctx->stream;
  const StepConst& c = ctx->c;
  const int o = m->cur, t = o ^ 1;
  const uint32_t n = m->n;
  // Populate the scratch buffer t (the next step's sorted state) once; the
  // current state in buffer o is never written.
  // (the sorted keys and index go to the re-sort's second mover arrays, n
  // words each, which every re-sort rewrites before reading; the re-sort is
  // off after tuning, and the last step's sorted keys and index stay as
  // they are)
  DEM_HIP(ctx, dem::launch_hash(c, s, ctx->bs_hash, m->pos[o], multi(ctx) ? m->type[o] : nullptr, m->key, m->idx, n), "tune");
  uint32_t* sk = m->resort.mkey[1];
  uint32_t* si = m->resort.mpos[1];
  DEM_HIP(ctx, dem::launch_sort_to(s, m->key, m->idx, sk, si, n, ctx->key_bits, m->sort), "tune");
  DEM_HIP(ctx, dem::launch_fill(s, m->cell_start, dem::kEmpty, m->cells), "tune");
  // the fully sorted copy into kScratch (the fused layout); the gather
  // layout's reorder writes the sorted positions there too
  const int K = kScratch;
  int tper = 1;  // particles a lane under test (hash, reorder, integrate)
  auto reorder_as = [&](int bs, bool full) {
    return dem::launch_reorder(s, bs, m->cell_start, m->cell_end, m->pos[K], full ? m->vel[K] : nullptr,
                               full ? m->theta[K] : nullptr, full ? m->omega[K] : nullptr, m->id[t], m->loc[t],
                               m->type[t], m->rad ? m->radius[t] : nullptr, sk, si, m->pos[o], m->vel[o],
                               m->theta[o], m->omega[o], m->id[o], m->type[o], m->rad ? m->radius[o] : nullptr, n,
                               m->cells, full, tper);
  };
  DEM_HIP(ctx, reorder_as(ctx->bs_reorder, true), "tune");
  hipEvent_t e0, e1;
  DEM_HIP(ctx, hipEventCreate(&e0), "hipEventCreate");
  DEM_HIP(ctx, hipEventCreate(&e1), "hipEventCreate");
  const int cand[] = {64, 128, 256};
  auto time_it = [&](auto&& launch, int bs, float* out) -> hipError_t {
    hipError_t e = launch(bs);  // warm
    if (e != hipSuccess) return e;
    (void)hipEventRecord(e0, s);
    for (int k = 0; k < 5 && e == hipSuccess; ++k) e = launch(bs);
    (void)hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(out, e0, e1);
    return e;
  };
  // scratch outputs for the integrated state, so the live state is never written
  float2 *tpos = nullptr, *tvel = nullptr;
  float *ttheta = nullptr, *tomega = nullptr;
  hipError_t e = hipMalloc((void**)&tpos, 8 * (size_t)n);
  if (e == hipSuccess) e = hipMalloc((void**)&tvel, 8 * (size_t)n);
  if (e == hipSuccess) e = hipMalloc((void**)&ttheta, 4 * (size_t)n);
  if (e == hipSuccess) e = hipMalloc((void**)&tomega, 4 * (size_t)n);
  dem::NextKeys nk;
  nk.key = m->key;  // the sort's input, free after it
  nk.bits = m->resort.bits;
  nk.tile_count = m->resort.tile_count;
  const float2* S_pos = m->pos[K];
  const float2* S_vel = m->vel[K];
  const float* S_theta = m->theta[K];
  const float* S_omega = m->omega[K];
  const float* S_radius = m->rad ? m->radius[t] : nullptr;
  int pf = 0;  // candidate mode under test: 0 plain, 1 prefetch, 2 batch
  // the split layout as its re-sorted steps run it: in place, on the fully
  // sorted state; else velocity, spin and angle from the live state
  // (buffer o) through the sorted index
  const bool inplace = !ctx->copy;
  const float2* D_vel = inplace ? S_vel : m->vel[o];
  const float* D_omega = inplace ? S_omega : m->omega[o];
  const float* D_theta = inplace ? S_theta : m->theta[o];
  const uint32_t* gat = inplace ? nullptr : si;
  auto collide = [&](int b) {
    return dem::launch_collide(c, s, b, S_pos, D_vel, m->acc, D_omega, m->alpha, m->type[t], S_radius,
                               m->cell_start, m->cell_end, m->pressure, n, pf | (ctx->wide ? kWideMode : 0),
                               nullptr, gat);
  };
  auto integrate = [&](int b) {
    return dem::launch_integrate_to(c, s, b, S_pos, D_vel, D_theta, D_omega, tpos, tvel, ttheta, tomega, m->acc,
                                    m->alpha, m->type[t], S_radius, m->pressure, n, &nk, gat, tper);
  };
  auto fused = [&](int b) {
    return dem::launch_collide_integrate(c, s, b, S_pos, S_vel, S_omega, S_theta, m->type[t], S_radius,
                                         m->cell_start, m->cell_end, m->acc, m->alpha, m->pressure, tpos, tvel,
                                         tomega, ttheta, nk, n, pf);
  };
  int best[4] = {ctx->bs_hash, ctx->bs_reorder, ctx->bs_collide, ctx->bs_integrate};
  int best_per[4] = {ctx->per_hash, ctx->per_reorder, 1, ctx->per_integrate};
  float best_t[4] = {1e30f, 1e30f, 1e30f, 1e30f};
  float t_full = 1e30f;  // the fused layout's reorder (every field) at the best block size
  int bs_full = ctx->bs_reorder, per_full = ctx->per_reorder;
  // collide and fused block sizes for each candidate mode
  constexpr int kModes = 4;
  int best_c[kModes], best_f[kModes];
  float t_c[kModes], t_f[kModes];
  for (int p = 0; p < kModes; ++p) {
    best_c[p] = best_f[p] = ctx->bs_collide;
    t_c[p] = t_f[p] = 1e30f;
  }
  for (int bs : cand) {
    if (e != hipSuccess) break;
    float tm = 0;
    // the streaming kernels at 1, 2 and 4 particles a lane
    for (int per : {1, 2, 4}) {
      tper = per;
      e = time_it([&](int b) { return dem::launch_hash(c, s, b, m->pos[o], multi(ctx) ? m->type[o] : nullptr, m->key, m->idx, n, tper); }, bs, &tm);
      if (e != hipSuccess) break;
      if (tm < best_t[0]) { best_t[0] = tm; best[0] = bs; best_per[0] = per; }
      e = time_it([&](int b) { return reorder_as(b, true); }, bs, &tm);
      if (e != hipSuccess) break;
      if (tm < t_full) { t_full = tm; bs_full = bs; per_full = per; }
      e = time_it([&](int b) { return reorder_as(b, false); }, bs, &tm);
      if (e != hipSuccess) break;
      if (tm < best_t[1]) { best_t[1] = tm; best[1] = bs; best_per[1] = per; }
      e = time_it(integrate, bs, &tm);
      if (e != hipSuccess) break;
      if (tm < best_t[3]) { best_t[3] = tm; best[3] = bs; best_per[3] = per; }
    }
    if (e != hipSuccess) break;
    tper = best_per[3];  // the fused kernel and collide run one particle a lane
    for (int p = 0; p < kModes && e == hipSuccess; ++p) {
      pf = p;
      e = time_it(collide, bs, &tm);
      if (e != hipSuccess) break;
      if (tm < t_c[p]) { t_c[p] = tm; best_c[p] = bs; }
      if (p == 2) continue;  // the fused kernel has no batched mode
      e = time_it(fused, bs, &tm);
      if (e != hipSuccess) break;
      if (tm < t_f[p]) { t_f[p] = tm; best_f[p] = bs; }
    }
  }
  // the in-place layout's apply launches with nothing listed (their cost
  // when few particles change slot); the last in-place step's index is
  // materialized first, since its list is cleared for this
  float t_apply = 0.0f;
  if (e == hipSuccess && inplace) {
    if (m->perm_list) e = dem::launch_shift_index(s, m->shift, m->sidx, n);
    m->perm_list = false;
    if (e == hipSuccess) e = dem::launch_fill(s, m->shift.count, 0u, 2);
    const dem::ShiftState st = shift_state(m);
    if (e == hipSuccess)
      e = time_it([&](int) { return dem::launch_resort_shift_apply(s, st, m->shift, m->cells, n); }, 0, &t_apply);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (tpos) (void)hipFree(tpos);
  if (tvel) (void)hipFree(tvel);
  if (ttheta) (void)hipFree(ttheta);
  if (tomega) (void)hipFree(tomega);
  m->resort_ok = false;  // the cell arrays and keys now hold the tuning step's data
  m->acc_lazy = false;   // and acc, alpha the tuning collide's (every slot written)
  if (e != hipSuccess) return hip_fail(ctx, e, "autotune");
  // collide + integrate as one kernel or two, in whichever candidate mode is
  // fastest here (the split layout's reorder moves fewer bytes than the
  // fused layout's)
  bool fuse = false;
  int mode = 0;
  float tbest = 1e30f;
  const float t_move = inplace ? t_apply : best_t[1];  // the split layout's reorder
  for (int p = 0; p < kModes; ++p) {
    if (t_move + t_c[p] + best_t[3] < tbest) { tbest = t_move + t_c[p] + best_t[3]; fuse = false; mode = p; }
    if (p != 2 && t_full + t_f[p] < tbest) { tbest = t_full + t_f[p]; fuse = true; mode = p; }
  }
  ctx->bs_hash = best[0];
  ctx->bs_reorder = fuse ? bs_full : best[1];
  ctx->bs_collide = fuse ? best_f[mode] : best_c[mode];
  ctx->bs_integrate = best[3];
  ctx->per_hash = best_per[0];
  ctx->per_reorder = fuse ? per_full : best_per[1];
  ctx->per_integrate = best_per[3];
  ctx->fuse = fuse;
  ctx->cmode = mode;
  ctx->generation++;
  if (block) {
    block[0] = (uint32_t)best[0];
    block[1] = (uint32_t)best[1];
    block[2] = (uint32_t)ctx->bs_collide;
    block[3] = (uint32_t)best[3];
    auto lg = [](int per) { return per == 4 ? 2u : (per == 2 ? 1u : 0u); };
    block[4] = (fuse ? 1u : 0u) | (mode == 1 ? 2u : 0u) | (mode == 2 ? 4u : 0u) | (ctx->copy ? 8u : 0u) |
               (mode == 3 ? (uint32_t)DEM_STEP_LIST : 0u) | (lg(ctx->per_hash) << 8) |
               (lg(ctx->per_reorder) << 10) | (lg(ctx->per_integrate) << 12);
  }
  return DEM_OK;
}

}  // extern "C"
