// scene.hip -- DemSo's scene generators on the device: createRectangles,
// createTriangles and createSingleParticles (src/functions.cu:198-311,
// kernels src/particles_kernel.cuh:43-111), one thread per particle.
//
// The rectangle jitter is the reference's formula with its curand_normal
// draws (time(NULL)-seeded, so irreproducible) replaced by the normal draws
// of scene_rng.h: draw p of streams `stream` (x) and `stream + 1` (y) under
// `seed`.  Everything else is the reference's FP32 arithmetic, so a layout
// built here equals the host driver's and demso_amd/scene.py's bit for bit.
#include "capi_internal.h"
#include "scene_rng.h"

#include <vector>

namespace {

__global__ __launch_bounds__(256) void k_create_rectangles(float2* __restrict__ pos, uint32_t* __restrict__ id,
                                                           uint32_t* __restrict__ loc, uint32_t* __restrict__ type,
                                                           float2 start, float2 side_length, uint2 side,
                                                           uint32_t start_id, uint32_t num_types,
                                                           const uint32_t* __restrict__ type_vec, uint64_t seed,
                                                           uint64_t stream, uint32_t n) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t x = p % side.x, y = p / side.x;
  const float nx = (float)dem_rng::jitter_normal(seed, stream, p);
  const float ny = (float)dem_rng::jitter_normal(seed, stream + 1u, p);
  float2 q;
  q.x = start.x + side_length.x / (float)(side.x - 1u) * ((float)x + (nx - 0.5f) / 100.0f);
  q.y = start.y + side_length.y / (float)(side.y - 1u) * ((float)y + (ny - 0.5f) / 100.0f);
  pos[p] = q;
  id[p] = start_id + p;
  loc[p] = start_id + p;
  type[p] = type_vec[(x + y) % num_types];
}

__global__ __launch_bounds__(256) void k_create_triangles(float2* __restrict__ pos, uint32_t* __restrict__ id,
                                                          uint32_t* __restrict__ loc, uint32_t* __restrict__ type,
                                                          float2 start, uint32_t N, uint32_t num_types,
                                                          const uint32_t* __restrict__ type_vec, float space,
                                                          float height, uint32_t start_id, uint32_t n) {
  const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
  if (index >= n) return;
  // row L from the top of the triangle's index order, column C in the row
  const uint32_t radicand = N * N + N - 2u * (index + 1u) + 1u;
  const int L1 = (int)((1.0f + 2.0f * sqrtf((float)radicand)) / 2.0f);
  const int L = (int)N - L1;
  const int s = (int)(((uint32_t)L1 + N + N * N - (uint32_t)(L1 * L1)) / 2u);
  const int C = (int)index - s + L1;
  const int A = (int)N - 1;
  float2 q;
  q.x = start.x + (float)(L + 2 * C - A) * space / 2.0f;
  q.y = start.y + (float)L * height;
  pos[index] = q;
  id[index] = start_id + index;
  loc[index] = start_id + index;
  type[index] = type_vec[(uint32_t)(C + L) % num_types];
}

__global__ __launch_bounds__(256) void k_ids(uint32_t* __restrict__ id, uint32_t* __restrict__ loc, uint32_t start_id,
                                             uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  id[i] = start_id + i;
  loc[i] = start_id + i;
}

// The particle type vector in device memory (the reference allocates and
// frees one per call); the calls synchronize before returning, so one
// buffer per context is enough.
int upload_type_vec(dem_context* ctx, const uint32_t* host, uint32_t count, uint32_t** out) {
  if (ctx->typevec_cap < count) {
    if (ctx->d_typevec) (void)hipFree(ctx->d_typevec);
    ctx->d_typevec = nullptr;
    ctx->typevec_cap = 0;
    DEM_HIP(ctx, hipMalloc((void**)&ctx->d_typevec, sizeof(uint32_t) * count), "hipMalloc(type vector)");
    ctx->typevec_cap = count;
  }
  DEM_HIP(ctx, hipMemcpyAsync(ctx->d_typevec, host, sizeof(uint32_t) * count, hipMemcpyHostToDevice, ctx->stream),
          "type vector upload");
  *out = ctx->d_typevec;
  return DEM_OK;
}

int check_types(dem_context* ctx, uint32_t count, const uint32_t* vec) {
  if (count == 0 || !vec) return dem::fail(ctx, DEM_ERR_INVALID, "empty particle type vector");
  if (ctx->params_set)
    for (uint32_t k = 0; k < count; ++k)
      if (vec[k] >= ctx->c.num_types) return dem::fail(ctx, DEM_ERR_INVALID, "particle type out of range");
  return DEM_OK;
}

}  // namespace

extern "C" {

int dem_create_rectangles_ex(dem_context* ctx, float* pos, uint32_t* ID, uint32_t* loc, uint32_t* type,
                             const float start[2], const float sideLength[2], const uint32_t side[2],
                             uint32_t startID, uint32_t numParticleTypes, const uint32_t* particleTypeVec,
                             uint64_t seed, uint64_t stream) {
  if (!ctx) return DEM_ERR_INVALID;  // parameters are not needed here
  if (!start || !sideLength || !side) return dem::fail(ctx, DEM_ERR_INVALID, "NULL rectangle geometry");
  const uint64_t n64 = (uint64_t)side[0] * side[1];
  if (n64 == 0) return DEM_OK;
  if (n64 > 0xFFFFFFFFull - startID) return dem::fail(ctx, DEM_ERR_INVALID, "rectangle too large");
  if (!pos || !ID || !loc || !type) return dem::fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (int r = check_types(ctx, numParticleTypes, particleTypeVec)) return r;
  if (int r = dem::ensure_device(ctx)) return r;
  uint32_t* tv = nullptr;
  if (int r = upload_type_vec(ctx, particleTypeVec, numParticleTypes, &tv)) return r;
  const uint32_t n = (uint32_t)n64;
  hipLaunchKernelGGL(k_create_rectangles, dim3((n + 255u) / 256u), dim3(256), 0, ctx->stream, (float2*)pos, ID, loc,
                     type, make_float2(start[0], start[1]), make_float2(sideLength[0], sideLength[1]),
                     make_uint2(side[0], side[1]), startID, numParticleTypes, tv, seed, stream, n);
  DEM_HIP(ctx, hipGetLastError(), "rectangle kernel");
  DEM_HIP(ctx, hipStreamSynchronize(ctx->stream), "rectangle kernel");
  return DEM_OK;
}

int dem_create_rectangles(dem_context* ctx, float* pos, uint32_t* ID, uint32_t* loc, uint32_t* type,
                          const float start[2], const float sideLength[2], const uint32_t side[2], uint32_t startID,
                          uint32_t numParticleTypes, const uint32_t* particleTypeVec, uint64_t seed) {
  return dem_create_rectangles_ex(ctx, pos, ID, loc, type, start, sideLength, side, startID, numParticleTypes,
                                  particleTypeVec, seed, DEM_RECT_STREAM);
}

int dem_create_triangles(dem_context* ctx, float* pos, uint32_t* ID, uint32_t* loc, uint32_t* type,
                         const float start[2], uint32_t N, uint32_t numParticleTypes, const uint32_t* particleTypeVec,
                         float space, float height, uint32_t startID, uint32_t numParticles) {
  if (!ctx) return DEM_ERR_INVALID;  // parameters are not needed here
  if (numParticles == 0) return DEM_OK;
  if (!start) return dem::fail(ctx, DEM_ERR_INVALID, "NULL triangle start");
  // N rows hold N(N+1)/2 particles; beyond that the row formula has no meaning
  if (N == 0 || (uint64_t)N * (N + 1ull) / 2ull < numParticles || N > 65535u)
    return dem::fail(ctx, DEM_ERR_INVALID, "numParticles exceeds the N-row triangle");
  if ((uint64_t)numParticles > 0xFFFFFFFFull - startID) return dem::fail(ctx, DEM_ERR_INVALID, "IDs overflow");
  if (!pos || !ID || !loc || !type) return dem::fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (int r = check_types(ctx, numParticleTypes, particleTypeVec)) return r;
  if (int r = dem::ensure_device(ctx)) return r;
  uint32_t* tv = nullptr;
  if (int r = upload_type_vec(ctx, particleTypeVec, numParticleTypes, &tv)) return r;
  hipLaunchKernelGGL(k_create_triangles, dim3((numParticles + 255u) / 256u), dim3(256), 0, ctx->stream, (float2*)pos,
                     ID, loc, type, make_float2(start[0], start[1]), N, numParticleTypes, tv, space, height, startID,
                     numParticles);
  DEM_HIP(ctx, hipGetLastError(), "triangle kernel");
  DEM_HIP(ctx, hipStreamSynchronize(ctx->stream), "triangle kernel");
  return DEM_OK;
}

int dem_create_single_particles(dem_context* ctx, float* pos, float* vel, float* theta, float* omega, uint32_t* ID,
                                uint32_t* loc, uint32_t* type, const float* usrPos, const float* usrVel,
                                const float* usrTheta, const float* usrOmega, const uint32_t* usrType,
                                uint32_t numParticles, uint32_t startID) {
  if (!ctx) return DEM_ERR_INVALID;  // parameters are not needed here
  if (numParticles == 0) return DEM_OK;
  if (!pos || !vel || !theta || !omega || !ID || !loc || !type || !usrPos || !usrVel || !usrTheta || !usrOmega ||
      !usrType)
    return dem::fail(ctx, DEM_ERR_INVALID, "NULL array");
  if ((uint64_t)numParticles > 0xFFFFFFFFull - startID) return dem::fail(ctx, DEM_ERR_INVALID, "IDs overflow");
  if (ctx->params_set)
    for (uint32_t k = 0; k < numParticles; ++k)
      if (usrType[k] >= ctx->c.num_types) return dem::fail(ctx, DEM_ERR_INVALID, "particle type out of range");
  if (int r = dem::ensure_device(ctx)) return r;
  hipStream_t s = ctx->stream;
  const size_t n = numParticles;
  DEM_HIP(ctx, hipMemcpyAsync(pos, usrPos, 8 * n, hipMemcpyDefault, s), "single particles");
  DEM_HIP(ctx, hipMemcpyAsync(vel, usrVel, 8 * n, hipMemcpyDefault, s), "single particles");
  DEM_HIP(ctx, hipMemcpyAsync(theta, usrTheta, 4 * n, hipMemcpyDefault, s), "single particles");
  DEM_HIP(ctx, hipMemcpyAsync(omega, usrOmega, 4 * n, hipMemcpyDefault, s), "single particles");
  DEM_HIP(ctx, hipMemcpyAsync(type, usrType, 4 * n, hipMemcpyDefault, s), "single particles");
  hipLaunchKernelGGL(k_ids, dim3((numParticles + 255u) / 256u), dim3(256), 0, s, ID, loc, startID, numParticles);
  DEM_HIP(ctx, hipGetLastError(), "id kernel");
  DEM_HIP(ctx, hipStreamSynchronize(s), "single particles");
  return DEM_OK;
}

}  // extern "C"
