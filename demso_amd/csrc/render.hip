// render.hip -- DemSo's headless image output: plotParticles
// (src/functions.cu:557-600) with plotSpheresD and plotControlParticleD
// (src/particles_kernel.cuh:629-775), and updatePressureScale
// (src/functions.cu:684-693).
//
// The reference paints every disc from its own thread, so where discs
// overlap the pixel keeps whichever write landed last.  Here the result is
// fixed: a pixel covered by several discs takes the one with the highest
// index (the last in a sequential loop over the particles), found with one
// atomicMax per covered pixel into an owner map before the discs are
// painted.  Rotation lines are drawn over all discs, the controlled particle
// over everything.  The pixel arithmetic is the reference's (float pixel
// scale, truncation, unsigned wrap and the clamp to the last pixel);
// conversions it leaves undefined (negative or NaN float to integer, colour
// values outside a byte) are pinned by px_int and to_byte, and the rotation
// line's cos/sin come from the correctly rounded series of scene_rng.h, so a
// sequential host restatement gives the same bytes.
#include "capi_internal.h"
#include "scene_rng.h"

#include <string.h>

#include <algorithm>
#include <vector>

namespace dem {
namespace {

struct RenderType {
  float radius, r, g, b;
};

struct RenderConst {
  int32_t dimx, dimy;        // image size in pixels
  int32_t cdimx, cdimy;      // the controlled particle's pixel square
  float scale_x, scale_y;    // imageDIM / cubeDimension (pixels per metre)
  uint32_t npix;             // dimx * dimy
  uint32_t rotations, by_pressure;
  float max_pressure;
  uint32_t num_types;
};

__host__ __device__ inline int px_int(float v) {
  if (!(v == v)) return 0;
  if (v >= 2147483520.0f) return 2147483647;
  if (v <= -2147483648.0f) return (-2147483647 - 1);
  return (int)v;
}

__host__ __device__ inline uint8_t to_byte(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 255.0f) return 255;
  return (uint8_t)(int)v;
}

// colorbar (src/particles_kernel.cuh:629-648)
__device__ inline float colorbar(float x) {
  const float top = 1.5f, bottom = -0.5f;
  float rv;
  if (x < 0.5f) rv = x / 0.5f * (top - bottom) + bottom;
  else rv = (x - 0.5f) / 0.5f * (bottom - top) + top;
  if (rv > 1.0f) return 255.0f;
  if (rv < 0.0f) return 0.0f;
  return rv * 255.0f;
}

__device__ inline uint32_t clamp_pixel(const RenderConst& rc, uint32_t pixel) {
  return pixel >= rc.npix ? rc.npix - 1u : pixel;
}

// Discs wider than 2 kMaxHalf pixels are drawn as that wide (the reference
// would loop over every pixel of an arbitrarily large square).
constexpr int kMaxHalf = 8192;

struct Disc {
  int cx, cy, h;
  float pr;
};

__device__ inline Disc disc_of(const RenderConst& rc, const RenderType* __restrict__ tt, float2 p, uint32_t t,
                               const float* __restrict__ radius, uint32_t i) {
  const uint32_t tc = t < rc.num_types ? t : rc.num_types - 1u;
  Disc d;
  const float r = radius ? radius[i] : tt[tc].radius;
  d.pr = rc.scale_y * r;
  d.cx = px_int(rc.scale_x * p.x);
  d.cy = px_int(rc.scale_y * p.y);
  const float c = ceilf(d.pr);
  d.h = c > 0.0f ? min(px_int(c), kMaxHalf) : 0;
  return d;
}

__global__ __launch_bounds__(256) void k_plot_clear(uchar4* __restrict__ img, uint32_t* __restrict__ owner,
                                                    uint8_t bg, uint32_t npix) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  img[i] = make_uchar4(bg, bg, bg, bg);
  owner[i] = 0u;
}

// Pass 1: every covered pixel records the highest particle index + 1.
__global__ __launch_bounds__(256) void k_plot_claim(RenderConst rc, const RenderType* __restrict__ tt,
                                                    const float2* __restrict__ pos, const uint32_t* __restrict__ type,
                                                    const float* __restrict__ radius, uint32_t* __restrict__ owner,
                                                    uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Disc d = disc_of(rc, tt, pos[i], type[i], radius, i);
  const float pr2 = d.pr * d.pr;
  for (int x = -d.h; x < d.h; ++x)
    for (int y = -d.h; y < d.h; ++y)
      if ((float)(x * x + y * y) < pr2) {
        const uint32_t gx = (uint32_t)(d.cx + x), gy = (uint32_t)(d.cy + y);
        const uint32_t pixel = clamp_pixel(rc, gx + gy * (uint32_t)rc.dimx);
        atomicMax(&owner[pixel], i + 1u);
      }
}

// Pass 2: the owner paints (plotSpheresD's colours and shading).
__global__ __launch_bounds__(256) void k_plot_paint(RenderConst rc, const RenderType* __restrict__ tt,
                                                    const float2* __restrict__ pos, const uint32_t* __restrict__ type,
                                                    const float* __restrict__ radius,
                                                    const float* __restrict__ pressure,
                                                    const uint32_t* __restrict__ owner, uchar4* __restrict__ img,
                                                    uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t t = type[i];
  const Disc d = disc_of(rc, tt, pos[i], t, radius, i);
  const uint32_t tc = t < rc.num_types ? t : rc.num_types - 1u;
  float cr = tt[tc].r, cg = tt[tc].g, cb = tt[tc].b;
  if (rc.by_pressure) {
    const float pn = pressure[i] / rc.max_pressure;
    cr = colorbar(pn - 0.25f);
    cg = colorbar(pn);
    cb = colorbar(pn + 0.25f);
  }
  const float pr2 = d.pr * d.pr;
  for (int x = -d.h; x < d.h; ++x)
    for (int y = -d.h; y < d.h; ++y)
      if ((float)(x * x + y * y) < pr2) {
        const uint32_t gx = (uint32_t)(d.cx + x), gy = (uint32_t)(d.cy + y);
        const uint32_t pixel = clamp_pixel(rc, gx + gy * (uint32_t)rc.dimx);
        if (owner[pixel] != i + 1u) continue;
        const float fscale = sqrtf((pr2 - (float)(x * x) - (float)(y * y)) / pr2);
        img[pixel] = make_uchar4(to_byte(cr * fscale), to_byte(cg * fscale), to_byte(cb * fscale),
                                 to_byte(255.0f * fscale));
      }
}

// Pass 3: a black line from the centre along theta, one pixel per radius
// step (plotSpheresD's viewRotations branch); every line pixel gets the same
// value, so the order of the writes does not matter.
__global__ __launch_bounds__(256) void k_plot_lines(RenderConst rc, const RenderType* __restrict__ tt,
                                                    const float2* __restrict__ pos, const float* __restrict__ theta,
                                                    const uint32_t* __restrict__ type,
                                                    const float* __restrict__ radius, uchar4* __restrict__ img,
                                                    uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Disc d = disc_of(rc, tt, pos[i], type[i], radius, i);
  double u = (double)theta[i] / 6.28318530717958647692;
  u = u - floor(u);
  if (!(u < 1.0)) u = 0.0;  // 1 - tiny rounding, inf, NaN
  double us = u - 0.25;
  if (us < 0.0) us = us + 1.0;
  const float c = (float)dem_rng::series_cos2pi(u);
  const float s = (float)dem_rng::series_cos2pi(us);
  for (int r = 0; (float)r < d.pr && r < kMaxHalf; ++r) {
    const uint32_t gx = (uint32_t)px_int((float)d.cx + (float)r * c);
    const uint32_t gy = (uint32_t)px_int((float)d.cy + (float)r * s);
    const uint32_t pixel = clamp_pixel(rc, gx + gy * (uint32_t)rc.dimx);
    img[pixel] = make_uchar4(0, 0, 0, 0);
  }
}

// plotControlParticleD: a tx x ty thread square (the reference's 16-wide
// blocks rounded up) centred on the controlled particle.
__global__ __launch_bounds__(256) void k_plot_control(RenderConst rc, const RenderType* __restrict__ tt,
                                                      const dem_controlled* __restrict__ ctl, uint32_t tx, uint32_t ty,
                                                      uchar4* __restrict__ img) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= tx * ty) return;
  const int x = (int)(k % tx), y = (int)(k / tx);
  const int hx = (int)((float)tx / 2.0f), hy = (int)((float)ty / 2.0f);
  const float rx = (float)(x - hx), ry = (float)(y - hy);
  const uint32_t t = ctl->type < rc.num_types ? ctl->type : rc.num_types - 1u;
  const float pr = rc.scale_y * tt[t].radius;
  if (!(rx * rx + ry * ry < (float)(hx * hy))) return;
  const int cx = px_int(rc.scale_x * ctl->pos[0]);
  const int cy = px_int(rc.scale_y * ctl->pos[1]);
  const uint32_t gx = (uint32_t)(cx + x) - tx / 2u, gy = (uint32_t)(cy + y) - ty / 2u;
  if (gx >= (uint32_t)rc.dimx || gy >= (uint32_t)rc.dimy) return;
  const float fscale = sqrtf((pr * pr - rx * rx - ry * ry) / (pr * pr));
  const uint32_t pixel = clamp_pixel(rc, gx + gy * (uint32_t)rc.dimx);
  img[pixel] = make_uchar4(to_byte(tt[t].r * fscale), to_byte(tt[t].g * fscale), to_byte(tt[t].b * fscale),
                           to_byte(255.0f * fscale));
}

// Largest pressure as an order-preserving unsigned key (NaN ignored).
__device__ inline uint32_t order_key(float v) {
  const uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ __launch_bounds__(256) void k_max_key(const float* __restrict__ v, uint32_t n, uint32_t* __restrict__ out) {
  uint32_t best = 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float x = v[i];
    if (x == x) best = max(best, order_key(x));
  }
  for (int o = 32; o > 0; o >>= 1) best = max(best, (uint32_t)__shfl_xor((int)best, o, kWave));
  if ((threadIdx.x & (kWave - 1)) == 0 && best) atomicMax(out, best);
}

__global__ void k_zero_word(uint32_t* p) { *p = 0u; }

// control square side in threads: computeGridSize(dim, 16) blocks x threads
uint32_t control_threads(int32_t dim) {
  if (dim <= 0) return 0u;
  const uint32_t t = dim < 16 ? (uint32_t)dim : 16u;
  return t * (((uint32_t)dim + t - 1u) / t);
}

int grow_u32(dem_context* ctx, uint32_t** p, size_t* cap, size_t need, const char* what) {
  if (*cap >= need) return DEM_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  DEM_HIP(ctx, hipMalloc((void**)p, sizeof(uint32_t) * need), what);
  *cap = need;
  return DEM_OK;
}

int plot(dem_context* ctx, uint8_t* image, const float* pos, const float* theta, const uint32_t* type,
         const float* radius, uint32_t n, const dem_render_params* rp, const dem_controlled* d_control,
         const float* pressure) {
  if (!image || !rp) return fail(ctx, DEM_ERR_INVALID, "NULL image or render parameters");
  if (n && (!pos || !theta || !type)) return fail(ctx, DEM_ERR_INVALID, "NULL array");
  if (rp->color_by_pressure && n && !pressure) return fail(ctx, DEM_ERR_INVALID, "pressure colouring needs pressure");
  if (rp->image_dim[0] <= 0 || rp->image_dim[1] <= 0 || (uint64_t)rp->image_dim[0] * rp->image_dim[1] > (1ull << 28))
    return fail(ctx, DEM_ERR_INVALID, "image size out of range (1 .. 2^28 pixels)");
  if (rp->control_dim[0] > 16384 || rp->control_dim[1] > 16384)
    return fail(ctx, DEM_ERR_INVALID, "control_dim out of range (<= 16384)");
  if (ctx->c.num_instances > 1) return fail(ctx, DEM_ERR_INVALID, "images are drawn for a single instance");
  if (ctx->h_color.size() != 3 * ctx->h_mat.size()) return fail(ctx, DEM_ERR_STATE, "particle colours not set");
  const Instance& inst = ctx->h_inst[0];
  RenderConst rc{};
  rc.dimx = rp->image_dim[0];
  rc.dimy = rp->image_dim[1];
  rc.cdimx = rp->control_dim[0];
  rc.cdimy = rp->control_dim[1];
  rc.scale_x = (float)rc.dimx / inst.box[0];
  rc.scale_y = (float)rc.dimy / inst.box[1];
  rc.npix = (uint32_t)(rc.dimx * rc.dimy);
  rc.rotations = rp->view_rotations ? 1u : 0u;
  rc.by_pressure = rp->color_by_pressure ? 1u : 0u;
  rc.max_pressure = rp->max_pressure;
  rc.num_types = (uint32_t)ctx->h_mat.size();
  std::vector<RenderType> h(rc.num_types);
  for (uint32_t t = 0; t < rc.num_types; ++t)
    h[t] = RenderType{ctx->h_mat[t].radius, ctx->h_color[3 * t], ctx->h_color[3 * t + 1], ctx->h_color[3 * t + 2]};
  size_t tcap = ctx->render_types_cap;
  if (int r = grow_u32(ctx, &ctx->d_render_types, &tcap, 4 * (size_t)rc.num_types, "hipMalloc(render)"))
    return r;
  ctx->render_types_cap = tcap;
  if (int r = grow_u32(ctx, &ctx->d_owner, &ctx->owner_cap, rc.npix, "hipMalloc(owner map)")) return r;
  hipStream_t s = ctx->stream;
  const RenderType* tt = (const RenderType*)ctx->d_render_types;
  DEM_HIP(ctx, hipMemcpyAsync(ctx->d_render_types, h.data(), sizeof(RenderType) * h.size(), hipMemcpyHostToDevice, s),
          "render table upload");
  uchar4* img = (uchar4*)image;
  hipLaunchKernelGGL(k_plot_clear, dim3((rc.npix + 255u) / 256u), dim3(256), 0, s, img, ctx->d_owner,
                     (uint8_t)(rp->bg_color & 0xFF), rc.npix);
  if (n) {
    const dim3 grid((n + 255u) / 256u);
    hipLaunchKernelGGL(k_plot_claim, grid, dim3(256), 0, s, rc, tt, (const float2*)pos, type, radius, ctx->d_owner, n);
    hipLaunchKernelGGL(k_plot_paint, grid, dim3(256), 0, s, rc, tt, (const float2*)pos, type, radius, pressure,
                       ctx->d_owner, img, n);
    if (rc.rotations)
      hipLaunchKernelGGL(k_plot_lines, grid, dim3(256), 0, s, rc, tt, (const float2*)pos, theta, type, radius, img, n);
  }
  if (d_control) {
    const uint32_t tx = control_threads(rc.cdimx), ty = control_threads(rc.cdimy);
    if (tx && ty)
      hipLaunchKernelGGL(k_plot_control, dim3((tx * ty + 255u) / 256u), dim3(256), 0, s, rc, tt, d_control, tx, ty,
                         img);
  }
  DEM_HIP(ctx, hipGetLastError(), "plot kernels");
  DEM_HIP(ctx, hipStreamSynchronize(s), "plot");
  return DEM_OK;
}

}  // namespace
}  // namespace dem

using namespace dem;

extern "C" {

int dem_plot_particles(dem_context* ctx, uint8_t* image, const float* pos, const float* theta, const uint32_t* type,
                       const float* radius, uint32_t numParticles, const dem_render_params* rp,
                       const dem_controlled* control, const float* pressure) {
  if (int r = check_ctx(ctx)) return r;
  if (int r = ensure_device(ctx)) return r;
  const dem_controlled* d_control = nullptr;
  if (control && control->enabled) {
    if (!ctx->d_stage_control) {
      DEM_HIP(ctx, hipMalloc((void**)&ctx->d_stage_control, sizeof(dem_controlled) + 4 * sizeof(float)),
              "hipMalloc(controlled)");
      ctx->d_stage_reaction = (float*)(ctx->d_stage_control + 1);
    }
    DEM_HIP(ctx, hipMemcpyAsync(ctx->d_stage_control, control, sizeof(dem_controlled), hipMemcpyHostToDevice,
                                ctx->stream),
            "controlled upload");
    d_control = ctx->d_stage_control;
  }
  return plot(ctx, image, pos, theta, type, radius, numParticles, rp, d_control, pressure);
}

int dem_sim_plot(dem_sim* m, uint8_t* image, const dem_render_params* rp) {
  if (int r = sim_check(m)) return r;
  dem_context* ctx = m->ctx;
  if (int r = ensure_device(ctx)) return r;
  const dem_controlled* d_control = ctx->h_control.enabled ? ctx->d_control : nullptr;
  return plot(ctx, image, (const float*)m->pos[m->cur], m->theta[m->cur], m->type[m->cur],
              m->rad ? m->radius[m->cur] : nullptr, m->n, rp, d_control, m->pressure);
}

int dem_update_pressure_scale(dem_context* ctx, const float* pressure, uint32_t numParticles, float* maxPressure) {
  if (!ctx || !maxPressure) return DEM_ERR_INVALID;
  if (numParticles && !pressure) return fail(ctx, DEM_ERR_INVALID, "NULL pressure");
  if (int r = ensure_device(ctx)) return r;
  size_t cap = ctx->scalar_cap;
  if (int r = grow_u32(ctx, &ctx->d_scalar, &cap, 1, "hipMalloc(scalar)")) return r;
  ctx->scalar_cap = cap;
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, s, ctx->d_scalar);
  if (numParticles) {
    const uint32_t blocks = std::min<uint32_t>((numParticles + 255u) / 256u, 2048u);
    hipLaunchKernelGGL(k_max_key, dim3(blocks), dim3(256), 0, s, pressure, numParticles, ctx->d_scalar);
  }
  uint32_t key = 0;
  DEM_HIP(ctx, hipGetLastError(), "pressure scale kernel");
  DEM_HIP(ctx, hipMemcpyAsync(&key, ctx->d_scalar, 4, hipMemcpyDeviceToHost, s), "pressure scale");
  DEM_HIP(ctx, hipStreamSynchronize(s), "pressure scale");
  float mx = 0.0f;
  if (key) {
    const uint32_t b = (key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key;
    memcpy(&mx, &b, 4);
  }
  *maxPressure = (key && mx > 0.0f) ? mx : 1.0f;
  return DEM_OK;
}

int dem_sim_pressure_scale(dem_sim* m, float* maxPressure) {
  if (int r = sim_check(m)) return r;
  return dem_update_pressure_scale(m->ctx, m->pressure, m->n, maxPressure);
}

}  // extern "C"
