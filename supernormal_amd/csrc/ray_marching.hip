// Occupancy-grid ray marching (AABB contraction), restating the semantics of
// csrc/ray_marching.cu:9-192 (calc_dt, grid_idx_at, grid_occupied_at,
// distance_to_next_voxel, advance_to_next_voxel, ray_marching_kernel).
//
// Floating point: every operation is an individually rounded IEEE fp32 op in
// the reference's order; this file is compiled with -ffp-contract=off so that
// no multiply-add is fused, which makes the sample positions bit-identical to
// the C oracle (oracle/sn_oracle.c) built the same way.
//
// Termination: the reference loops can spin forever when t + dt == t in fp32
// (e.g. t ~ 1e10 with dt ~ 1e-3).  Each ray here stops after kMaxIter loop
// trips; that only changes results where the reference would hang.
#include "sn_common.h"

namespace sn {

constexpr int64_t kMaxIter = 1 << 22;

struct RoiGrid {
  float mn[3], mx[3];
  int32_t res[3];
};

__device__ __forceinline__ float calc_dt(float t, float cone, float dt_min, float dt_max) {
  return fmaxf(dt_min, fminf(t * cone, dt_max));
}

// Cell index of a point, or -1 outside the roi (then unoccupied).
__device__ __forceinline__ int grid_index(const float xyz[3], const RoiGrid& g) {
  if (xyz[0] < g.mn[0] || xyz[0] > g.mx[0] || xyz[1] < g.mn[1] || xyz[1] > g.mx[1] ||
      xyz[2] < g.mn[2] || xyz[2] > g.mx[2])
    return -1;
  int ix[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float u = (xyz[d] - g.mn[d]) / (g.mx[d] - g.mn[d]);
    int v = (int)(u * (float)g.res[d]);  // truncation, helpers_math.h:177-180
    v = max(0, min(v, g.res[d] - 1));
    ix[d] = v;
  }
  return ix[0] * (g.res[1] * g.res[2]) + ix[1] * g.res[2] + ix[2];
}

// Occupancy lookups along one ray: consecutive samples usually fall in the
// same cell (the step is far below the cell size), so the last cell's value
// is kept and memory is read only when the cell changes.
struct CellCache {
  int idx = -2;
  bool occ = false;
  __device__ __forceinline__ bool lookup(const float xyz[3], const RoiGrid& g,
                                         const uint8_t* __restrict__ grid) {
    const int i = grid_index(xyz, g);
    if (i != idx) {
      idx = i;
      occ = i >= 0 && grid[i] != 0;
    }
    return occ;
  }
};

__device__ __forceinline__ float distance_to_next_voxel(const float xyz[3], const float dir[3],
                                                        const float inv_dir[3],
                                                        const RoiGrid& g) {
  float t[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float r = (float)g.res[d];
    float u = ((xyz[d] - g.mn[d]) / (g.mx[d] - g.mn[d])) * r;
    float sg = copysignf(1.0f, dir[d]);
    float nb = floorf((u + 0.5f) + 0.5f * sg);
    t[d] = (((nb - u) * inv_dir[d]) / r) * (g.mx[d] - g.mn[d]);
  }
  float tm = fminf(fminf(t[0], t[1]), t[2]);
  return fmaxf(tm, 0.0f);
}

// The roi is a device tensor in the reference's interface; every lane loads
// the 6 floats (one cached line) instead of a host round trip.
// kMode: kCount writes num_steps; kFill writes the packed samples at
// packed_info's bases; kPadded writes ray i's samples at i * max_per_ray and
// num_steps in one march (march_gather packs them afterwards).
enum { kCount = 0, kFill = 1, kPadded = 2 };
template <int kMode>
__global__ void __launch_bounds__(64)
ray_march_kernel_roi(const float* __restrict__ rays_o, const float* __restrict__ rays_d,
                     const float* __restrict__ t_min, const float* __restrict__ t_max,
                     int32_t n_rays, const float* __restrict__ roi, int32_t rx, int32_t ry,
                     int32_t rz, const uint8_t* __restrict__ grid, float step,
                     const float* __restrict__ step_dev, float cone,
                     int32_t max_per_ray, int32_t* __restrict__ num_steps,
                     const int32_t* __restrict__ packed_info, int64_t* __restrict__ ray_indices,
                     float* __restrict__ t_starts, float* __restrict__ t_ends);

}  // namespace sn

using namespace sn;

template <int kMode>
__global__ void __launch_bounds__(64)
sn::ray_march_kernel_roi(const float* __restrict__ rays_o, const float* __restrict__ rays_d,
                         const float* __restrict__ t_min, const float* __restrict__ t_max,
                         int32_t n_rays, const float* __restrict__ roi, int32_t rx,
                         int32_t ry, int32_t rz, const uint8_t* __restrict__ grid, float step,
                         const float* __restrict__ step_dev, float cone, int32_t max_per_ray,
                         int32_t* __restrict__ num_steps,
                         const int32_t* __restrict__ packed_info,
                         int64_t* __restrict__ ray_indices, float* __restrict__ t_starts,
                         float* __restrict__ t_ends) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rays) return;
  RoiGrid g;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    g.mn[d] = roi[d];
    g.mx[d] = roi[3 + d];
  }
  g.res[0] = rx;
  g.res[1] = ry;
  g.res[2] = rz;
  float o[3], dir[3], inv_dir[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    o[d] = rays_o[3 * i + d];
    dir[d] = rays_d[3 * i + d];
    inv_dir[d] = 1.0f / dir[d];
  }
  const float near = t_min[i], far = t_max[i];
  // the step may live in device memory (set per iteration without a host sync)
  const float dt_min = step_dev ? *step_dev : step, dt_max = 1e10f;
  if (!(dt_min > 0.f)) {
    if (kMode != kFill) num_steps[i] = 0;
    return;
  }
  int64_t base = 0;
  int32_t cap = max_per_ray > 0 ? max_per_ray : 0x7fffffff;
  if (kMode == kFill) {
    base = packed_info[2 * i];
    int32_t c = packed_info[2 * i + 1];
    if (c < cap) cap = c;
  } else if (kMode == kPadded) {
    base = (int64_t)i * max_per_ray;
  }
  int32_t j = 0;
  float t0 = near;
  float dt = calc_dt(t0, cone, dt_min, dt_max);
  float t1 = t0 + dt;
  float t_mid = (t0 + t1) * 0.5f;
  int64_t iter = 0;
  CellCache cells;
  while (t_mid < far && j < cap && iter < kMaxIter) {
    float xyz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xyz[d] = o[d] + t_mid * dir[d];
    if (cells.lookup(xyz, g, grid)) {
      if (kMode != kCount) {
        t_starts[base + j] = t0;
        t_ends[base + j] = t1;
        if (kMode == kFill) ray_indices[base + j] = i;
      }
      ++j;
      ++iter;
      t0 = t1;
      t1 = t0 + calc_dt(t0, cone, dt_min, dt_max);
      t_mid = (t0 + t1) * 0.5f;
    } else {
      float t_target = t_mid + distance_to_next_voxel(xyz, dir, inv_dir, g);
      t_target = fminf(t_target, far);
      float tt = t_mid;
      do {
        tt += dt_min;
        ++iter;
      } while (tt < t_target && iter < kMaxIter);
      t_mid = tt;
      dt = calc_dt(t_mid, cone, dt_min, dt_max);
      t0 = t_mid - dt * 0.5f;
      t1 = t_mid + dt * 0.5f;
    }
  }
  if (kMode != kFill) num_steps[i] = j;
}

// Packs the padded per-ray samples at packed_info's bases (exclusive scan of
// num_steps): the result equals what the fill pass writes.
__global__ void __launch_bounds__(256)
march_gather_kernel(const int32_t* __restrict__ packed_info, int32_t n_rays, int32_t slot,
                    const float* __restrict__ t0p, const float* __restrict__ t1p,
                    int64_t* __restrict__ ray_indices, float* __restrict__ t_starts,
                    float* __restrict__ t_ends) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= (int64_t)n_rays * slot) return;
  const int32_t i = (int32_t)(g / slot), j = (int32_t)(g - (int64_t)i * slot);
  const int32_t base = packed_info[2 * i], cnt = packed_info[2 * i + 1];
  if (j >= cnt) return;
  t_starts[base + j] = t0p[g];
  t_ends[base + j] = t1p[g];
  ray_indices[base + j] = i;
}

static int check_march_args(const char* who, const float* o, const float* d, const float* tmin,
                            const float* tmax, int32_t n_rays, const float* roi,
                            const uint8_t* grid, int32_t rx, int32_t ry, int32_t rz,
                            float step, float cone) {
  SN_REQUIRE(n_rays >= 0, "%s: n_rays < 0", who);
  SN_REQUIRE(rx > 0 && ry > 0 && rz > 0, "%s: bad grid resolution", who);
  SN_REQUIRE((int64_t)rx * ry * rz < (1ll << 31), "%s: grid too large", who);
  SN_REQUIRE(step > 0.f, "%s: render_step_size must be > 0 (got %g)", who, (double)step);
  SN_REQUIRE(cone >= 0.f, "%s: cone_angle must be >= 0", who);
  SN_REQUIRE(n_rays == 0 || (o && d && tmin && tmax && roi && grid), "%s: null pointer", who);
  return SN_OK;
}

extern "C" {

int sn_ray_march_count(const float* rays_o, const float* rays_d, const float* t_min,
                       const float* t_max, int32_t n_rays, const float* roi_aabb,
                       const uint8_t* grid, int32_t rx, int32_t ry, int32_t rz,
                       float step_size, const float* step_dev, float cone_angle,
                       int32_t max_per_ray, int32_t* num_steps, void* stream) {
  clear_error();
  int rc = check_march_args("sn_ray_march_count", rays_o, rays_d, t_min, t_max, n_rays,
                            roi_aabb, grid, rx, ry, rz, step_dev ? 1.f : step_size, cone_angle);
  if (rc) return rc;
  if (n_rays == 0) return SN_OK;
  SN_REQUIRE(num_steps, "sn_ray_march_count: null num_steps");
  ray_march_kernel_roi<kCount><<<(unsigned)ceil_div(n_rays, 64), 64, 0, as_stream(stream)>>>(
      rays_o, rays_d, t_min, t_max, n_rays, roi_aabb, rx, ry, rz, grid, step_size, step_dev,
      cone_angle, max_per_ray, num_steps, nullptr, nullptr, nullptr, nullptr);
  return check_launch("sn_ray_march_count");
}

int sn_ray_march_fill(const float* rays_o, const float* rays_d, const float* t_min,
                      const float* t_max, int32_t n_rays, const float* roi_aabb,
                      const uint8_t* grid, int32_t rx, int32_t ry, int32_t rz,
                      float step_size, const float* step_dev, float cone_angle,
                      int32_t max_per_ray, const int32_t* packed_info, int64_t* ray_indices,
                      float* t_starts, float* t_ends, void* stream) {
  clear_error();
  int rc = check_march_args("sn_ray_march_fill", rays_o, rays_d, t_min, t_max, n_rays,
                            roi_aabb, grid, rx, ry, rz, step_dev ? 1.f : step_size, cone_angle);
  if (rc) return rc;
  if (n_rays == 0) return SN_OK;
  SN_REQUIRE(packed_info && ray_indices && t_starts && t_ends,
             "sn_ray_march_fill: null output");
  ray_march_kernel_roi<kFill><<<(unsigned)ceil_div(n_rays, 64), 64, 0, as_stream(stream)>>>(
      rays_o, rays_d, t_min, t_max, n_rays, roi_aabb, rx, ry, rz, grid, step_size, step_dev,
      cone_angle, max_per_ray, nullptr, packed_info, ray_indices, t_starts, t_ends);
  return check_launch("sn_ray_march_fill");
}

int sn_ray_march_padded(const float* rays_o, const float* rays_d, const float* t_min,
                        const float* t_max, int32_t n_rays, const float* roi_aabb,
                        const uint8_t* grid, int32_t rx, int32_t ry, int32_t rz,
                        float step_size, const float* step_dev, float cone_angle,
                        int32_t max_per_ray, int32_t* num_steps, float* t_starts_pad,
                        float* t_ends_pad, void* stream) {
  clear_error();
  int rc = check_march_args("sn_ray_march_padded", rays_o, rays_d, t_min, t_max, n_rays,
                            roi_aabb, grid, rx, ry, rz, step_dev ? 1.f : step_size, cone_angle);
  if (rc) return rc;
  SN_REQUIRE(max_per_ray > 0, "sn_ray_march_padded: max_per_ray must be > 0");
  if (n_rays == 0) return SN_OK;
  SN_REQUIRE(num_steps && t_starts_pad && t_ends_pad, "sn_ray_march_padded: null output");
  ray_march_kernel_roi<kPadded><<<(unsigned)ceil_div(n_rays, 64), 64, 0, as_stream(stream)>>>(
      rays_o, rays_d, t_min, t_max, n_rays, roi_aabb, rx, ry, rz, grid, step_size, step_dev,
      cone_angle, max_per_ray, num_steps, nullptr, nullptr, t_starts_pad, t_ends_pad);
  return check_launch("sn_ray_march_padded");
}

int sn_march_gather(const int32_t* packed_info, int32_t n_rays, int32_t max_per_ray,
                    const float* t_starts_pad, const float* t_ends_pad, int64_t* ray_indices,
                    float* t_starts, float* t_ends, void* stream) {
  clear_error();
  SN_REQUIRE(n_rays >= 0 && max_per_ray > 0, "sn_march_gather: bad sizes");
  if (n_rays == 0) return SN_OK;
  SN_REQUIRE(packed_info && t_starts_pad && t_ends_pad && ray_indices && t_starts && t_ends,
             "sn_march_gather: null pointer");
  const int64_t total = (int64_t)n_rays * max_per_ray;
  march_gather_kernel<<<(unsigned)ceil_div(total, 256), 256, 0, as_stream(stream)>>>(
      packed_info, n_rays, max_per_ray, t_starts_pad, t_ends_pad, ray_indices, t_starts, t_ends);
  return check_launch("sn_march_gather");
}

}  // extern "C"
