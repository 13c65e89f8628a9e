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
  // 1/(mx - mn) and 1/res where those are powers of two (kPow2 paths): then
  // x / v == x * (1/v) exactly in IEEE arithmetic, so the divisions of the
  // reference's formulas become multiplications with identical results.
  float inv_span[3], inv_res[3];
};

__device__ __forceinline__ bool is_pow2f(float v) {
  const uint32_t b = __float_as_uint(v);
  return (b & 0x7fffffu) == 0u && ((b >> 23) & 0xffu) != 0u && ((b >> 23) & 0xffu) < 0xfeu &&
         !(b >> 31);
}

__device__ __forceinline__ float calc_dt(float t, float cone, float dt_min, float dt_max) {
  return fmaxf(dt_min, fminf(t * cone, dt_max));
}

// Cell of a point (false outside the roi: then unoccupied).
template <bool kPow2>
__device__ __forceinline__ bool grid_cell(const float xyz[3], const RoiGrid& g, int ix[3]) {
  if (xyz[0] < g.mn[0] || xyz[0] > g.mx[0] || xyz[1] < g.mn[1] || xyz[1] > g.mx[1] ||
      xyz[2] < g.mn[2] || xyz[2] > g.mx[2])
    return false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float u = kPow2 ? (xyz[d] - g.mn[d]) * g.inv_span[d]
                          : (xyz[d] - g.mn[d]) / (g.mx[d] - g.mn[d]);
    int v = (int)(u * (float)g.res[d]);  // truncation, helpers_math.h:177-180
    ix[d] = max(0, min(v, g.res[d] - 1));
  }
  return true;
}

// Occupancy of the cell holding xyz: from the byte grid, or from 4x4x4 brick
// masks (one 64-bit word per brick) when given.
template <bool kPow2>
__device__ __forceinline__ bool occupied(const float xyz[3], const RoiGrid& g,
                                         const uint8_t* __restrict__ grid,
                                         const uint64_t* __restrict__ bricks) {
  int ix[3];
  if (!grid_cell<kPow2>(xyz, g, ix)) return false;
  if (bricks) {
    const int b = ((ix[0] >> 2) * (g.res[1] >> 2) + (ix[1] >> 2)) * (g.res[2] >> 2) + (ix[2] >> 2);
    return (bricks[b] >> ((ix[0] & 3) | ((ix[1] & 3) << 2) | ((ix[2] & 3) << 4))) & 1ull;
  }
  return grid[ix[0] * (g.res[1] * g.res[2]) + ix[1] * g.res[2] + ix[2]] != 0;
}

template <bool kPow2>
__device__ __forceinline__ float distance_to_next_voxel(const float xyz[3], const float dir[3],
                                                        const float inv_dir[3],
                                                        const RoiGrid& g) {
  float t[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float r = (float)g.res[d];
    float u = kPow2 ? ((xyz[d] - g.mn[d]) * g.inv_span[d]) * r
                    : ((xyz[d] - g.mn[d]) / (g.mx[d] - g.mn[d])) * r;
    float sg = copysignf(1.0f, dir[d]);
    float nb = floorf((u + 0.5f) + 0.5f * sg);
    t[d] = kPow2 ? (((nb - u) * inv_dir[d]) * g.inv_res[d]) * (g.mx[d] - g.mn[d])
                 : (((nb - u) * inv_dir[d]) / r) * (g.mx[d] - g.mn[d]);
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

// One wavefront per ray.  The reference's loop is sequential in t; here a
// trip is first tested by the whole wavefront (uniform values), and a run of
// occupied trips is then evaluated up to L trips at once, with the very same
// fp32 operations, so the samples are identical:
//  * from an occupied trip (t0, t1, t_mid), lane k plays the trip reached
//    after k occupied steps (k times t0 <- t1, t1 <- t0 + dt,
//    t_mid <- (t0 + t1) / 2; lane 0 is the trip itself).  The first lane whose
//    trip fails the loop condition or finds its cell empty ends the run: the
//    lanes before it are samples, and the wavefront resumes at that trip.  L
//    is 16 for a run's first round and 64 after a full round;
//  * at an empty cell the skip to the next voxel adds dt_min to t_mid until
//    it reaches the target; lane k holds the value after k + 1 additions and
//    the first lane at or past the target (or at the trip limit) gives the
//    new t_mid.
template <int kMode, bool kPow2>
__device__ __forceinline__ void march_ray_wave(int32_t i, const float o[3], const float dir[3],
                                               const float inv_dir[3], float near, float far,
                                               const RoiGrid& g, const uint8_t* __restrict__ grid,
                                               const uint64_t* __restrict__ bricks, float dt_min,
                                               float dt_max, float cone, int32_t cap,
                                               int64_t base, int32_t* __restrict__ num_steps,
                                               int64_t* __restrict__ ray_indices,
                                               float* __restrict__ t_starts,
                                               float* __restrict__ t_ends) {
  const int lane = threadIdx.x & 63;
  int32_t j = 0;
  int64_t iter = 0;
  float t0 = near;
  float dt = calc_dt(t0, cone, dt_min, dt_max);
  float t1 = t0 + dt;
  float t_mid = (t0 + t1) * 0.5f;
  int L = 16;
  for (;;) {  // uniform: every value that is not per-lane is the same in all lanes
    if (!(t_mid < far && j < cap && iter < kMaxIter)) break;
    float xyz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xyz[d] = o[d] + t_mid * dir[d];
    if (!occupied<kPow2>(xyz, g, grid, bricks)) {
      // empty cell: skip to the next voxel
      float t_target = t_mid + distance_to_next_voxel<kPow2>(xyz, dir, inv_dir, g);
      t_target = fminf(t_target, far);
      float tt = t_mid;
      for (;;) {
        // additions needed, estimated (a low estimate only costs another round)
        const float est = (t_target - tt) / dt_min;
        const int n = est < 62.f ? max(1, (int)est + 2) : 64;
        float tk = tt;
        for (int s = 0; s < n; ++s)
          if (s <= lane) tk += dt_min;
        const bool done = lane < n && !(tk < t_target && iter + lane + 1 < kMaxIter);
        const uint64_t dm = __ballot(done);
        if (dm) {
          const int k = __ffsll((unsigned long long)dm) - 1;
          tt = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tk), k));
          iter += k + 1;
          break;
        }
        tt = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tk), n - 1));
        iter += n;
      }
      t_mid = tt;
      dt = calc_dt(t_mid, cone, dt_min, dt_max);
      t0 = t_mid - dt * 0.5f;
      t1 = t_mid + dt * 0.5f;
      L = 16;
      continue;
    }
    // an occupied trip: evaluate the run from it, L trips at once
    float a0 = t0, a1 = t1, am = t_mid;
    for (int s = 0; s < L - 1; ++s) {
      if (s < lane) {
        a0 = a1;
        a1 = a0 + calc_dt(a0, cone, dt_min, dt_max);
        am = (a0 + a1) * 0.5f;
      }
    }
    bool occ = lane == 0;
    if (lane > 0 && lane < L && am < far && j + lane < cap && iter + lane < kMaxIter) {
#pragma unroll
      for (int d = 0; d < 3; ++d) xyz[d] = o[d] + am * dir[d];
      occ = occupied<kPow2>(xyz, g, grid, bricks);
    }
    const uint64_t stop = __ballot(!occ);
    // 1 <= f <= L: lanes >= L stop, and with L = 64 every lane may be a sample
    const int f = stop ? __ffsll((unsigned long long)stop) - 1 : 64;
    if (kMode != kCount && lane < f) {
      t_starts[base + j + lane] = a0;
      t_ends[base + j + lane] = a1;
      if (kMode == kFill) ray_indices[base + j + lane] = i;
    }
    j += f;
    iter += f;
    if (f == L) {
      // a full round: resume after lane L-1's trip, with rounds of 64
      t0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a1), L - 1));
      t1 = t0 + calc_dt(t0, cone, dt_min, dt_max);
      t_mid = (t0 + t1) * 0.5f;
      L = 64;
      continue;
    }
    // resume at trip f (it failed the condition or is empty: the next pass decides)
    t0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a0), f));
    t1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a1), f));
    t_mid = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(am), f));
  }
  if (kMode != kFill && lane == 0) num_steps[i] = j;
}

template <int kMode>
__global__ void __launch_bounds__(256)
ray_march_kernel_roi(const float* __restrict__ rays_o, const float* __restrict__ rays_d,
                     const float* __restrict__ t_min, const float* __restrict__ t_max,
                     int32_t n_rays, const float* __restrict__ roi, int32_t rx, int32_t ry,
                     int32_t rz, const uint8_t* __restrict__ grid,
                     const uint64_t* __restrict__ bricks, float step,
                     const float* __restrict__ step_dev, float cone, int32_t max_per_ray,
                     int32_t* __restrict__ num_steps, const int32_t* __restrict__ packed_info,
                     int64_t* __restrict__ ray_indices, float* __restrict__ t_starts,
                     float* __restrict__ t_ends) {
  // one wavefront per ray (uniform: taken from the wavefront's first lane)
  const int64_t ii = (int64_t)blockIdx.x * 4 + (__builtin_amdgcn_readfirstlane(threadIdx.x) >> 6);
  if (ii >= n_rays) return;
  const int32_t i = (int32_t)ii;
  RoiGrid g;
  bool pow2 = true;
  g.res[0] = rx;
  g.res[1] = ry;
  g.res[2] = rz;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    g.mn[d] = roi[d];
    g.mx[d] = roi[3 + d];
    const float span = g.mx[d] - g.mn[d], r = (float)g.res[d];
    pow2 = pow2 && is_pow2f(span) && is_pow2f(r);
    g.inv_span[d] = 1.0f / span;
    g.inv_res[d] = 1.0f / r;
  }
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
    if (kMode != kFill && (threadIdx.x & 63) == 0) num_steps[i] = 0;
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
  if (pow2)
    march_ray_wave<kMode, true>(i, o, dir, inv_dir, near, far, g, grid, bricks, dt_min, dt_max,
                                cone, cap, base, num_steps, ray_indices, t_starts, t_ends);
  else
    march_ray_wave<kMode, false>(i, o, dir, inv_dir, near, far, g, grid, bricks, dt_min, dt_max,
                                 cone, cap, base, num_steps, ray_indices, t_starts, t_ends);
}

// Brick masks of the occupancy grid: bit (x&3) | (y&3)<<2 | (z&3)<<4 of brick
// ((x>>2) * ry/4 + (y>>2)) * rz/4 + (z>>2) is binary[x][y][z] != 0.
__global__ void __launch_bounds__(256)
occ_pack_bricks_kernel(const uint8_t* __restrict__ binary, int32_t rx, int32_t ry, int32_t rz,
                       uint64_t* __restrict__ bricks) {
  const int bx = rx >> 2, by = ry >> 2, bz = rz >> 2;
  const int64_t nb = (int64_t)bx * by * bz;
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nb;
       b += (int64_t)gridDim.x * blockDim.x) {
    const int x0 = (int)(b / ((int64_t)by * bz)) * 4, y0 = (int)((b / bz) % by) * 4,
              z0 = (int)(b % bz) * 4;
    uint64_t m = 0;
    for (int x = 0; x < 4; ++x)
      for (int y = 0; y < 4; ++y) {
        const uint8_t* row = binary + ((int64_t)(x0 + x) * ry + (y0 + y)) * rz + z0;
#pragma unroll
        for (int z = 0; z < 4; ++z)
          if (row[z]) m |= 1ull << (x | (y << 2) | (z << 4));
      }
    bricks[b] = m;
  }
}

// Packs the padded per-ray samples at packed_info's bases (exclusive scan of
// num_steps): the result equals what the fill pass writes.
__global__ void __launch_bounds__(256)
march_gather_kernel(const int32_t* __restrict__ packed_info, int32_t n_rays, int32_t slot,
                    const float* __restrict__ t0p, const float* __restrict__ t1p,
                    int64_t* __restrict__ ray_indices, float* __restrict__ t_starts,
                    float* __restrict__ t_ends, int32_t* __restrict__ n_capped) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= (int64_t)n_rays * slot) return;
  const int32_t i = (int32_t)(g / slot), j = (int32_t)(g - (int64_t)i * slot);
  const int32_t base = packed_info[2 * i], cnt = packed_info[2 * i + 1];
  if (j >= cnt) return;
  if (n_capped && j == slot - 1) atomicAdd(n_capped, 1);  // the ray filled its slots
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

}  // namespace sn

using namespace sn;

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
  ray_march_kernel_roi<kCount><<<(unsigned)ceil_div(n_rays, 4), 256, 0, as_stream(stream)>>>(
      rays_o, rays_d, t_min, t_max, n_rays, roi_aabb, rx, ry, rz, grid, nullptr, step_size, step_dev,
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
  ray_march_kernel_roi<kFill><<<(unsigned)ceil_div(n_rays, 4), 256, 0, as_stream(stream)>>>(
      rays_o, rays_d, t_min, t_max, n_rays, roi_aabb, rx, ry, rz, grid, nullptr, step_size, step_dev,
      cone_angle, max_per_ray, nullptr, packed_info, ray_indices, t_starts, t_ends);
  return check_launch("sn_ray_march_fill");
}

int sn_ray_march_padded_bricks(const float* rays_o, const float* rays_d, const float* t_min,
                               const float* t_max, int32_t n_rays, const float* roi_aabb,
                               const uint8_t* grid, const uint64_t* bricks, int32_t rx,
                               int32_t ry, int32_t rz, float step_size, const float* step_dev,
                               float cone_angle, int32_t max_per_ray, int32_t* num_steps,
                               float* t_starts_pad, float* t_ends_pad, void* stream) {
  clear_error();
  int rc = check_march_args("sn_ray_march_padded", rays_o, rays_d, t_min, t_max, n_rays,
                            roi_aabb, grid, rx, ry, rz, step_dev ? 1.f : step_size, cone_angle);
  if (rc) return rc;
  SN_REQUIRE(max_per_ray > 0, "sn_ray_march_padded: max_per_ray must be > 0");
  SN_REQUIRE(!bricks || (rx % 4 == 0 && ry % 4 == 0 && rz % 4 == 0),
             "sn_ray_march_padded: brick masks need resolutions divisible by 4");
  if (n_rays == 0) return SN_OK;
  SN_REQUIRE(num_steps && t_starts_pad && t_ends_pad, "sn_ray_march_padded: null output");
  ray_march_kernel_roi<kPadded><<<(unsigned)ceil_div(n_rays, 4), 256, 0, as_stream(stream)>>>(
      rays_o, rays_d, t_min, t_max, n_rays, roi_aabb, rx, ry, rz, grid, bricks, step_size,
      step_dev, cone_angle, max_per_ray, num_steps, nullptr, nullptr, t_starts_pad, t_ends_pad);
  return check_launch("sn_ray_march_padded");
}

int sn_ray_march_padded(const float* rays_o, const float* rays_d, const float* t_min,
                        const float* t_max, int32_t n_rays, const float* roi_aabb,
                        const uint8_t* grid, int32_t rx, int32_t ry, int32_t rz,
                        float step_size, const float* step_dev, float cone_angle,
                        int32_t max_per_ray, int32_t* num_steps, float* t_starts_pad,
                        float* t_ends_pad, void* stream) {
  return sn_ray_march_padded_bricks(rays_o, rays_d, t_min, t_max, n_rays, roi_aabb, grid,
                                    nullptr, rx, ry, rz, step_size, step_dev, cone_angle,
                                    max_per_ray, num_steps, t_starts_pad, t_ends_pad, stream);
}

int sn_occ_pack_bricks(const uint8_t* binary, int32_t rx, int32_t ry, int32_t rz,
                       uint64_t* bricks, void* stream) {
  clear_error();
  SN_REQUIRE(rx > 0 && ry > 0 && rz > 0 && rx % 4 == 0 && ry % 4 == 0 && rz % 4 == 0,
             "sn_occ_pack_bricks: resolutions must be positive multiples of 4");
  SN_REQUIRE((int64_t)rx * ry * rz < (1ll << 31), "sn_occ_pack_bricks: grid too large");
  SN_REQUIRE(binary && bricks, "sn_occ_pack_bricks: null pointer");
  const int64_t nb = (int64_t)(rx / 4) * (ry / 4) * (rz / 4);
  occ_pack_bricks_kernel<<<grid_for(nb, 256), 256, 0, as_stream(stream)>>>(binary, rx, ry, rz,
                                                                           bricks);
  return check_launch("sn_occ_pack_bricks");
}

int sn_march_gather(const int32_t* packed_info, int32_t n_rays, int32_t max_per_ray,
                    const float* t_starts_pad, const float* t_ends_pad, int64_t* ray_indices,
                    float* t_starts, float* t_ends, int32_t* n_capped, void* stream) {
  clear_error();
  SN_REQUIRE(n_rays >= 0 && max_per_ray > 0, "sn_march_gather: bad sizes");
  if (n_rays == 0) return SN_OK;
  SN_REQUIRE(packed_info && t_starts_pad && t_ends_pad && ray_indices && t_starts && t_ends,
             "sn_march_gather: null pointer");
  const int64_t total = (int64_t)n_rays * max_per_ray;
  march_gather_kernel<<<(unsigned)ceil_div(total, 256), 256, 0, as_stream(stream)>>>(
      packed_info, n_rays, max_per_ray, t_starts_pad, t_ends_pad, ray_indices, t_starts, t_ends,
      n_capped);
  return check_launch("sn_march_gather");
}

}  // extern "C"
