/*
 * sn_oracle.c -- CPU oracle for SuperNormal's training hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP kernels are
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never calls it.
 *
 * Sequential restatements, one function per reference routine (paths relative
 * to the SuperNormal repository, NA/ = third_parties/nerfacc-0.3.5/nerfacc-0.3.5/,
 * csrc/ = NA/nerfacc/cuda/csrc/).  Built with -ffp-contract=off: every fp32
 * operation is individually rounded, in the reference's order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* pack_info: NA/nerfacc/pack.py:63-74 (scatter_add of ones, cumsum, base =
 * cumsum - count).  Out-of-range indices are skipped and counted. */
int or_pack_info(const int64_t* ri, int64_t n, int32_t n_rays, int32_t* packed) {
  int bad = 0;
  for (int32_t r = 0; r < n_rays; ++r) packed[2 * r + 1] = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (ri[i] < 0 || ri[i] >= n_rays) { ++bad; continue; }
    packed[2 * ri[i] + 1] += 1;
  }
  int32_t c = 0;
  for (int32_t r = 0; r < n_rays; ++r) {
    packed[2 * r] = c;
    c += packed[2 * r + 1];
  }
  return bad;
}

/* Patch weights + transmittance forward: csrc/render_weight.cu:87-118 and
 * 120-152 (loop over the patch's samples j for each ray k). */
void or_patch_weight_fwd(const int32_t* packed, int32_t n_patches, int32_t P,
                         const float* alphas, float* w, float* trans) {
  for (int32_t p = 0; p < n_patches; ++p) {
    const int32_t base = packed[2 * p], steps = packed[2 * p + 1];
    for (int32_t k = 0; k < P; ++k) {
      float T = 1.f;
      for (int32_t j = 0; j < steps; ++j) {
        const int64_t idx = ((int64_t)base + j) * P + k;
        const float a = alphas[idx];
        if (trans) trans[idx] = T;
        if (w) w[idx] = a * T;
        T *= (1.f - a);
      }
    }
  }
}

/* Reference backward formula, csrc/render_weight.cu:298-340:
 *   accum = sum_j gw*w ; ga = (gw*T - accum)/max(1-a,1e-10); accum -= gw*w. */
void or_patch_weight_bwd_ref(const int32_t* packed, int32_t n_patches, int32_t P,
                             const float* alphas, const float* w, const float* gw, float* ga) {
  for (int32_t p = 0; p < n_patches; ++p) {
    const int32_t base = packed[2 * p], steps = packed[2 * p + 1];
    for (int32_t k = 0; k < P; ++k) {
      float accum = 0.f;
      for (int32_t j = 0; j < steps; ++j) {
        const int64_t idx = ((int64_t)base + j) * P + k;
        accum += gw[idx] * w[idx];
      }
      float T = 1.f;
      for (int32_t j = 0; j < steps; ++j) {
        const int64_t idx = ((int64_t)base + j) * P + k;
        const float a = alphas[idx];
        ga[idx] = (gw[idx] * T - accum) / fmaxf(1.f - a, 1e-10f);
        accum -= gw[idx] * w[idx];
        T *= (1.f - a);
      }
    }
  }
}

/* Division-free form of the same gradient (plus the transmittance term):
 * g_a[j] = T[j] (gw[j] - Q[j]), Q[j] = (gw[j+1] a[j+1] + gT[j+1]) + (1-a[j+1]) Q[j+1].
 * Computed in double so that it is a reference, not a copy, of the fp32 kernel. */
void or_patch_weight_bwd_f64(const int32_t* packed, int32_t n_patches, int32_t P,
                             const double* alphas, const double* gw, const double* gT,
                             double* ga) {
  for (int32_t p = 0; p < n_patches; ++p) {
    const int32_t base = packed[2 * p], steps = packed[2 * p + 1];
    for (int32_t k = 0; k < P; ++k) {
      double T = 1.0;
      for (int32_t j = 0; j < steps; ++j) {
        const int64_t idx = ((int64_t)base + j) * P + k;
        ga[idx] = T;
        T *= (1.0 - alphas[idx]);
      }
      double Q = 0.0;
      for (int32_t j = steps - 1; j >= 0; --j) {
        const int64_t idx = ((int64_t)base + j) * P + k;
        const double a = alphas[idx];
        const double g1 = gw ? gw[idx] : 0.0, g2 = gT ? gT[idx] : 0.0;
        ga[idx] = ga[idx] * (g1 - Q);
        Q = (g1 * a + g2) + (1.0 - a) * Q;
      }
    }
  }
}

/* Visibility: sequential exclusive product csrc/render_transmittance.cu:85-112
 * and the test of NA/nerfacc/vol_rendering.py:744-747. */
void or_visibility(const int32_t* packed, int32_t n_rays, const float* alphas, float eps,
                   float alpha_thre, uint8_t* keep) {
  for (int32_t r = 0; r < n_rays; ++r) {
    const int32_t base = packed[2 * r], steps = packed[2 * r + 1];
    float T = 1.f;
    for (int32_t j = 0; j < steps; ++j) {
      const float a = alphas[base + j];
      int v = T >= eps;
      if (alpha_thre > 0.f) v = v && (a >= alpha_thre);
      keep[base + j] = (uint8_t)v;
      T *= (1.f - a);
    }
  }
}

/* ---- ray marching: csrc/ray_marching.cu:9-192 ---------------------------- */
static float calc_dt(float t, float cone, float dt_min, float dt_max) {
  /* clamp(t*cone, dt_min, dt_max) = fmaxf(dt_min, fminf(t*cone, dt_max)), helpers_math.h:1167-1170 */
  return fmaxf(dt_min, fminf(t * cone, dt_max));
}

static int occupied_at(const float xyz[3], const float* roi, const uint8_t* grid,
                       const int32_t res[3]) {
  for (int d = 0; d < 3; ++d)
    if (xyz[d] < roi[d] || xyz[d] > roi[3 + d]) return 0;
  int ix[3];
  for (int d = 0; d < 3; ++d) {
    const float u = (xyz[d] - roi[d]) / (roi[3 + d] - roi[d]);
    int v = (int)(u * (float)res[d]); /* make_int3(float3): truncation */
    if (v < 0) v = 0;
    if (v > res[d] - 1) v = res[d] - 1;
    ix[d] = v;
  }
  return grid[ix[0] * (res[1] * res[2]) + ix[1] * res[2] + ix[2]] != 0;
}

static float distance_to_next_voxel(const float xyz[3], const float dir[3],
                                    const float inv_dir[3], const float* roi,
                                    const int32_t res[3]) {
  float t[3];
  for (int d = 0; d < 3; ++d) {
    const float r = (float)res[d];
    const float u = ((xyz[d] - roi[d]) / (roi[3 + d] - roi[d])) * r;
    const float sg = copysignf(1.0f, dir[d]);
    const float nb = floorf((u + 0.5f) + 0.5f * sg);
    t[d] = (((nb - u) * inv_dir[d]) / r) * (roi[3 + d] - roi[d]);
  }
  return fmaxf(fminf(fminf(t[0], t[1]), t[2]), 0.0f);
}

#define OR_MAX_ITER (1LL << 22)

/* One ray.  If out_t0 is NULL only counts.  Returns the number of samples. */
static int32_t march_one(const float o[3], const float dir[3], float near, float far,
                         const float* roi, const uint8_t* grid, const int32_t res[3],
                         float step, float cone, int32_t cap, int64_t ray, int64_t* out_ri,
                         float* out_t0, float* out_t1) {
  float inv_dir[3];
  for (int d = 0; d < 3; ++d) inv_dir[d] = 1.0f / dir[d];
  const float dt_min = step, dt_max = 1e10f;
  int32_t j = 0;
  float t0 = near;
  float dt = calc_dt(t0, cone, dt_min, dt_max);
  float t1 = t0 + dt;
  float t_mid = (t0 + t1) * 0.5f;
  long long iter = 0;
  while (t_mid < far && j < cap && iter < OR_MAX_ITER) {
    float xyz[3];
    for (int d = 0; d < 3; ++d) xyz[d] = o[d] + t_mid * dir[d];
    if (occupied_at(xyz, roi, grid, res)) {
      if (out_t0) {
        out_t0[j] = t0;
        out_t1[j] = t1;
        out_ri[j] = ray;
      }
      ++j;
      ++iter;
      t0 = t1;
      t1 = t0 + calc_dt(t0, cone, dt_min, dt_max);
      t_mid = (t0 + t1) * 0.5f;
    } else {
      float t_target = t_mid + distance_to_next_voxel(xyz, dir, inv_dir, roi, res);
      t_target = fminf(t_target, far);
      float tt = t_mid;
      do {
        tt += dt_min;
        ++iter;
      } while (tt < t_target && iter < OR_MAX_ITER);
      t_mid = tt;
      dt = calc_dt(t_mid, cone, dt_min, dt_max);
      t0 = t_mid - dt * 0.5f;
      t1 = t_mid + dt * 0.5f;
    }
  }
  return j;
}

/* Pass 1: num_steps per ray (csrc/ray_marching.cu:235-255). */
void or_ray_march_count(const float* rays_o, const float* rays_d, const float* tmin,
                        const float* tmax, int32_t n_rays, const float* roi,
                        const uint8_t* grid, int32_t rx, int32_t ry, int32_t rz, float step,
                        float cone, int32_t max_per_ray, int32_t* num_steps) {
  const int32_t res[3] = {rx, ry, rz};
  const int32_t cap = max_per_ray > 0 ? max_per_ray : 0x7fffffff;
  for (int32_t i = 0; i < n_rays; ++i)
    num_steps[i] = march_one(rays_o + 3 * i, rays_d + 3 * i, tmin[i], tmax[i], roi, grid, res,
                             step, cone, cap, i, NULL, NULL, NULL);
}

/* Pass 2: fill at packed_info bases (csrc/ray_marching.cu:257-286). */
void or_ray_march_fill(const float* rays_o, const float* rays_d, const float* tmin,
                       const float* tmax, int32_t n_rays, const float* roi, const uint8_t* grid,
                       int32_t rx, int32_t ry, int32_t rz, float step, float cone,
                       int32_t max_per_ray, const int32_t* packed, int64_t* ri, float* t0,
                       float* t1) {
  const int32_t res[3] = {rx, ry, rz};
  for (int32_t i = 0; i < n_rays; ++i) {
    int32_t cap = packed[2 * i + 1];
    if (max_per_ray > 0 && max_per_ray < cap) cap = max_per_ray;
    const int64_t b = packed[2 * i];
    march_one(rays_o + 3 * i, rays_d + 3 * i, tmin[i], tmax[i], roi, grid, res, step, cone, cap,
              i, ri + b, t0 + b, t1 + b);
  }
}

/* ---- occupancy grid: NA/nerfacc/grid.py:196-239 -------------------------- */
/* x = ((coord + jitter)/res) * (max - min) + min  (grid.py:214-228, helpers_contraction.h:23-28) */
void or_occ_cell_points(const int64_t* idx, int32_t n, const float* jit, int32_t rx, int32_t ry,
                        int32_t rz, const float* roi, float* x) {
  const float res[3] = {(float)rx, (float)ry, (float)rz};
  for (int32_t i = 0; i < n; ++i) {
    const int64_t c = idx[i];
    const int64_t ryz = (int64_t)ry * rz;
    const float coord[3] = {(float)(c / ryz), (float)((c / rz) % ry), (float)(c % rz)};
    for (int d = 0; d < 3; ++d) {
      const float u = (coord[d] + jit[3 * i + d]) / res[d];
      x[3 * i + d] = u * (roi[3 + d] - roi[d]) + roi[d];
    }
  }
}

/* occs[c] = max(decay*occs[c], max occ over samples hitting c) for sampled c (grid.py:232,
 * as the scatter-max its comment at 233-236 intends). */
void or_occ_ema_max(float* occs, int32_t n_cells, const int64_t* idx, const float* occ, int32_t n,
                    float decay) {
  uint8_t* seen = (uint8_t*)calloc((size_t)n_cells, 1);
  float* mx = (float*)malloc((size_t)n_cells * sizeof(float));
  for (int32_t i = 0; i < n; ++i) {
    const int64_t c = idx[i];
    if (c < 0 || c >= n_cells) continue;
    if (!seen[c]) {
      mx[c] = occ[i];
      seen[c] = 1;
    } else if (mx[c] == mx[c] && (occ[i] != occ[i] || occ[i] > mx[c])) {
      mx[c] = occ[i]; /* NaN sticks once seen, like torch.maximum */
    }
  }
  for (int32_t c = 0; c < n_cells; ++c) {
    if (!seen[c]) continue;
    const float a = occs[c] * decay, b = mx[c];
    occs[c] = (a != a || b != b) ? NAN : (a > b ? a : b);
  }
  free(seen);
  free(mx);
}

/* compensated (Kahan) fp32 accumulation: s + x with the running error c */
static void kahan_add(float* s, float* c, float x) {
  const float y = x - *c;
  const float t = *s + y;
  *c = (t - *s) - y;
  *s = t;
}

/* binary = occs > min(mean(occs), thre) (grid.py:237-239); mean = compensated
 * fp32 sum over 1024-cell chunks, each sequential, chunk sums summed
 * sequentially (compensated too), divided by n_cells in fp32. */
float or_occ_binarize(const float* occs, int32_t n_cells, float thre, uint8_t* binary) {
  float s = 0.f, sc = 0.f;
  for (int64_t b = 0; b < n_cells; b += 1024) {
    float p = 0.f, pc = 0.f;
    int64_t e = b + 1024 < n_cells ? b + 1024 : n_cells;
    for (int64_t c = b; c < e; ++c) kahan_add(&p, &pc, occs[c]);
    kahan_add(&s, &sc, p);
  }
  const float mean = s / (float)n_cells;
  const float t = fminf(mean, thre);
  for (int32_t c = 0; c < n_cells; ++c) binary[c] = occs[c] > t;
  return t;
}

/* ---- hash-grid corner selection (integer part of tcnn's GridEncoding) ----- */
/* For each point and level: 8 corner table indices (level-local) and the fp32
 * fractional position, computed with fmaf exactly like the device code.  The
 * floating-point interpolation is done by the fp64 torch oracle. */
void or_hashgrid_corners(const float* x, int64_t n, int32_t n_levels, const float* scale,
                         const uint32_t* res, const uint32_t* hsize, const uint8_t* dense,
                         uint32_t* idx_out /* [n, L, 8] */, float* frac_out /* [n, L, 3] */,
                         float* pos_out /* [n, L, 3] (nullable) */) {
  for (int64_t i = 0; i < n; ++i) {
    for (int32_t l = 0; l < n_levels; ++l) {
      uint32_t g[3];
      for (int k = 0; k < 3; ++k) {
        const float p = fmaf(scale[l], x[3 * i + k], 0.5f);
        const float fl = floorf(p);
        g[k] = (uint32_t)(int32_t)fl;
        frac_out[(i * n_levels + l) * 3 + k] = p - fl;
        if (pos_out) pos_out[(i * n_levels + l) * 3 + k] = fl;
      }
      for (int c = 0; c < 8; ++c) {
        const uint32_t gx = g[0] + (c & 1), gy = g[1] + ((c >> 1) & 1), gz = g[2] + ((c >> 2) & 1);
        uint32_t id;
        if (dense[l])
          id = (gx + gy * res[l] + gz * res[l] * res[l]) % hsize[l];
        else
          id = (gx ^ (gy * 2654435761u) ^ (gz * 805459861u)) % hsize[l];
        idx_out[(i * n_levels + l) * 8 + c] = id;
      }
    }
  }
}
