// Multiresolution hash-grid encoding, 3-D input, linear interpolation.
//
// Semantics follow tiny-cuda-nn's GridEncoding (commit 2ec562e8, the version
// pinned by the reference's create_env.sh:11; its source is not vendored in
// the reference) as restated from the Instant-NGP method:
//   scale_l = base * per_level_scale^l - 1      (computed once on the host, fp64 -> fp32)
//   res_l   = ceil(scale_l) + 1
//   size_l  = min(next_multiple(res_l^3, 8), 2^log2_hashmap_size)
//   pos     = fmaf(scale_l, x, 0.5); cell = floor(pos); frac = pos - cell
//   corner index: dense (x + y*res + z*res^2) when res^3 <= size_l, else
//   x ^ y*2654435761 ^ z*805459861 (uint32 wrap), modulo size_l.
// Cells of negative inputs wrap through uint32, which is what the reference
// gets when it feeds [-1, 1] positions (renderer.py:44) to tcnn.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sn {

constexpr int kMaxLevels = 32;

struct HashGridDesc {
  int32_t n_levels;
  int32_t n_active;  // levels >= n_active are masked (output 0, no gradient)
  // optional device copy of the active level count (read by the kernels,
  // bounded by n_active): lets a captured launch follow a changing count
  const int32_t* n_active_dev;
  uint32_t offset[kMaxLevels + 1];  // table entry offsets (not floats)
  uint32_t hsize[kMaxLevels];
  uint32_t res[kMaxLevels];
  float scale[kMaxLevels];
  uint32_t dense_mask;
};

// Host: fills the descriptor; returns total entries or -1 on bad arguments.
int64_t make_hashgrid_desc(int32_t n_levels, int32_t log2_T, int32_t base_res,
                           float per_level_scale, int32_t n_active, HashGridDesc* d);

#ifdef __HIPCC__
// Corner indices/weights of one level.  frac in [0,1); w[c] in corner order
// c = bx | by<<1 | bz<<2, computed as ((w_x) * w_y) * w_z like the reference loop;
// g = the cell's integer coordinates (corner 0).
__device__ __forceinline__ void hg_level(const float x[3], const HashGridDesc& d, int l,
                                         uint32_t idx[8], float w[8], float frac[3],
                                         uint32_t g[3]) {
  const float sc = d.scale[l];
  const uint32_t res = d.res[l];
  const uint32_t hs = d.hsize[l];
  const bool dense = (d.dense_mask >> l) & 1u;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float p = fmaf(sc, x[k], 0.5f);
    float fl = floorf(p);
    g[k] = (uint32_t)(int32_t)fl;
    frac[k] = p - fl;
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t gx = g[0] + (c & 1), gy = g[1] + ((c >> 1) & 1), gz = g[2] + ((c >> 2) & 1);
    uint32_t id;
    if (dense) {
      id = gx + gy * res + gz * res * res;
      id = id < hs ? id : id % hs;
    } else {
      id = (gx ^ (gy * 2654435761u) ^ (gz * 805459861u)) & (hs - 1u);
    }
    idx[c] = id;
    float wt = (c & 1) ? frac[0] : 1.f - frac[0];
    wt *= ((c >> 1) & 1) ? frac[1] : 1.f - frac[1];
    wt *= ((c >> 2) & 1) ? frac[2] : 1.f - frac[2];
    w[c] = wt;
  }
}

__device__ __forceinline__ void hg_level(const float x[3], const HashGridDesc& d, int l,
                                         uint32_t idx[8], float w[8], float frac[3]) {
  uint32_t g[3];
  hg_level(x, d, l, idx, w, frac, g);
}

// Active levels seen by a kernel.
__device__ __forceinline__ int active_levels(const HashGridDesc& d) {
  if (!d.n_active_dev) return d.n_active;
  const int v = *d.n_active_dev;
  return v < 0 ? 0 : (v < d.n_active ? v : d.n_active);
}

// d w[c] / d frac[k]
__device__ __forceinline__ void hg_dweights(const float frac[3], float dw[8][3]) {
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float a0 = (c & 1) ? frac[0] : 1.f - frac[0];
    const float a1 = ((c >> 1) & 1) ? frac[1] : 1.f - frac[1];
    const float a2 = ((c >> 2) & 1) ? frac[2] : 1.f - frac[2];
    const float s0 = (c & 1) ? 1.f : -1.f;
    const float s1 = ((c >> 1) & 1) ? 1.f : -1.f;
    const float s2 = ((c >> 2) & 1) ? 1.f : -1.f;
    dw[c][0] = s0 * a1 * a2;
    dw[c][1] = a0 * s1 * a2;
    dw[c][2] = a0 * a1 * s2;
  }
}
#endif

}  // namespace sn
