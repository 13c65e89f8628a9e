// Patch-based render weights / transmittance, per-ray visibility and the
// scatter-add used by accumulate_along_rays_patch_based.
//
// One lane per (patch, ray-in-patch): lane g -> patch g / P, ray k = g % P,
// so the P lanes of a patch read P consecutive floats of each sample row and
// a 64-lane wavefront covers ~7 patches.  The recurrence along a patch's
// samples is sequential per lane, in the same order as the reference loop
// (csrc/render_weight.cu:108-116), so forward results are bit-identical to
// that loop.  Built with -ffp-contract=off.
#include "sn_common.h"

namespace sn {

__global__ void __launch_bounds__(256)
patch_weight_fwd_kernel(const int32_t* __restrict__ packed_info, int32_t n_patches, int32_t P,
                        int64_t n_samples, const float* __restrict__ alphas, float* __restrict__ weights,
                        float* __restrict__ trans) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= (int64_t)n_patches * P) return;
  const int32_t p = (int32_t)(g / P);
  const int32_t k = (int32_t)(g - (int64_t)p * P);
  const int32_t base = packed_info[2 * p];
  const int32_t steps = packed_info[2 * p + 1];
  if (steps <= 0 || base < 0 || (int64_t)base + steps > n_samples) return;
  const int64_t off = (int64_t)base * P + k;
  float T = 1.f;
  for (int32_t j = 0; j < steps; ++j) {
    const int64_t idx = off + (int64_t)j * P;
    const float a = alphas[idx];
    if (trans) trans[idx] = T;
    if (weights) weights[idx] = a * T;
    T *= (1.f - a);
  }
}

// g_alpha[j] = T[j] * (gw[j] - Q[j]),
// Q[j] = (gw[j+1] a[j+1] + gT[j+1]) + (1 - a[j+1]) Q[j+1], Q[last] = 0.
// T[j] is first written into grad_alphas (pass 1), then scaled in place.
__global__ void __launch_bounds__(256)
patch_weight_bwd_kernel(const int32_t* __restrict__ packed_info, int32_t n_patches, int32_t P,
                        int64_t n_samples, const float* __restrict__ alphas, const float* __restrict__ gw,
                        const float* __restrict__ gT, float* __restrict__ ga) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= (int64_t)n_patches * P) return;
  const int32_t p = (int32_t)(g / P);
  const int32_t k = (int32_t)(g - (int64_t)p * P);
  const int32_t base = packed_info[2 * p];
  const int32_t steps = packed_info[2 * p + 1];
  if (steps <= 0 || base < 0 || (int64_t)base + steps > n_samples) return;
  const int64_t off = (int64_t)base * P + k;
  float T = 1.f;
  for (int32_t j = 0; j < steps; ++j) {
    const int64_t idx = off + (int64_t)j * P;
    ga[idx] = T;
    T *= (1.f - alphas[idx]);
  }
  float Q = 0.f;
  for (int32_t j = steps - 1; j >= 0; --j) {
    const int64_t idx = off + (int64_t)j * P;
    const float a = alphas[idx];
    const float gwj = gw ? gw[idx] : 0.f;
    const float gTj = gT ? gT[idx] : 0.f;
    ga[idx] = ga[idx] * (gwj - Q);
    Q = (gwj * a + gTj) + (1.f - a) * Q;
  }
}

// Per-ray exclusive product (csrc/render_transmittance.cu:85-112) and the
// visibility test of vol_rendering.py:744-747.
__global__ void __launch_bounds__(256)
visibility_kernel(const int32_t* __restrict__ packed_info, int32_t n_rays, int64_t n_samples,
                  const float* __restrict__ alphas, float eps, float alpha_thre,
                  uint8_t* __restrict__ keep) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rays) return;
  const int32_t base = packed_info[2 * r];
  const int32_t steps = packed_info[2 * r + 1];
  if (steps <= 0 || base < 0 || (int64_t)base + steps > n_samples) return;
  float T = 1.f;
  for (int32_t j = 0; j < steps; ++j) {
    const float a = alphas[base + j];
    bool v = T >= eps;
    if (alpha_thre > 0.f) v = v && (a >= alpha_thre);
    keep[base + j] = v ? 1 : 0;
    T *= (1.f - a);
  }
}

__global__ void __launch_bounds__(256)
scatter_add_rows_kernel(const int64_t* __restrict__ idx, int64_t n, int32_t row,
                        const float* __restrict__ src, float* __restrict__ out,
                        int32_t n_out_rows) {
  const int64_t total = n * row;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = g / row;
    const int32_t e = (int32_t)(g - s * row);
    const int64_t r = idx[s];
    if (r >= 0 && r < n_out_rows) atomicAdd(&out[r * row + e], src[g]);
  }
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_patch_weight_fwd(const int32_t* packed_info, int32_t n_patches, int32_t patch_size,
                        int64_t n_samples, const float* alphas, float* weights,
                        float* transmittance, void* stream) {
  clear_error();
  SN_REQUIRE(n_patches >= 0 && patch_size > 0 && n_samples >= 0,
             "sn_patch_weight_fwd: bad sizes (%d, %d)", n_patches, patch_size);
  if (n_patches == 0 || n_samples == 0) return SN_OK;
  SN_REQUIRE(packed_info && alphas, "sn_patch_weight_fwd: null input");
  SN_REQUIRE(weights || transmittance, "sn_patch_weight_fwd: no output");
  int64_t n = (int64_t)n_patches * patch_size;
  patch_weight_fwd_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, as_stream(stream)>>>(
      packed_info, n_patches, patch_size, n_samples, alphas, weights, transmittance);
  return check_launch("sn_patch_weight_fwd");
}

int sn_patch_transmittance_fwd(const int32_t* packed_info, int32_t n_patches,
                               int32_t patch_size, int64_t n_samples, const float* alphas,
                               float* transmittance, void* stream) {
  return sn_patch_weight_fwd(packed_info, n_patches, patch_size, n_samples, alphas, nullptr,
                             transmittance, stream);
}

int sn_patch_weight_bwd(const int32_t* packed_info, int32_t n_patches, int32_t patch_size,
                        int64_t n_samples, const float* alphas, const float* grad_weights,
                        const float* grad_transmittance, float* grad_alphas, void* stream) {
  clear_error();
  SN_REQUIRE(n_patches >= 0 && patch_size > 0 && n_samples >= 0,
             "sn_patch_weight_bwd: bad sizes (%d, %d)", n_patches, patch_size);
  if (n_patches == 0 || n_samples == 0) return SN_OK;
  SN_REQUIRE(packed_info && alphas && grad_alphas, "sn_patch_weight_bwd: null pointer");
  int64_t n = (int64_t)n_patches * patch_size;
  patch_weight_bwd_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, as_stream(stream)>>>(
      packed_info, n_patches, patch_size, n_samples, alphas, grad_weights, grad_transmittance,
      grad_alphas);
  return check_launch("sn_patch_weight_bwd");
}

int sn_visibility(const int32_t* packed_info, int32_t n_rays, int64_t n_samples,
                  const float* alphas, float early_stop_eps, float alpha_thre, uint8_t* keep,
                  void* stream) {
  clear_error();
  SN_REQUIRE(n_rays >= 0 && n_samples >= 0, "sn_visibility: negative size");
  if (n_rays == 0 || n_samples == 0) return SN_OK;
  SN_REQUIRE(packed_info && alphas && keep, "sn_visibility: null pointer");
  visibility_kernel<<<(unsigned)ceil_div(n_rays, 256), 256, 0, as_stream(stream)>>>(
      packed_info, n_rays, n_samples, alphas, early_stop_eps, alpha_thre, keep);
  return check_launch("sn_visibility");
}

int sn_scatter_add_rows(const int64_t* ray_indices, int64_t n, int32_t row_elems,
                        const float* src, float* out, int32_t n_out_rows, void* stream) {
  clear_error();
  SN_REQUIRE(n >= 0 && row_elems > 0 && n_out_rows >= 0, "sn_scatter_add_rows: bad sizes");
  if (n == 0) return SN_OK;
  SN_REQUIRE(ray_indices && src && out, "sn_scatter_add_rows: null pointer");
  scatter_add_rows_kernel<<<grid_for(n * row_elems, 256), 256, 0, as_stream(stream)>>>(
      ray_indices, n, row_elems, src, out, n_out_rows);
  return check_launch("sn_scatter_add_rows");
}

}  // extern "C"
