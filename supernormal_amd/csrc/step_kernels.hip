// Glue of the sync-free training iteration: every size that depends on the
// data stays on the device (launches are sized by capacities, kernels read
// the valid counts), so an iteration needs no host round trip and can be
// captured as a graph.
//
//   sn_centre_points    eval points of the visibility pre-pass: the starts of
//                       the marched centre samples and the ends that are not
//                       the next sample's start (renderer.py:80-100)
//   sn_alpha_visibility NeuS alpha from those SDF values and the per-ray
//                       transmittance test (renderer.py:102-118,
//                       ray_marching.py:193-215)
//   sn_patch_points     patch expansion onto the marching planes
//                       (renderer.py:142-159): starts of every patch ray and
//                       the non-contiguous ends, laid out [S*P | Nd*P]
//   sn_occ_sample_cells cells of a post-warm-up occupancy update: N/4 uniform
//                       + up to N/4 occupied (grid.py:181-194)
//
// A sample's end is its own point unless the next sample is on the same ray
// and starts exactly there.  Compiled with -ffp-contract=off.
#include "sn_common.h"

namespace sn {

__global__ void __launch_bounds__(256)
diff_flags_kernel(const int64_t* __restrict__ ri, const float* __restrict__ t0,
                  const float* __restrict__ t1, int64_t n_cap, const int32_t* __restrict__ n_dev,
                  uint8_t* __restrict__ flags) {
  const int64_t n = dev_count(n_cap, n_dev);
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    bool d = true;
    if (s + 1 < n) d = !(ri[s + 1] == ri[s] && t1[s] == t0[s + 1]);
    flags[s] = d ? 1 : 0;
  }
}

__global__ void __launch_bounds__(256)
centre_points_kernel(const int64_t* __restrict__ ri, const float* __restrict__ t0,
                     const float* __restrict__ t1, int64_t n_cap,
                     const int32_t* __restrict__ n_dev, const float* __restrict__ o_c,
                     const float* __restrict__ d_c, int32_t c_stride,
                     const int32_t* __restrict__ didx, const int32_t* __restrict__ nd_dev,
                     float* __restrict__ pts, int32_t* __restrict__ n_eval) {
  const int64_t n = dev_count(n_cap, n_dev);
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_eval = (int32_t)n + *nd_dev;
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = ri[s];
    const float* o = o_c + r * c_stride;
    const float* d = d_c + r * c_stride;
    const float a = t0[s];
#pragma unroll
    for (int c = 0; c < 3; ++c) pts[3 * s + c] = o[c] + d[c] * a;
    const int32_t di = didx[s];
    if (di >= 0) {
      const float b = t1[s];
#pragma unroll
      for (int c = 0; c < 3; ++c) pts[3 * (n + di) + c] = o[c] + d[c] * b;
    }
  }
}

__device__ __forceinline__ float sigm(float x) { return 1.f / (1.f + expf(-x)); }

// One wavefront per ray: the lanes evaluate 64 consecutive samples' alphas at
// once, then the transmittance runs as the sequential product T <- T (1 - a)
// in sample order (each factor read from its sample's lane), so the flags are
// those of a one-lane-per-ray loop, bit for bit.
__global__ void __launch_bounds__(256)
alpha_visibility_kernel(const int32_t* __restrict__ packed, int32_t n_rays,
                        const float* __restrict__ sdf, const int32_t* __restrict__ didx,
                        int64_t n_cap, const int32_t* __restrict__ n_dev,
                        const float* __restrict__ inv_s_dev, float eps,
                        uint8_t* __restrict__ keep) {
  const int lane = threadIdx.x & 63;
  // the wavefront's ray (uniform: taken from its first lane)
  const int64_t r = (int64_t)blockIdx.x * 4 + (__builtin_amdgcn_readfirstlane(threadIdx.x) >> 6);
  if (r >= n_rays) return;
  const int64_t n = dev_count(n_cap, n_dev);
  const int32_t base = packed[2 * r], cnt = packed[2 * r + 1];
  if (cnt <= 0 || base < 0 || (int64_t)base + cnt > n) return;
  const float inv_s = *inv_s_dev;
  float T = 1.f;
  for (int32_t j0 = 0; j0 < cnt; j0 += 64) {
    const int32_t j = j0 + lane;
    float om = 1.f;
    if (j < cnt) {
      const int64_t s = base + j;
      const float fs = sdf[s];
      const int32_t di = didx[s];
      const float fe = di >= 0 ? sdf[n + di] : sdf[s + 1];
      const float c = sigm(fs * inv_s), e = sigm(fe * inv_s);
      const float a = fminf(fmaxf(((c - e) + 1e-5f) / (c + 1e-5f), 0.f), 1.f);
      om = 1.f - a;
    }
    const int m = min(64, cnt - j0);
    float before = 1.f;
    for (int k = 0; k < m; ++k) {
      if (lane == k) before = T;
      T *= __int_as_float(__builtin_amdgcn_readlane(__float_as_int(om), k));
    }
    if (j < cnt) keep[base + j] = before >= eps ? 1 : 0;
  }
}

__global__ void __launch_bounds__(256)
patch_points_kernel(const int64_t* __restrict__ ri, const float* __restrict__ t0,
                    const float* __restrict__ t1, int64_t n_cap,
                    const int32_t* __restrict__ n_dev, const float* __restrict__ rays_o,
                    const float* __restrict__ rays_d, const float* __restrict__ plane_n, int32_t P,
                    int32_t kc, const int32_t* __restrict__ didx,
                    const int32_t* __restrict__ nd_dev, float* __restrict__ pts,
                    int32_t* __restrict__ n_eval) {
  const int64_t n = dev_count(n_cap, n_dev);
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_eval = (int32_t)((n + *nd_dev) * P);
  const int64_t total = n * P;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = g / P;
    const int32_t k = (int32_t)(g - s * P);
    const int64_t p = ri[s];
    const float* nn = plane_n + 3 * p;
    const float* dc = rays_d + (p * P + kc) * 3;
    const float* o = rays_o + (p * P + k) * 3;
    const float* d = rays_d + (p * P + k) * 3;
    const float dcn = (dc[0] * nn[0] + dc[1] * nn[1]) + dc[2] * nn[2];
    const float dkn = (d[0] * nn[0] + d[1] * nn[1]) + d[2] * nn[2];
    const float ts = (t0[s] * dcn) / dkn;
#pragma unroll
    for (int c = 0; c < 3; ++c) pts[3 * g + c] = o[c] + d[c] * ts;
    const int32_t di = didx[s];
    if (di >= 0) {
      const float te = (t1[s] * dcn) / dkn;
      const int64_t row = (n + di) * P + k;
#pragma unroll
      for (int c = 0; c < 3; ++c) pts[3 * row + c] = o[c] + d[c] * te;
    }
  }
}

__global__ void __launch_bounds__(256)
occ_sample_kernel(const int64_t* __restrict__ occ_list, const int32_t* __restrict__ occ_count,
                  int64_t n_cells, int64_t quarter, const float* __restrict__ u_uniform,
                  const float* __restrict__ u_select, int64_t* __restrict__ out,
                  int32_t* __restrict__ n_out) {
  const int64_t c = *occ_count;
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_out = (int32_t)(quarter + (c < quarter ? c : quarter));
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < quarter;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t u = (int64_t)(u_uniform[i] * (float)n_cells);
    out[i] = u < n_cells ? u : n_cells - 1;
    if (c > quarter) {
      int64_t j = (int64_t)(u_select[i] * (float)c);
      out[quarter + i] = occ_list[j < c ? j : c - 1];
    } else if (i < c) {
      out[quarter + i] = occ_list[i];
    }
  }
}

static size_t flags_bytes(int64_t n) { return align_up((size_t)(n > 0 ? n : 1), 256); }

}  // namespace sn

using namespace sn;

extern "C" {

size_t sn_step_workspace_bytes(int64_t n_cap) {
  return flags_bytes(n_cap) + align_up(sizeof(int32_t) * 4, 256) + scan_ws_bytes(n_cap);
}

int sn_centre_points(const int64_t* ray_indices, const float* t_starts, const float* t_ends,
                     int64_t n_cap, const int32_t* n_dev, const float* rays_o_c,
                     const float* rays_d_c, int32_t centre_stride, int32_t* diff_idx,
                     int32_t* n_diff, float* points, int32_t* n_eval, void* ws, size_t ws_bytes,
                     void* stream) {
  clear_error();
  SN_REQUIRE(n_cap >= 0 && centre_stride >= 3, "sn_centre_points: bad sizes");
  SN_REQUIRE(n_dev && n_diff && n_eval, "sn_centre_points: device counts required");
  if (ws_bytes < sn_step_workspace_bytes(n_cap)) {
    set_error("sn_centre_points: workspace %zu < %zu", ws_bytes, sn_step_workspace_bytes(n_cap));
    return SN_ESPACE;
  }
  hipStream_t s = as_stream(stream);
  if (n_cap == 0) {
    int rc = fill_i32(n_diff, 1, 0, s);
    return rc ? rc : fill_i32(n_eval, 1, 0, s);
  }
  uint8_t* flags = reinterpret_cast<uint8_t*>(ws);
  char* rest = (char*)ws + flags_bytes(n_cap) + align_up(sizeof(int32_t) * 4, 256);
  diff_flags_kernel<<<grid_for(n_cap, 256), 256, 0, s>>>(ray_indices, t_starts, t_ends, n_cap,
                                                         n_dev, flags);
  int rc = scan_flags_dev(flags, n_cap, n_dev, diff_idx, n_diff, rest,
                          ws_bytes - (rest - (char*)ws), s);
  if (rc) return rc;
  centre_points_kernel<<<grid_for(n_cap, 256), 256, 0, s>>>(
      ray_indices, t_starts, t_ends, n_cap, n_dev, rays_o_c, rays_d_c, centre_stride, diff_idx,
      n_diff, points, n_eval);
  return check_launch("sn_centre_points");
}

int sn_alpha_visibility(const int32_t* packed_info, int32_t n_rays, const float* sdf,
                        const int32_t* diff_idx, int64_t n_cap, const int32_t* n_dev,
                        const float* inv_s, float early_stop_eps, uint8_t* keep, void* stream) {
  clear_error();
  SN_REQUIRE(n_rays >= 0 && n_cap >= 0, "sn_alpha_visibility: bad sizes");
  if (n_rays == 0 || n_cap == 0) return SN_OK;
  SN_REQUIRE(packed_info && sdf && diff_idx && inv_s && keep, "sn_alpha_visibility: null");
  alpha_visibility_kernel<<<(unsigned)ceil_div((int64_t)n_rays * 64, 256), 256, 0,
                            as_stream(stream)>>>(packed_info, n_rays, sdf, diff_idx, n_cap, n_dev,
                                                 inv_s, early_stop_eps, keep);
  return check_launch("sn_alpha_visibility");
}

int sn_patch_points(const int64_t* ray_indices, const float* t_starts, const float* t_ends,
                    int64_t n_cap, const int32_t* n_dev, const float* rays_o,
                    const float* rays_d, const float* plane_n, int32_t patch_h, int32_t patch_w,
                    int32_t* diff_idx, int32_t* n_diff, float* points, int32_t* n_eval,
                    void* ws, size_t ws_bytes, void* stream) {
  clear_error();
  const int32_t P = patch_h * patch_w;
  SN_REQUIRE(n_cap >= 0 && patch_h > 0 && patch_w > 0, "sn_patch_points: bad sizes");
  SN_REQUIRE(n_dev && n_diff && n_eval, "sn_patch_points: device counts required");
  if (ws_bytes < sn_step_workspace_bytes(n_cap)) {
    set_error("sn_patch_points: workspace %zu < %zu", ws_bytes, sn_step_workspace_bytes(n_cap));
    return SN_ESPACE;
  }
  hipStream_t s = as_stream(stream);
  if (n_cap == 0) {
    int rc = fill_i32(n_diff, 1, 0, s);
    return rc ? rc : fill_i32(n_eval, 1, 0, s);
  }
  uint8_t* flags = reinterpret_cast<uint8_t*>(ws);
  char* rest = (char*)ws + flags_bytes(n_cap) + align_up(sizeof(int32_t) * 4, 256);
  diff_flags_kernel<<<grid_for(n_cap, 256), 256, 0, s>>>(ray_indices, t_starts, t_ends, n_cap,
                                                         n_dev, flags);
  int rc = scan_flags_dev(flags, n_cap, n_dev, diff_idx, n_diff, rest,
                          ws_bytes - (rest - (char*)ws), s);
  if (rc) return rc;
  const int32_t kc = (patch_h / 2) * patch_w + patch_w / 2;
  patch_points_kernel<<<grid_for(n_cap * P, 256), 256, 0, s>>>(
      ray_indices, t_starts, t_ends, n_cap, n_dev, rays_o, rays_d, plane_n, P, kc, diff_idx,
      n_diff, points, n_eval);
  return check_launch("sn_patch_points");
}

int sn_occ_sample_cells(const uint8_t* binary, int32_t n_cells, const float* u_uniform,
                        const float* u_select, int64_t* out_idx, int32_t* n_out, void* ws,
                        size_t ws_bytes, void* stream) {
  clear_error();
  SN_REQUIRE(n_cells > 0, "sn_occ_sample_cells: n_cells <= 0");
  SN_REQUIRE(binary && u_uniform && u_select && out_idx && n_out, "sn_occ_sample_cells: null");
  const size_t list_b = align_up((size_t)n_cells * sizeof(int64_t), 256);
  const size_t cnt_b = 256;
  const size_t need = list_b + cnt_b + align_up((size_t)n_cells * sizeof(int32_t), 256) +
                      scan_ws_bytes(n_cells);
  if (ws_bytes < need) {
    set_error("sn_occ_sample_cells: workspace %zu < %zu", ws_bytes, need);
    return SN_ESPACE;
  }
  hipStream_t s = as_stream(stream);
  int64_t* occ_list = reinterpret_cast<int64_t*>(ws);
  int32_t* occ_cnt = reinterpret_cast<int32_t*>((char*)ws + list_b);
  int rc = sn_occ_nonzero(binary, n_cells, occ_list, occ_cnt, (char*)ws + list_b + cnt_b,
                          ws_bytes - list_b - cnt_b, stream);
  if (rc) return rc;
  const int64_t quarter = n_cells / 4;
  occ_sample_kernel<<<grid_for(quarter, 256), 256, 0, s>>>(occ_list, occ_cnt, n_cells, quarter,
                                                           u_uniform, u_select, out_idx, n_out);
  return check_launch("sn_occ_sample_cells");
}

size_t sn_occ_sample_workspace_bytes(int32_t n_cells) {
  return align_up((size_t)(n_cells > 0 ? n_cells : 1) * sizeof(int64_t), 256) + 256 +
         align_up((size_t)(n_cells > 0 ? n_cells : 1) * sizeof(int32_t), 256) +
         scan_ws_bytes(n_cells);
}

}  // extern "C"
