// Patch sampling of the training loop (dataset_loader.py:223-297): given the
// random patch centres and views, gathers each pixel's supervision (normal,
// V^-1, mask), builds its ray (o = camera centre, d = R K^-1 [x y 1] /
// ||K^-1 [x y 1]||), the patch's marching-plane normal (3rd column of the
// camera rotation), and the centre ray's near/far on the unit sphere with the
// stratified start near + u * step (renderer.py:69-71).  The random numbers
// come in (the caller draws them from its generator, as the reference does
// with torch.randint / np.random.choice).
//
// Pixel (b, ky, kx) of a patch: x = cx[b] + (-Pw // 2 + 1) + kx, y likewise
// (the reference's arange(-Pw//2 + 1, Pw//2 + 1) with Python's floor
// division).  Indices are clamped into the image and the view table, so a
// bad random number cannot read out of bounds.
#include "sn_common.h"

namespace sn {

struct SampleArgs {
  const int64_t* cx;
  const int64_t* cy;
  const int64_t* sel;       // index into train_idx
  const float* u;           // [B] stratified jitter in [0,1) or null
  const float* step;        // device step size (with u)
  const int64_t* train_idx;
  const float* normals;     // [n, H, W, 3]
  const float* v_inv;       // [n, H, W, 3, 3]
  const float* masks;       // [n, H, W]
  const float* k_inv;       // [n, ks, ks] (ks = 3 or 4; the top-left 3x3 is used)
  const float* pose;        // [n, 4, 4]
  int32_t H, W, B, Ph, Pw, mask_mode, n_img, n_train, ks;
  float* rays_o;            // [B, Ph, Pw, 3]
  float* rays_d;
  float* plane_n;           // [B, 3]
  float* v_out;             // [B, Ph, Pw, 3, 3]
  float* n_out;             // [B, Ph, Pw, 3]
  float* mask_out;          // [B, Ph * Pw]
  float* o_c;               // [B, 3] centre pixel's ray
  float* d_c;
  float* t_min;             // [B]
  float* t_max;
};

__device__ __forceinline__ int py_floordiv(int a, int b) {  // b > 0
  return a >= 0 ? a / b : -((-a + b - 1) / b);
}

__global__ void __launch_bounds__(256) patch_sample_kernel(SampleArgs a) {
  const int P = a.Ph * a.Pw;
  const int64_t total = (int64_t)a.B * P;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(t / P), k = (int)(t - (int64_t)b * P);
    const int ky = k / a.Pw, kx = k - ky * a.Pw;
    const int64_t si = min(max(a.sel[b], (int64_t)0), (int64_t)a.n_train - 1);
    const int64_t img = min(max(a.train_idx[si], (int64_t)0), (int64_t)a.n_img - 1);
    const int x = (int)min(max(a.cx[b] + py_floordiv(-a.Pw, 2) + 1 + kx, (int64_t)0),
                           (int64_t)a.W - 1);
    const int y = (int)min(max(a.cy[b] + py_floordiv(-a.Ph, 2) + 1 + ky, (int64_t)0),
                           (int64_t)a.H - 1);
    const int64_t pix = (img * a.H + y) * a.W + x;
#pragma unroll
    for (int c = 0; c < 3; ++c) a.n_out[t * 3 + c] = a.normals[pix * 3 + c];
#pragma unroll
    for (int c = 0; c < 9; ++c) a.v_out[t * 9 + c] = a.v_inv[pix * 9 + c];
    const float m = a.masks[pix];
    a.mask_out[t] = a.mask_mode ? (m > 0.5f ? 1.f : 0.f) : 1.f;
    // p = K^-1 [x, y, 1], d = p / ||p||, d = R d
    const float* Ki = a.k_inv + img * a.ks * a.ks;
    const float* R = a.pose + img * 16;
    const float fx = (float)x, fy = (float)y;
    float p[3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
      p[r] = (Ki[a.ks * r] * fx + Ki[a.ks * r + 1] * fy) + Ki[a.ks * r + 2];
    const float nrm = sqrtf((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]);
    float d[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) d[r] = p[r] / nrm;
    float dw[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) dw[r] = (R[4 * r] * d[0] + R[4 * r + 1] * d[1]) + R[4 * r + 2] * d[2];
    const float o[3] = {R[3], R[7], R[11]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      a.rays_d[t * 3 + c] = dw[c];
      a.rays_o[t * 3 + c] = o[c];
    }
    if (k == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) a.plane_n[b * 3 + c] = R[4 * c + 2];
    }
    if (ky == a.Ph / 2 && kx == a.Pw / 2) {
      // the centre ray: near / far on the unit sphere (dataset_loader.py:279-297)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        a.o_c[b * 3 + c] = o[c];
        a.d_c[b * 3 + c] = dw[c];
      }
      const float aa = (dw[0] * dw[0] + dw[1] * dw[1]) + dw[2] * dw[2];
      const float bb = 2.f * ((o[0] * dw[0] + o[1] * dw[1]) + o[2] * dw[2]);
      const float cc = ((o[0] * o[0] + o[1] * o[1]) + o[2] * o[2]) - 1.f;
      const float mid = 0.5f * (-bb) / aa;
      const float q = sqrtf(bb * bb - 4.f * aa * cc) / (2.f * aa);
      float near = mid - q;
      if (a.u) near = near + a.u[b] * *a.step;
      a.t_min[b] = near;
      a.t_max[b] = mid + q;
    }
  }
}

// sum of the mask in a fixed order (one workgroup): per-thread contiguous
// runs, then a fixed tree; out = sum + eps
__global__ void __launch_bounds__(1024)
mask_sum_kernel(const float* __restrict__ m, int64_t n, float eps, float* __restrict__ out) {
  __shared__ float part[1024];
  const int tid = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  float s = 0.f;
  for (int64_t i = tid * per; i < min(n, (int64_t)(tid + 1) * per); ++i) s += m[i];
  part[tid] = s;
  __syncthreads();
  for (int w = 512; w >= 1; w >>= 1) {
    if (tid < w) part[tid] += part[tid + w];
    __syncthreads();
  }
  if (tid == 0) out[0] = part[0] + eps;
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_patch_sample(const int64_t* cx, const int64_t* cy, const int64_t* sel, const float* u,
                    const float* step_dev, const int64_t* train_idx, const float* normals,
                    const float* v_inv_all, const float* masks, const float* k_inv,
                    int32_t k_stride, const float* pose, int32_t n_img, int32_t n_train,
                    int32_t H, int32_t W,
                    int32_t B, int32_t patch_h,
                    int32_t patch_w, int32_t mask_mode, float* rays_o, float* rays_d,
                    float* plane_n, float* v_inv, float* normals_out, float* mask,
                    float* mask_sum, float* o_c, float* d_c, float* t_min, float* t_max,
                    void* stream) {
  clear_error();
  SN_REQUIRE(B >= 0 && H > 0 && W > 0 && patch_h > 0 && patch_w > 0 && n_img > 0 &&
                 n_train > 0,
             "sn_patch_sample: bad sizes");
  SN_REQUIRE(patch_w <= W && patch_h <= H, "sn_patch_sample: patch larger than the image");
  SN_REQUIRE(k_stride == 3 || k_stride == 4, "sn_patch_sample: K^-1 must be 3x3 or 4x4");
  if (B == 0) return SN_OK;
  SN_REQUIRE(cx && cy && sel && train_idx && normals && v_inv_all && masks && k_inv && pose &&
                 rays_o && rays_d && plane_n && v_inv && normals_out && mask && mask_sum &&
                 o_c && d_c && t_min && t_max && (!u || step_dev),
             "sn_patch_sample: null pointer");
  SampleArgs a{cx, cy, sel, u, step_dev, train_idx, normals, v_inv_all, masks, k_inv, pose,
               H, W, B, patch_h, patch_w, mask_mode, n_img, n_train, k_stride, rays_o, rays_d, plane_n, v_inv,
               normals_out, mask, o_c, d_c, t_min, t_max};
  hipStream_t s = as_stream(stream);
  const int64_t total = (int64_t)B * patch_h * patch_w;
  patch_sample_kernel<<<grid_for(total, 256), 256, 0, s>>>(a);
  int rc = check_launch("sn_patch_sample");
  if (rc) return rc;
  mask_sum_kernel<<<1, 1024, 0, s>>>(mask, total, 1e-5f, mask_sum);
  return check_launch("sn_patch_sample");
}

}  // extern "C"
