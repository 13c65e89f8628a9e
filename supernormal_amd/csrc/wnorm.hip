// Weight-normalised MLP weights packed for the MLP kernels, and the backward
// of that packing, on the device (fields.py:69-70 wraps each nn.Linear in
// torch.nn.utils.weight_norm, dim 0: W[r] = g[r] * v[r] / ||v[r]||), plus the
// variance's inv_s = clip(exp(10 * variance), 1e-6, 1e6) (fields.py:133-139).
//
// Packed layout (as pack_mlp): per layer W [out, in] row-major, then b [out].
// Parameters and their gradients live in one flat buffer each; a layer
// names its tensors by offsets into them (v, g, b; g < 0: no weight norm,
// W = v).  One wavefront per weight row: the row's squared norm is a
// wavefront sum; the backward is
//   dL/dg[r]  += sum_i dW[r,i] v[r,i] / n_r
//   dL/dv[r,i] += (g[r] / n_r) * (dW[r,i] - v[r,i] * (sum_j dW[r,j] v[r,j]) / n_r^2)
//   dL/db     += dL/db_packed,   dL/dvariance += dL/dinv_s * 10 * inv_s (unclipped).
#include "sn_common.h"

namespace sn {

constexpr int kMaxPackLayers = 8;

struct PackLayers {
  int32_t n;
  int64_t v[kMaxPackLayers], g[kMaxPackLayers], b[kMaxPackLayers], packed[kMaxPackLayers];
  int32_t out[kMaxPackLayers], in[kMaxPackLayers];
  int32_t row0[kMaxPackLayers + 1];  // first global row of each layer
};

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int layer_of(const PackLayers& L, int row) {
  int l = 0;
  while (l + 1 < L.n && L.row0[l + 1] <= row) ++l;
  return l;
}

__global__ void __launch_bounds__(256)
mlp_pack_fwd_kernel(const float* __restrict__ flat, PackLayers L, float* __restrict__ packed,
                    const float* __restrict__ variance, float* __restrict__ inv_s) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (__builtin_amdgcn_readfirstlane(threadIdx.x) >> 6);
  if (row == 0 && lane == 0 && variance && inv_s)
    *inv_s = fminf(fmaxf(expf(*variance * 10.f), 1e-6f), 1e6f);
  if (row >= L.row0[L.n]) return;
  const int l = layer_of(L, row), r = row - L.row0[l], in = L.in[l];
  const float* v = flat + L.v[l] + (int64_t)r * in;
  float* W = packed + L.packed[l] + (int64_t)r * in;
  if (L.g[l] >= 0) {
    float ss = 0.f;
    for (int i = lane; i < in; i += 64) ss += v[i] * v[i];
    const float nrm = sqrtf(wave_sum_f(ss));
    const float g = flat[L.g[l] + r];
    for (int i = lane; i < in; i += 64) W[i] = g * v[i] / nrm;
  } else {
    for (int i = lane; i < in; i += 64) W[i] = v[i];
  }
  if (lane == 0) packed[L.packed[l] + (int64_t)L.out[l] * in + r] = flat[L.b[l] + r];
}

__global__ void __launch_bounds__(256)
mlp_pack_bwd_kernel(const float* __restrict__ flat, PackLayers L,
                    const float* __restrict__ g_packed, float* __restrict__ grad,
                    const float* __restrict__ variance, const float* __restrict__ g_inv_s,
                    float* __restrict__ g_variance) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (__builtin_amdgcn_readfirstlane(threadIdx.x) >> 6);
  if (row == 0 && lane == 0 && variance && g_inv_s && g_variance) {
    const float e = expf(*variance * 10.f);
    if (e >= 1e-6f && e <= 1e6f) *g_variance += *g_inv_s * 10.f * e;
  }
  if (row >= L.row0[L.n]) return;
  const int l = layer_of(L, row), r = row - L.row0[l], in = L.in[l];
  const float* v = flat + L.v[l] + (int64_t)r * in;
  const float* dW = g_packed + L.packed[l] + (int64_t)r * in;
  float* dv = grad + L.v[l] + (int64_t)r * in;
  if (L.g[l] >= 0) {
    float ss = 0.f, dot = 0.f;
    for (int i = lane; i < in; i += 64) {
      ss += v[i] * v[i];
      dot += dW[i] * v[i];
    }
    ss = wave_sum_f(ss);
    dot = wave_sum_f(dot);
    const float nrm = sqrtf(ss);
    const float g = flat[L.g[l] + r];
    const float a = g / nrm, c = dot / ss;
    for (int i = lane; i < in; i += 64) dv[i] += a * (dW[i] - v[i] * c);
    if (lane == 0) grad[L.g[l] + r] += dot / nrm;
  } else {
    for (int i = lane; i < in; i += 64) dv[i] += dW[i];
  }
  if (lane == 0) grad[L.b[l] + r] += g_packed[L.packed[l] + (int64_t)L.out[l] * in + r];
}

static int make_layers(const char* who, int32_t n_layers, const int64_t* desc, PackLayers* L) {
  SN_REQUIRE(n_layers > 0 && n_layers <= kMaxPackLayers, "%s: 1..%d layers", who,
             kMaxPackLayers);
  SN_REQUIRE(desc, "%s: null layer table", who);
  L->n = n_layers;
  int32_t rows = 0;
  for (int l = 0; l < n_layers; ++l) {
    const int64_t* e = desc + 6 * l;  // v, g, b, packed, out, in
    SN_REQUIRE(e[0] >= 0 && e[2] >= 0 && e[3] >= 0 && e[4] > 0 && e[5] > 0,
               "%s: bad layer %d", who, l);
    L->v[l] = e[0];
    L->g[l] = e[1];
    L->b[l] = e[2];
    L->packed[l] = e[3];
    L->out[l] = (int32_t)e[4];
    L->in[l] = (int32_t)e[5];
    L->row0[l] = rows;
    rows += (int32_t)e[4];
  }
  L->row0[n_layers] = rows;
  return SN_OK;
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_mlp_pack_fwd(const float* flat, int32_t n_layers, const int64_t* layers, float* packed,
                    const float* variance, float* inv_s, void* stream) {
  clear_error();
  PackLayers L;
  int rc = make_layers("sn_mlp_pack_fwd", n_layers, layers, &L);
  if (rc) return rc;
  SN_REQUIRE(flat && packed, "sn_mlp_pack_fwd: null pointer");
  const int rows = L.row0[L.n];
  mlp_pack_fwd_kernel<<<(unsigned)ceil_div(rows, 4), 256, 0, as_stream(stream)>>>(
      flat, L, packed, variance, inv_s);
  return check_launch("sn_mlp_pack_fwd");
}

int sn_mlp_pack_bwd(const float* flat, int32_t n_layers, const int64_t* layers,
                    const float* g_packed, float* grad, const float* variance,
                    const float* g_inv_s, float* g_variance, void* stream) {
  clear_error();
  PackLayers L;
  int rc = make_layers("sn_mlp_pack_bwd", n_layers, layers, &L);
  if (rc) return rc;
  SN_REQUIRE(flat && g_packed && grad, "sn_mlp_pack_bwd: null pointer");
  const int rows = L.row0[L.n];
  mlp_pack_bwd_kernel<<<(unsigned)ceil_div(rows, 4), 256, 0, as_stream(stream)>>>(
      flat, L, g_packed, grad, variance, g_inv_s, g_variance);
  return check_launch("sn_mlp_pack_bwd");
}

}  // extern "C"
