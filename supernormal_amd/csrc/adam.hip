// Adam over one flat fp32 parameter buffer (torch.optim.Adam defaults used by
// exp_runner.py:97, 205-207).  The buffer holds every trainable tensor of the
// SDF (hash table, MLP, variance) back to back, so one launch updates all of
// them; 28 bytes of HBM traffic per parameter (p, g, m, v read; p, m, v
// written), float4-vectorised.
#include <cmath>

#include "sn_common.h"

namespace sn {

// 1 - beta is formed in double on the host: in fp32, 1 - 0.999f is off by
// 1.3e-5 relative, which torch (Python floats) does not do.  The bias
// corrections of a step count held on the device are 1 - beta^t =
// -expm1(t ln beta) in fp32 (ln beta from the host's double): accurate for
// small t, where 1 - exp would cancel.
struct AdamK {
  float b1, b2, omb1, omb2, eps, step_size, inv_sqrt_bc2;
  float ln_b1, ln_b2;
};

__device__ __forceinline__ void adam1(float& p, float g, float& m, float& v, const AdamK& k) {
  m = k.b1 * m + k.omb1 * g;
  v = k.b2 * v + k.omb2 * g * g;
  const float denom = sqrtf(v) * k.inv_sqrt_bc2 + k.eps;
  p = p - k.step_size * (m / denom);
}

// copy16 (nullable): fp16 copy of parameters [copy_off, copy_off + copy_n),
// refreshed with the update (the hash table's gather copy).
__global__ void __launch_bounds__(256)
adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
            float* __restrict__ v, int64_t n, AdamK k, float lr_scalar,
            const float* __restrict__ lr_dev, float bc1, const int32_t* __restrict__ step_dev,
            _Float16* __restrict__ copy16, int64_t copy_off, int64_t copy_n,
            const int32_t* __restrict__ run_dev) {
  // an iteration with no sample leaves parameters and moments as they are
  // (the reference's iteration never reaches its optimiser step then)
  if (run_dev && *run_dev <= 0) return;  // uniform
  if (step_dev) {  // bias corrections of the step held on the device
    const float st = (float)*step_dev;
    bc1 = -expm1f(st * k.ln_b1);
    k.inv_sqrt_bc2 = 1.f / sqrtf(-expm1f(st * k.ln_b2));
  }
  if (lr_dev) k.step_size = lr_dev[0] / bc1;
  else k.step_size = lr_scalar / bc1;
  const int64_t n4 = n / 4;
  float4* p4 = reinterpret_cast<float4*>(p);
  const float4* g4 = reinterpret_cast<const float4*>(g);
  float4* m4 = reinterpret_cast<float4*>(m);
  float4* v4 = reinterpret_cast<float4*>(v);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4; i += stride) {
    float4 pp = p4[i], gg = g4[i], mm = m4[i], vv = v4[i];
    adam1(pp.x, gg.x, mm.x, vv.x, k);
    adam1(pp.y, gg.y, mm.y, vv.y, k);
    adam1(pp.z, gg.z, mm.z, vv.z, k);
    adam1(pp.w, gg.w, mm.w, vv.w, k);
    p4[i] = pp;
    m4[i] = mm;
    v4[i] = vv;
    if (copy16) {
      const float q[4] = {pp.x, pp.y, pp.z, pp.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t c = 4 * i + e - copy_off;
        if (c >= 0 && c < copy_n) copy16[c] = (_Float16)q[e];
      }
    }
  }
  for (int64_t i = n4 * 4 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    adam1(p[i], g[i], m[i], v[i], k);
    if (copy16 && i - copy_off >= 0 && i - copy_off < copy_n) copy16[i - copy_off] = (_Float16)p[i];
  }
}

// Several optimisers' Adam steps in one launch (the serving loop's resident
// jobs): segment s owns blocks [block0, block0 + n_blocks) of the grid and
// runs adam_kernel's body over its own flat buffers with its own learning
// rate, step count, run flag and fp16 copy.
__global__ void __launch_bounds__(256)
adam_multi_kernel(const SnAdamSegment* __restrict__ segs, int32_t n_segs, AdamK k) {
  const int b = blockIdx.x;
  int si = 0;
  while (si + 1 < n_segs && segs[si + 1].block0 <= b) ++si;  // uniform
  const SnAdamSegment sg = segs[si];
  if (b >= sg.block0 + sg.n_blocks || !sg.step_dev || !sg.lr_dev) return;  // uniform
  if (sg.run_dev && *sg.run_dev <= 0) return;
  float bc1 = 1.f;
  {
    const float st = (float)*sg.step_dev;
    bc1 = -expm1f(st * k.ln_b1);
    k.inv_sqrt_bc2 = 1.f / sqrtf(-expm1f(st * k.ln_b2));
  }
  k.step_size = sg.lr_dev[0] / bc1;
  float* p = sg.params;
  const float* g = sg.grads;
  float* m = sg.exp_avg;
  float* v = sg.exp_avg_sq;
  const int64_t n = sg.n;
  _Float16* copy16 = reinterpret_cast<_Float16*>(sg.copy16);
  const int64_t n4 = n / 4;
  float4* p4 = reinterpret_cast<float4*>(p);
  const float4* g4 = reinterpret_cast<const float4*>(g);
  float4* m4 = reinterpret_cast<float4*>(m);
  float4* v4 = reinterpret_cast<float4*>(v);
  const int64_t stride = (int64_t)sg.n_blocks * blockDim.x;
  const int64_t first = (int64_t)(b - sg.block0) * blockDim.x + threadIdx.x;
  for (int64_t i = first; i < n4; i += stride) {
    float4 pp = p4[i], gg = g4[i], mm = m4[i], vv = v4[i];
    adam1(pp.x, gg.x, mm.x, vv.x, k);
    adam1(pp.y, gg.y, mm.y, vv.y, k);
    adam1(pp.z, gg.z, mm.z, vv.z, k);
    adam1(pp.w, gg.w, mm.w, vv.w, k);
    p4[i] = pp;
    m4[i] = mm;
    v4[i] = vv;
    if (copy16) {
      const float q[4] = {pp.x, pp.y, pp.z, pp.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t c = 4 * i + e - sg.copy_off;
        if (c >= 0 && c < sg.copy_n) copy16[c] = (_Float16)q[e];
      }
    }
  }
  for (int64_t i = n4 * 4 + first; i < n; i += stride) {
    adam1(p[i], g[i], m[i], v[i], k);
    if (copy16 && i - sg.copy_off >= 0 && i - sg.copy_off < sg.copy_n)
      copy16[i - sg.copy_off] = (_Float16)p[i];
  }
}

// Advances the device step count of an iteration that updates (run > 0) or
// counts it as skipped.
__global__ void step_advance_kernel(int32_t* __restrict__ step_dev,
                                    const int32_t* __restrict__ run_dev,
                                    int32_t* __restrict__ skipped_dev) {
  if (threadIdx.x != 0) return;
  if (*run_dev > 0) ++*step_dev;
  else if (skipped_dev) ++*skipped_dev;
}

// Step guard: run = 1 when the iteration has samples and its losses and
// every gradient are finite, else 0 (the optimiser step is then skipped).
__global__ void guard_init_kernel(const int32_t* __restrict__ n_samples,
                                  const float* __restrict__ losses, int32_t n_losses,
                                  int32_t* __restrict__ run) {
  if (threadIdx.x != 0) return;
  bool ok = *n_samples > 0;
  for (int i = 0; i < n_losses; ++i) ok = ok && isfinite(losses[i]);
  *run = ok ? 1 : 0;
}

__device__ __forceinline__ bool nonfinite_bits(float x) {
  return (__float_as_uint(x) & 0x7f800000u) == 0x7f800000u;
}

__global__ void __launch_bounds__(256)
guard_grads_kernel(const float* __restrict__ g, int64_t n, int32_t* __restrict__ run) {
  if (*run == 0) return;  // uniform: nothing left to decide
  const int64_t n4 = n / 4;
  const float4* g4 = reinterpret_cast<const float4*>(g);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  bool bad = false;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4; i += stride) {
    const float4 v = g4[i];
    bad |= nonfinite_bits(v.x) | nonfinite_bits(v.y) | nonfinite_bits(v.z) | nonfinite_bits(v.w);
  }
  for (int64_t i = n4 * 4 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    bad |= nonfinite_bits(g[i]);
  // every writer stores the same 0
  if (__any(bad) && (threadIdx.x & (kWave - 1)) == 0) *run = 0;
}

__global__ void guard_count_kernel(const int32_t* __restrict__ n_samples,
                                   const int32_t* __restrict__ run,
                                   int32_t* __restrict__ n_nonfinite) {
  if (threadIdx.x == 0 && *n_samples > 0 && *run == 0) ++*n_nonfinite;
}

__global__ void __launch_bounds__(256)
to_half_kernel(const float* __restrict__ src, _Float16* __restrict__ dst, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (_Float16)src[i];
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_adam_step_copy16(float* params, const float* grads, float* exp_avg, float* exp_avg_sq,
                        int64_t n, float lr, const float* lr_dev, double beta1, double beta2,
                        double eps, int32_t step, const int32_t* step_dev, void* copy16,
                        int64_t copy_offset, int64_t copy_n, const int32_t* run_dev,
                        void* stream);

int sn_adam_step(float* params, const float* grads, float* exp_avg, float* exp_avg_sq,
                 int64_t n, float lr, const float* lr_dev, double beta1, double beta2,
                 double eps, int32_t step, const int32_t* step_dev, void* stream) {
  return sn_adam_step_copy16(params, grads, exp_avg, exp_avg_sq, n, lr, lr_dev, beta1, beta2,
                             eps, step, step_dev, nullptr, 0, 0, nullptr, stream);
}

int sn_step_advance(int32_t* step_dev, const int32_t* run_dev, int32_t* skipped_dev,
                    void* stream) {
  clear_error();
  SN_REQUIRE(step_dev && run_dev, "sn_step_advance: null pointer");
  step_advance_kernel<<<1, 64, 0, as_stream(stream)>>>(step_dev, run_dev, skipped_dev);
  return check_launch("sn_step_advance");
}

int32_t sn_adam_multi_blocks(int64_t n) {
  return n <= 0 ? 0 : (int32_t)grid_for(ceil_div(n, 4), 256, 256 * 8);
}

int sn_adam_multi(const SnAdamSegment* segments, int32_t n_segments, int32_t total_blocks,
                  double beta1, double beta2, double eps, void* stream) {
  clear_error();
  SN_REQUIRE(n_segments >= 0 && total_blocks >= 0, "sn_adam_multi: negative size");
  if (n_segments == 0 || total_blocks == 0) return SN_OK;
  SN_REQUIRE(segments, "sn_adam_multi: null segment table");
  AdamK k;
  k.b1 = (float)beta1;
  k.b2 = (float)beta2;
  k.omb1 = (float)(1.0 - beta1);
  k.omb2 = (float)(1.0 - beta2);
  k.eps = (float)eps;
  k.ln_b1 = (float)std::log(beta1);
  k.ln_b2 = (float)std::log(beta2);
  k.inv_sqrt_bc2 = 1.f;
  k.step_size = 0.f;
  adam_multi_kernel<<<(unsigned)total_blocks, 256, 0, as_stream(stream)>>>(segments, n_segments,
                                                                          k);
  return check_launch("sn_adam_multi");
}

int sn_step_guard(const int32_t* n_samples, const float* losses, int32_t n_losses,
                  const float* grads, int64_t n, int32_t* run, int32_t* n_nonfinite,
                  void* stream) {
  clear_error();
  SN_REQUIRE(n_samples && run && n >= 0 && n_losses >= 0 && (n_losses == 0 || losses),
             "sn_step_guard: bad arguments");
  SN_REQUIRE(n == 0 || (grads && (uintptr_t)grads % 16 == 0),
             "sn_step_guard: gradients must be a 16-byte aligned buffer");
  hipStream_t s = as_stream(stream);
  guard_init_kernel<<<1, 64, 0, s>>>(n_samples, losses, n_losses, run);
  if (n > 0)
    guard_grads_kernel<<<grid_for(ceil_div(n, 4), 256, 256 * 8), 256, 0, s>>>(grads, n, run);
  if (n_nonfinite) guard_count_kernel<<<1, 64, 0, s>>>(n_samples, run, n_nonfinite);
  return check_launch("sn_step_guard");
}

int sn_to_half(const float* src, void* dst, int64_t n, void* stream) {
  clear_error();
  SN_REQUIRE(n >= 0, "sn_to_half: n < 0");
  if (n == 0) return SN_OK;
  SN_REQUIRE(src && dst, "sn_to_half: null pointer");
  to_half_kernel<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(
      src, reinterpret_cast<_Float16*>(dst), n);
  return check_launch("sn_to_half");
}

int sn_adam_step_copy16(float* params, const float* grads, float* exp_avg, float* exp_avg_sq,
                        int64_t n, float lr, const float* lr_dev, double beta1, double beta2,
                        double eps, int32_t step, const int32_t* step_dev, void* copy16,
                        int64_t copy_offset, int64_t copy_n, const int32_t* run_dev,
                        void* stream) {
  clear_error();
  SN_REQUIRE(!copy16 || (copy_offset >= 0 && copy_n >= 0 && copy_offset + copy_n <= n),
             "sn_adam_step: copy range [%lld, %lld) outside [0, %lld)", (long long)copy_offset,
             (long long)(copy_offset + copy_n), (long long)n);
  SN_REQUIRE(n >= 0 && (step >= 1 || step_dev), "sn_adam_step: bad n=%lld / step=%d",
             (long long)n, step);
  if (n == 0) return SN_OK;
  SN_REQUIRE(params && grads && exp_avg && exp_avg_sq, "sn_adam_step: null pointer");
  SN_REQUIRE(((uintptr_t)params | (uintptr_t)grads | (uintptr_t)exp_avg |
              (uintptr_t)exp_avg_sq) % 16 == 0,
             "sn_adam_step: buffers must be 16-byte aligned");
  AdamK k;
  k.b1 = (float)beta1;
  k.b2 = (float)beta2;
  k.omb1 = (float)(1.0 - beta1);
  k.omb2 = (float)(1.0 - beta2);
  k.eps = (float)eps;
  k.ln_b1 = (float)std::log(beta1);
  k.ln_b2 = (float)std::log(beta2);
  const double st = step >= 1 ? (double)step : 1.0;
  const double bc1 = 1.0 - std::pow(beta1, st);
  const double bc2 = 1.0 - std::pow(beta2, st);
  k.inv_sqrt_bc2 = (float)(1.0 / std::sqrt(bc2));
  k.step_size = 0.f;
  adam_kernel<<<grid_for(ceil_div(n, 4), 256, 256 * 8), 256, 0, as_stream(stream)>>>(
      params, grads, exp_avg, exp_avg_sq, n, k, lr, lr_dev, (float)bc1, step_dev,
      reinterpret_cast<_Float16*>(copy16), copy_offset, copy_n, run_dev);
  return check_launch("sn_adam_step");
}

}  // extern "C"
