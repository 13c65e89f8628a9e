// Hash-grid encode kernels behind the tinycudann.Encoding drop-in
// (models/fields.py:25-31, 78): forward, backward to (params, x) and the
// backward of the x-gradient (needed when the SDF gradient is trained,
// fields.py:107-119 with create_graph=True).
//
// One lane per point looping over levels: all lanes of a wavefront work on
// the same level at the same time, so a level's table slice is what the
// wavefront touches (coarse dense levels stay in L2; fine hashed levels are
// random 8-byte gathers).  Parameter gradients are fp32 atomics, like the
// reference's tcnn backward.
#include <cmath>

#include "hashgrid.h"
#include "sn_common.h"

namespace sn {

int64_t make_hashgrid_desc(int32_t n_levels, int32_t log2_T, int32_t base_res,
                           float per_level_scale, int32_t n_active, HashGridDesc* d) {
  if (n_levels <= 0 || n_levels > kMaxLevels || log2_T <= 0 || log2_T > 30 || base_res <= 0 ||
      !(per_level_scale > 0.f))
    return -1;
  d->n_levels = n_levels;
  d->n_active = n_active < 0 ? 0 : (n_active > n_levels ? n_levels : n_active);
  d->dense_mask = 0;
  const uint64_t T = 1ull << log2_T;
  const double l2s = std::log2((double)per_level_scale);
  uint64_t off = 0;
  for (int l = 0; l < n_levels; ++l) {
    const float sc = (float)(std::exp2((double)l * l2s) * (double)base_res - 1.0);
    const uint32_t res = (uint32_t)std::ceil(sc) + 1u;
    uint64_t p = (uint64_t)res * res * res;
    const uint64_t max_params = 0xffffffffull / 2;
    if (p > max_params) p = max_params;
    p = (p + 7) / 8 * 8;
    if (p > T) p = T;
    d->scale[l] = sc;
    d->res[l] = res;
    d->hsize[l] = (uint32_t)p;
    d->offset[l] = (uint32_t)off;
    if ((uint64_t)res * res * res <= p) d->dense_mask |= (1u << l);
    off += p;
    if (off > 0xffffffffull) return -1;
  }
  d->offset[n_levels] = (uint32_t)off;
  return (int64_t)off;
}

// Table element type TP: float (the parameters) or _Float16 (a copy refreshed
// by the optimizer step: half the gather bytes; blending stays fp32).
template <int F, typename TP>
__device__ __forceinline__ void hash_fwd_point(int64_t i, const float* __restrict__ x,
                                               const TP* __restrict__ params,
                                               const HashGridDesc& d, float* __restrict__ enc) {
  const float p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  float* out = enc + i * (int64_t)(d.n_levels * F);
  for (int l = 0; l < d.n_levels; ++l) {
    float acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.f;
    if (l < d.n_active) {
      uint32_t idx[8];
      float w[8], fr[3];
      hg_level(p, d, l, idx, w, fr);
      const TP* tab = params + (int64_t)d.offset[l] * F;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
#pragma unroll
        for (int f = 0; f < F; ++f)
          acc[f] = fmaf(w[c], (float)tab[(int64_t)idx[c] * F + f], acc[f]);
      }
    }
#pragma unroll
    for (int f = 0; f < F; ++f) out[l * F + f] = acc[f];
  }
}

// Launches are sized by a capacity and bounded (kHashBlocks); the lanes
// stride over the count read from the device.
constexpr int kHashBlocks = 256 * 16;

template <int F, typename TP = float>
__global__ void __launch_bounds__(256)
hash_fwd_kernel(const float* __restrict__ x, int64_t n_cap, const int32_t* __restrict__ n_dev,
                const TP* __restrict__ params, HashGridDesc d, float* __restrict__ enc) {
  const int64_t n = dev_count(n_cap, n_dev);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    hash_fwd_point<F, TP>(i, x, params, d, enc);
}

// Wavefront-aggregated atomic add of F values at table row idx: lanes that
// target the same row sum their values first and one lane issues the atomics.
// Every lane of the wavefront must call it (inactive lanes with active=false).
template <int F>
__device__ __forceinline__ void wave_row_atomic_add(float* __restrict__ tab, uint32_t idx,
                                                    const float (&v)[F], bool active) {
  const int lane = threadIdx.x & 63;
  uint64_t pending = __ballot(active);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t lidx = __shfl(idx, leader, 64);
    const bool mine = active && idx == lidx;
    const uint64_t same = __ballot(mine);
#pragma unroll
    for (int f = 0; f < F; ++f) {
      float s = mine ? v[f] : 0.f;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == leader) atomicAdd(&tab[(int64_t)lidx * F + f], s);
    }
    pending &= ~same;
  }
}

// Levels whose cells are larger than ~1/kAggScale of the unit interval put many
// lanes of a wavefront (consecutive samples of a patch) on the same corners:
// those levels aggregate across the wavefront before the atomics.
constexpr float kAggScale = 192.f;

template <int F>
__device__ __forceinline__ void hash_bwd_point(int64_t i0, int64_t n,
                                               const float* __restrict__ x,
                                               const float* __restrict__ params,
                                               const HashGridDesc& d,
                                               const float* __restrict__ genc,
                                               float* __restrict__ gparams,
                                               float* __restrict__ gx, int aggregate) {
  const bool act = i0 < n;
  const int64_t i = act ? i0 : 0;  // inactive lanes stay for the wavefront collectives
  if (!aggregate && !act) return;
  const float p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  const float* g = genc + i * (int64_t)(d.n_levels * F);
  float dx[3] = {0.f, 0.f, 0.f};
  for (int l = 0; l < d.n_active; ++l) {
    float gl[F];
#pragma unroll
    for (int f = 0; f < F; ++f) gl[f] = act ? g[l * F + f] : 0.f;
    uint32_t idx[8];
    float w[8], fr[3];
    hg_level(p, d, l, idx, w, fr);
    if (gparams) {
      float* tg = gparams + (int64_t)d.offset[l] * F;
      if (aggregate && d.scale[l] < kAggScale) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          float v[F];
#pragma unroll
          for (int f = 0; f < F; ++f) v[f] = w[c] * gl[f];
          wave_row_atomic_add<F>(tg, idx[c], v, act);
        }
      } else if (act) {
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
          for (int f = 0; f < F; ++f) atomicAdd(&tg[(int64_t)idx[c] * F + f], w[c] * gl[f]);
      }
    }
    if (gx) {
      float dw[8][3];
      hg_dweights(fr, dw);
      const float* tab = params + (int64_t)d.offset[l] * F;
      float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        float s = 0.f;
#pragma unroll
        for (int f = 0; f < F; ++f) s += gl[f] * tab[(int64_t)idx[c] * F + f];
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += dw[c][k] * s;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) dx[k] += d.scale[l] * acc[k];
    }
  }
  if (gx && act) {
#pragma unroll
    for (int k = 0; k < 3; ++k) gx[3 * i + k] = dx[k];
  }
}

template <int F>
__global__ void __launch_bounds__(256)
hash_bwd_kernel(const float* __restrict__ x, int64_t n_cap, const int32_t* __restrict__ n_dev,
                const float* __restrict__ params, HashGridDesc d,
                const float* __restrict__ genc, float* __restrict__ gparams,
                float* __restrict__ gx, int aggregate) {
  const int64_t n = dev_count(n_cap, n_dev);
  // the stride is a multiple of 64, so a wavefront's lanes stay together and
  // leave the loop together (the aggregation uses wavefront collectives)
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; (i0 & ~(int64_t)63) < n;
       i0 += (int64_t)gridDim.x * blockDim.x)
    hash_bwd_point<F>(i0, n, x, params, d, genc, gparams, gx, aggregate);
}

// Backward with the table gradient merged per workgroup in LDS first.  A
// workgroup takes 256 consecutive points (a few samples of one patch's rays:
// spatially compact) and, level by level, adds every corner contribution into
// an LDS hash table keyed by the entry index (open addressing, kLdsSlots
// slots, LDS atomics); the distinct entries are then flushed to the global
// gradient with one fp32 atomic per feature.  Consecutive samples share most
// corners on all but the finest levels, so the global atomics drop several-
// fold; an insertion that finds no slot within kMaxProbe probes falls back to
// a global atomic.  Same sums as the plain kernel, in another order.
constexpr int kLdsSlots = 4096;
constexpr int kMaxProbe = 32;
constexpr uint32_t kEmpty = 0xffffffffu;

template <int F>
__global__ void __launch_bounds__(256)
hash_bwd_lds_kernel(const float* __restrict__ x, int64_t n_cap, const int32_t* __restrict__ n_dev,
                    const float* __restrict__ params, HashGridDesc d,
                    const float* __restrict__ genc, float* __restrict__ gparams,
                    float* __restrict__ gx, int level_lo, int level_hi) {
  __shared__ uint32_t keys[kLdsSlots];
  __shared__ float vals[kLdsSlots * F];
  const int tid = threadIdx.x;
  for (int sl = tid; sl < kLdsSlots; sl += 256) {
    keys[sl] = kEmpty;
#pragma unroll
    for (int f = 0; f < F; ++f) vals[sl * F + f] = 0.f;
  }
  __syncthreads();
  const int64_t n = dev_count(n_cap, n_dev);
  const int64_t n_chunks = ceil_div(n, 256);
  for (int64_t ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {  // uniform per workgroup
    const int64_t i = ch * 256 + tid;
    const bool act = i < n;
    float p[3] = {0.f, 0.f, 0.f};
    if (act) {
#pragma unroll
      for (int k = 0; k < 3; ++k) p[k] = x[3 * i + k];
    }
    const float* g = genc + (act ? i : 0) * (int64_t)(d.n_levels * F);
    float dx[3] = {0.f, 0.f, 0.f};
    for (int l = level_lo; l < level_hi; ++l) {
      float* tg = gparams + (int64_t)d.offset[l] * F;
      if (act) {
        float gl[F];
#pragma unroll
        for (int f = 0; f < F; ++f) gl[f] = g[l * F + f];
        uint32_t idx[8];
        float w[8], fr[3];
        hg_level(p, d, l, idx, w, fr);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          uint32_t h = (idx[c] * 2654435761u) >> 20;  // 12 bits: kLdsSlots
          int probe = 0;
          for (;;) {
            const uint32_t old = atomicCAS(&keys[h], kEmpty, idx[c]);
            if (old == kEmpty || old == idx[c]) {
#pragma unroll
              for (int f = 0; f < F; ++f) atomicAdd(&vals[h * F + f], w[c] * gl[f]);
              break;
            }
            if (++probe == kMaxProbe) {
#pragma unroll
              for (int f = 0; f < F; ++f) atomicAdd(&tg[(int64_t)idx[c] * F + f], w[c] * gl[f]);
              break;
            }
            h = (h + 1) & (kLdsSlots - 1);
          }
        }
        if (gx) {
          float dw[8][3];
          hg_dweights(fr, dw);
          const float* tab = params + (int64_t)d.offset[l] * F;
          float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            float sm = 0.f;
#pragma unroll
            for (int f = 0; f < F; ++f) sm += gl[f] * tab[(int64_t)idx[c] * F + f];
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += dw[c][k] * sm;
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) dx[k] += d.scale[l] * acc[k];
        }
      }
      __syncthreads();
      // flush the merged entries of this level and clear the table
      for (int sl = tid; sl < kLdsSlots; sl += 256) {
        const uint32_t key = keys[sl];
        if (key != kEmpty) {
#pragma unroll
          for (int f = 0; f < F; ++f) {
            atomicAdd(&tg[(int64_t)key * F + f], vals[sl * F + f]);
            vals[sl * F + f] = 0.f;
          }
          keys[sl] = kEmpty;
        }
      }
      __syncthreads();
    }
    if (gx && act) {
#pragma unroll
      for (int k = 0; k < 3; ++k) gx[3 * i + k] = dx[k];
    }
  }
}

// Cotangent g_gx of grad_x = sum_{l,f,c} genc[l,f] * scale_l * dw[c] . theta[c,f]:
//   grad_grad_enc[l,f] = scale_l * sum_c (g_gx . dw[c]) theta[c,f]
//   grad_params[c,f]  += scale_l * (g_gx . dw[c]) * genc[l,f]
//   grad_x[e]          = scale_l^2 * sum_c S_c sum_{d != e} g_gx[d] d2w[c]/dfrac_d dfrac_e,
//                        S_c = sum_f genc[l,f] theta[c,f]  (d2w/dfrac_d^2 = 0)
template <int F>
__global__ void __launch_bounds__(256)
hash_bwd_bwd_kernel(const float* __restrict__ x, int64_t n_cap,
                    const int32_t* __restrict__ n_dev, const float* __restrict__ params,
                    HashGridDesc d, const float* __restrict__ genc,
                    const float* __restrict__ ggx, float* __restrict__ gge,
                    float* __restrict__ gparams, float* __restrict__ gx2) {
  const int64_t n = dev_count(n_cap, n_dev);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
  const float p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  const float v[3] = {ggx[3 * i], ggx[3 * i + 1], ggx[3 * i + 2]};
  float acc2[3] = {0.f, 0.f, 0.f};
  for (int l = 0; l < d.n_levels; ++l) {
    float out[F];
#pragma unroll
    for (int f = 0; f < F; ++f) out[f] = 0.f;
    if (l < d.n_active) {
      uint32_t idx[8];
      float w[8], fr[3];
      hg_level(p, d, l, idx, w, fr);
      float dw[8][3];
      hg_dweights(fr, dw);
      const float sc = d.scale[l];
      const float* tab = params + (int64_t)d.offset[l] * F;
      float* tg = gparams ? gparams + (int64_t)d.offset[l] * F : nullptr;
      float gl[F];
#pragma unroll
      for (int f = 0; f < F; ++f)
        gl[f] = genc ? genc[i * (int64_t)(d.n_levels * F) + l * F + f] : 0.f;
      float l2[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float a = sc * (v[0] * dw[c][0] + v[1] * dw[c][1] + v[2] * dw[c][2]);
        float S = 0.f;
#pragma unroll
        for (int f = 0; f < F; ++f) {
          const float th = tab[(int64_t)idx[c] * F + f];
          if (gge) out[f] += a * th;
          if (tg) atomicAdd(&tg[(int64_t)idx[c] * F + f], a * gl[f]);
          S += gl[f] * th;
        }
        if (gx2) {
          // d2w/(dfrac_d dfrac_e) = s_d s_e a_o (o = the third axis)
          const float s0 = (c & 1) ? 1.f : -1.f, s1 = ((c >> 1) & 1) ? 1.f : -1.f,
                      s2 = ((c >> 2) & 1) ? 1.f : -1.f;
          const float a0 = (c & 1) ? fr[0] : 1.f - fr[0];
          const float a1 = ((c >> 1) & 1) ? fr[1] : 1.f - fr[1];
          const float a2 = ((c >> 2) & 1) ? fr[2] : 1.f - fr[2];
          const float h01 = s0 * s1 * a2, h02 = s0 * s2 * a1, h12 = s1 * s2 * a0;
          l2[0] += S * (v[1] * h01 + v[2] * h02);
          l2[1] += S * (v[0] * h01 + v[2] * h12);
          l2[2] += S * (v[0] * h02 + v[1] * h12);
        }
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) acc2[k] += sc * sc * l2[k];
    }
    if (gge) {
#pragma unroll
      for (int f = 0; f < F; ++f) gge[i * (int64_t)(d.n_levels * F) + l * F + f] = out[f];
    }
  }
  if (gx2) {
#pragma unroll
    for (int k = 0; k < 3; ++k) gx2[3 * i + k] = acc2[k];
  }
  }
}

static int hg_setup(const char* who, int32_t n_levels, int32_t F, int32_t log2_T, int32_t base,
                    float pls, int32_t n_active, HashGridDesc* d) {
  SN_REQUIRE(F == 1 || F == 2 || F == 4 || F == 8, "%s: n_features_per_level must be 1,2,4,8",
             who);
  SN_REQUIRE(make_hashgrid_desc(n_levels, log2_T, base, pls, n_active, d) > 0,
             "%s: bad hash-grid config (levels %d, log2 T %d, base %d, scale %g)", who,
             n_levels, log2_T, base, (double)pls);
  return SN_OK;
}

// 0: plain atomics, 1: wavefront aggregation of coarse levels, 2: workgroup
// LDS merge of every level (F <= 2; wider features use 1)
static int g_aggregate = 2;

#define SN_DISPATCH_F(F, ...)          \
  switch (F) {                         \
    case 1: { constexpr int kF = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int kF = 2; __VA_ARGS__; } break; \
    case 4: { constexpr int kF = 4; __VA_ARGS__; } break; \
    default: { constexpr int kF = 8; __VA_ARGS__; } break; \
  }

}  // namespace sn

using namespace sn;

extern "C" {

void sn_hashgrid_set_aggregation(int32_t mode) { g_aggregate = mode < 0 ? 0 : (mode > 2 ? 2 : mode); }

int64_t sn_hashgrid_layout(int32_t n_levels, int32_t log2_hashmap_size, int32_t base_resolution,
                           float per_level_scale, uint32_t* offsets) {
  HashGridDesc d;
  int64_t tot = make_hashgrid_desc(n_levels, log2_hashmap_size, base_resolution,
                                   per_level_scale, n_levels, &d);
  if (tot < 0) return -1;
  if (offsets)
    for (int l = 0; l <= n_levels; ++l) offsets[l] = d.offset[l];
  return tot;
}

int sn_hashgrid_fwd_dev(const float* x, int64_t n_cap, const int32_t* n_dev, const float* params,
                        int32_t n_levels, int32_t n_features, int32_t log2_hashmap_size,
                        int32_t base_resolution, float per_level_scale, int32_t n_active_levels,
                        float* enc, void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_fwd", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0, "sn_hashgrid_fwd: n < 0");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && params && enc, "sn_hashgrid_fwd: null pointer");
  SN_DISPATCH_F(n_features, hash_fwd_kernel<kF><<<grid_for(n_cap, 256, kHashBlocks), 256, 0,
                                                  as_stream(stream)>>>(x, n_cap, n_dev, params,
                                                                       d, enc));
  return check_launch("sn_hashgrid_fwd");
}

int sn_hashgrid_fwd16_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                          const void* params16, int32_t n_levels, int32_t n_features,
                          int32_t log2_hashmap_size, int32_t base_resolution,
                          float per_level_scale, int32_t n_active_levels, float* enc,
                          void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_fwd16", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0, "sn_hashgrid_fwd16: n < 0");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && params16 && enc, "sn_hashgrid_fwd16: null pointer");
  const _Float16* p16 = reinterpret_cast<const _Float16*>(params16);
  SN_DISPATCH_F(n_features, (hash_fwd_kernel<kF, _Float16><<<grid_for(n_cap, 256, kHashBlocks), 256,
                                                              0, as_stream(stream)>>>(
                                x, n_cap, n_dev, p16, d, enc)));
  return check_launch("sn_hashgrid_fwd16");
}

int sn_hashgrid_fwd(const float* x, int64_t n, const float* params, int32_t n_levels,
                    int32_t n_features, int32_t log2_hashmap_size, int32_t base_resolution,
                    float per_level_scale, int32_t n_active_levels, float* enc, void* stream) {
  return sn_hashgrid_fwd_dev(x, n, nullptr, params, n_levels, n_features, log2_hashmap_size,
                             base_resolution, per_level_scale, n_active_levels, enc, stream);
}

int sn_hashgrid_bwd_dev(const float* x, int64_t n_cap, const int32_t* n_dev, const float* params,
                        int32_t n_levels, int32_t n_features, int32_t log2_hashmap_size,
                        int32_t base_resolution, float per_level_scale, int32_t n_active_levels,
                        const float* grad_enc, float* grad_params, float* grad_x, void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_bwd", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0, "sn_hashgrid_bwd: n < 0");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && params && grad_enc, "sn_hashgrid_bwd: null input");
  SN_REQUIRE(grad_params || grad_x, "sn_hashgrid_bwd: no output");
  if (g_aggregate == 2 && grad_params && n_features <= 2) {
    // (one launch over all levels: measured faster than a launch per level)
    const unsigned grid = grid_for(n_cap, 256, kHashBlocks);
    hipStream_t s = as_stream(stream);
    if (n_features == 1)
      hash_bwd_lds_kernel<1><<<grid, 256, 0, s>>>(x, n_cap, n_dev, params, d, grad_enc,
                                                   grad_params, grad_x, 0, d.n_active);
    else
      hash_bwd_lds_kernel<2><<<grid, 256, 0, s>>>(x, n_cap, n_dev, params, d, grad_enc,
                                                   grad_params, grad_x, 0, d.n_active);
    return check_launch("sn_hashgrid_bwd");
  }
  const int agg = g_aggregate >= 1 ? 1 : 0;
  SN_DISPATCH_F(n_features,
                hash_bwd_kernel<kF><<<grid_for(n_cap, 256, kHashBlocks), 256, 0,
                                      as_stream(stream)>>>(x, n_cap, n_dev, params, d, grad_enc,
                                                           grad_params, grad_x, agg));
  return check_launch("sn_hashgrid_bwd");
}

int sn_hashgrid_bwd(const float* x, int64_t n, const float* params, int32_t n_levels,
                    int32_t n_features, int32_t log2_hashmap_size, int32_t base_resolution,
                    float per_level_scale, int32_t n_active_levels, const float* grad_enc,
                    float* grad_params, float* grad_x, void* stream) {
  return sn_hashgrid_bwd_dev(x, n, nullptr, params, n_levels, n_features, log2_hashmap_size,
                             base_resolution, per_level_scale, n_active_levels, grad_enc,
                             grad_params, grad_x, stream);
}

int sn_hashgrid_bwd_bwd_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                            const float* params, int32_t n_levels, int32_t n_features,
                            int32_t log2_hashmap_size, int32_t base_resolution,
                            float per_level_scale, int32_t n_active_levels,
                            const float* grad_enc, const float* g_gx, float* grad_grad_enc,
                            float* grad_params, float* grad_x, void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_bwd_bwd", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0, "sn_hashgrid_bwd_bwd: n < 0");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && params && g_gx, "sn_hashgrid_bwd_bwd: null input");
  SN_REQUIRE(grad_grad_enc || ((grad_params || grad_x) && grad_enc),
             "sn_hashgrid_bwd_bwd: no output");
  SN_DISPATCH_F(n_features, hash_bwd_bwd_kernel<kF><<<grid_for(n_cap, 256, kHashBlocks), 256, 0,
                                                      as_stream(stream)>>>(
                                x, n_cap, n_dev, params, d, grad_enc, g_gx, grad_grad_enc,
                                grad_params, grad_x));
  return check_launch("sn_hashgrid_bwd_bwd");
}

int sn_hashgrid_bwd_bwd(const float* x, int64_t n, const float* params, int32_t n_levels,
                        int32_t n_features, int32_t log2_hashmap_size, int32_t base_resolution,
                        float per_level_scale, int32_t n_active_levels, const float* grad_enc,
                        const float* g_gx, float* grad_grad_enc, float* grad_params,
                        float* grad_x, void* stream) {
  return sn_hashgrid_bwd_bwd_dev(x, n, nullptr, params, n_levels, n_features, log2_hashmap_size,
                                 base_resolution, per_level_scale, n_active_levels, grad_enc,
                                 g_gx, grad_grad_enc, grad_params, grad_x, stream);
}

}  // extern "C"
