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
  d->n_active_dev = nullptr;
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

// Levels per step of the paired forward (F = 2, 16-byte output rows): 2 or 4.
#ifndef SN_FWD_LEVELS
#define SN_FWD_LEVELS 4
#endif
constexpr int kFwdLevels = SN_FWD_LEVELS;
static_assert(kFwdLevels == 2 || kFwdLevels == 4, "forward level group of 2 or 4");

// Table element type TP: float (the parameters) or _Float16 (a copy refreshed
// by the optimizer step: half the gather bytes; blending stays fp32).
template <int F, typename TP, bool kPair>
__device__ __forceinline__ void hash_fwd_point(int64_t i, const float* __restrict__ x,
                                               const TP* __restrict__ params,
                                               const HashGridDesc& d, int na,
                                               float* __restrict__ enc) {
  const float p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  float* out = enc + i * (int64_t)(d.n_levels * F);
  // F = 2 pairs: the outputs of 4 levels are written together (two 16-byte
  // stores) when the rows allow it, instead of 8 bytes per level into lines
  // the L2 does not keep between levels
  const bool vec_out = F == 2 && kPair && (d.n_levels & 3) == 0 &&
                       (reinterpret_cast<uintptr_t>(enc) & 15) == 0;
  if constexpr (F == 2 && kPair) {
    if (d.n_levels % kFwdLevels == 0 && (reinterpret_cast<uintptr_t>(enc) & 15) == 0) {
      // kFwdLevels levels at a time: their corner gathers are all issued
      // before any is used (that many times the loads in flight of a
      // level-by-level loop; the gathers are what this kernel waits on), then
      // blended in the same order as below
      typedef TP pair_t __attribute__((ext_vector_type(2)));
      constexpr int G = kFwdLevels;
      for (int l0 = 0; l0 < d.n_levels; l0 += G) {
        uint32_t idx[G][8];
        float w[G][8];
        pair_t v[G][8];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          if (l0 + j < na) {  // uniform
            float fr[3];
            hg_level(p, d, l0 + j, idx[j], w[j], fr);
            const pair_t* tab2 =
                reinterpret_cast<const pair_t*>(params + (int64_t)d.offset[l0 + j] * F);
#pragma unroll
            for (int c = 0; c < 8; ++c) v[j][c] = tab2[idx[j][c]];
          }
        }
        float o[2 * G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          float a0 = 0.f, a1 = 0.f;
          if (l0 + j < na) {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
              a0 = fmaf(w[j][c], (float)v[j][c].x, a0);
              a1 = fmaf(w[j][c], (float)v[j][c].y, a1);
            }
          }
          o[2 * j] = a0;
          o[2 * j + 1] = a1;
        }
#pragma unroll
        for (int j = 0; j < G; j += 2)
          *reinterpret_cast<float4*>(out + (l0 + j) * F) =
              make_float4(o[2 * j], o[2 * j + 1], o[2 * j + 2], o[2 * j + 3]);
      }
      return;
    }
  }
  float o0 = 0.f, o1 = 0.f, o2 = 0.f, o3 = 0.f, o4 = 0.f, o5 = 0.f, o6 = 0.f, o7 = 0.f;
  for (int l = 0; l < d.n_levels; ++l) {
    float acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.f;
    if (l < na) {
      uint32_t idx[8];
      float w[8], fr[3];
      hg_level(p, d, l, idx, w, fr);
      const TP* tab = params + (int64_t)d.offset[l] * F;
      if constexpr (F == 2 && kPair) {
        // both features of a corner in one load (rows are 2 * sizeof(TP) aligned)
        typedef TP pair_t __attribute__((ext_vector_type(2)));
        const pair_t* tab2 = reinterpret_cast<const pair_t*>(tab);
        pair_t v[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) v[c] = tab2[idx[c]];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          acc[0] = fmaf(w[c], (float)v[c].x, acc[0]);
          acc[1] = fmaf(w[c], (float)v[c].y, acc[1]);
        }
      } else {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
#pragma unroll
          for (int f = 0; f < F; ++f)
            acc[f] = fmaf(w[c], (float)tab[(int64_t)idx[c] * F + f], acc[f]);
        }
      }
    }
    if constexpr (F == 2 && kPair) {
      if (vec_out) {
        const int k = l & 3;
        if (k == 0) {
          o0 = acc[0];
          o1 = acc[1];
        } else if (k == 1) {
          o2 = acc[0];
          o3 = acc[1];
        } else if (k == 2) {
          o4 = acc[0];
          o5 = acc[1];
        } else {
          o6 = acc[0];
          o7 = acc[1];
          float4* o4p = reinterpret_cast<float4*>(out + (l - 3) * F);
          o4p[0] = make_float4(o0, o1, o2, o3);
          o4p[1] = make_float4(o4, o5, o6, o7);
        }
      } else {
        typedef float f2_t __attribute__((ext_vector_type(2)));
        *reinterpret_cast<f2_t*>(out + l * F) = f2_t{acc[0], acc[1]};
      }
    } else {
#pragma unroll
      for (int f = 0; f < F; ++f) out[l * F + f] = acc[f];
    }
  }
}

// Launches are sized by a capacity and bounded (kHashBlocks); the lanes
// stride over the count read from the device.
constexpr int kHashBlocks = 256 * 16;

// The analytic x-gradient of the SDF (fields.py:107-119) from the MLP's
// d sdf / d h0 rows (h0 = [x, enc], row stride ld): the encoding's part goes
// through the table (a gather: sum over levels and corners of
// scale_l * d w_c / d frac * (dsdf/denc_l . table row)), the position's part
// is added last -- the order in which the backward to x (grad_x) plus the
// direct term sums it.
template <int F, typename TP>
__global__ void __launch_bounds__(256)
hash_dx_kernel(const float* __restrict__ x, int64_t n_cap, const int32_t* __restrict__ n_dev,
               const TP* __restrict__ params, HashGridDesc d, const float* __restrict__ dh0,
               int32_t ld, float* __restrict__ grad) {
  const int64_t n = dev_count(n_cap, n_dev);
  const int na = active_levels(d);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    const float* g = dh0 + i * (int64_t)ld;
    float dx[3] = {0.f, 0.f, 0.f};
    for (int l = 0; l < na; ++l) {
      uint32_t idx[8];
      float w[8], fr[3], dw[8][3];
      hg_level(p, d, l, idx, w, fr);
      hg_dweights(fr, dw);
      const TP* tab = params + (int64_t)d.offset[l] * F;
      float gl[F];
#pragma unroll
      for (int f = 0; f < F; ++f) gl[f] = g[3 + l * F + f];
      float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        float sm = 0.f;
#pragma unroll
        for (int f = 0; f < F; ++f) sm += gl[f] * (float)tab[(int64_t)idx[c] * F + f];
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += dw[c][k] * sm;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) dx[k] += d.scale[l] * acc[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) grad[3 * i + k] = dx[k] + g[k];
  }
}

// hash_dx_kernel for F = 2 with the levels a multiple of 4: 4 levels' 32
// corner gathers issued before any is used; the same sums in the same order.
template <typename TP>
__global__ void __launch_bounds__(256)
hash_dx2_kernel(const float* __restrict__ x, int64_t n_cap, const int32_t* __restrict__ n_dev,
                const TP* __restrict__ params, HashGridDesc d, const float* __restrict__ dh0,
                int32_t ld, float* __restrict__ grad) {
  typedef TP pair_t __attribute__((ext_vector_type(2)));
  const int64_t n = dev_count(n_cap, n_dev);
  const int na = active_levels(d);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    const float* g = dh0 + i * (int64_t)ld;
    float dx[3] = {0.f, 0.f, 0.f};
    for (int l0 = 0; l0 < na; l0 += 4) {
      float fr[4][3];
      pair_t t[4][8];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (l0 + j < na) {  // uniform
          uint32_t idx[8];
          float w[8];
          hg_level(p, d, l0 + j, idx, w, fr[j]);
          const pair_t* tab2 =
              reinterpret_cast<const pair_t*>(params + (int64_t)d.offset[l0 + j] * 2);
#pragma unroll
          for (int c = 0; c < 8; ++c) t[j][c] = tab2[idx[c]];
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (l0 + j < na) {
          const int l = l0 + j;
          float dw[8][3];
          hg_dweights(fr[j], dw);
          const float gl0 = g[3 + l * 2], gl1 = g[3 + l * 2 + 1];
          float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            float sm = 0.f;
            sm += gl0 * (float)t[j][c].x;
            sm += gl1 * (float)t[j][c].y;
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += dw[c][k] * sm;
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) dx[k] += d.scale[l] * acc[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) grad[3 * i + k] = dx[k] + g[k];
  }
}

template <int F, typename TP = float, bool kPair = false>
__global__ void __launch_bounds__(256)
hash_fwd_kernel(const float* __restrict__ x, int64_t n_cap, const int32_t* __restrict__ n_dev,
                const TP* __restrict__ params, HashGridDesc d, float* __restrict__ enc) {
  const int64_t n = dev_count(n_cap, n_dev);
  const int na = active_levels(d);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    hash_fwd_point<F, TP, kPair>(i, x, params, d, na, enc);
}

// The encoding's directional derivative along v (the tangent the double
// backward of the analytic gradient needs, hash_bwd_bwd_kernel's grad_grad_enc):
// out[l, f] = sum_c scale_l (v . d w_c / d frac) table_l[idx_c, f], from the
// fp32 table or its fp16 copy.  F = 2, rows of 16 bytes: 4 levels per step,
// their 32 corner gathers issued before any is used (as the forward).
template <typename TP>
__global__ void __launch_bounds__(256)
hash_tangent_kernel(const float* __restrict__ x, int64_t n_cap, const int32_t* __restrict__ n_dev,
                    const TP* __restrict__ params, HashGridDesc d, const float* __restrict__ gv,
                    float* __restrict__ out) {
  typedef TP pair_t __attribute__((ext_vector_type(2)));
  const int64_t n = dev_count(n_cap, n_dev);
  const int na = active_levels(d);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float p[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    const float v[3] = {gv[3 * i], gv[3 * i + 1], gv[3 * i + 2]};
    float* o = out + i * (int64_t)(d.n_levels * 2);
    for (int l0 = 0; l0 < d.n_levels; l0 += 4) {
      float a[4][8];
      pair_t t[4][8];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (l0 + j < na) {  // uniform
          uint32_t idx[8];
          float w[8], fr[3], dw[8][3];
          hg_level(p, d, l0 + j, idx, w, fr);
          hg_dweights(fr, dw);
          const float sc = d.scale[l0 + j];
#pragma unroll
          for (int c = 0; c < 8; ++c)
            a[j][c] = sc * (v[0] * dw[c][0] + v[1] * dw[c][1] + v[2] * dw[c][2]);
          const pair_t* tab2 = reinterpret_cast<const pair_t*>(params + (int64_t)d.offset[l0 + j] * 2);
#pragma unroll
          for (int c = 0; c < 8; ++c) t[j][c] = tab2[idx[c]];
        }
      }
      float r[8];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float s0 = 0.f, s1 = 0.f;
        if (l0 + j < na) {
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            s0 += a[j][c] * (float)t[j][c].x;
            s1 += a[j][c] * (float)t[j][c].y;
          }
        }
        r[2 * j] = s0;
        r[2 * j + 1] = s1;
      }
      float4* o4 = reinterpret_cast<float4*>(o + l0 * 2);
      o4[0] = make_float4(r[0], r[1], r[2], r[3]);
      o4[1] = make_float4(r[4], r[5], r[6], r[7]);
    }
  }
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
  for (int l = 0, na_ = active_levels(d); l < na_; ++l) {
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
// gradient with one fp32 atomic per feature.  An insertion that finds no slot
// within kMaxProbe probes falls back to a global atomic.  Same sums as the
// plain kernel, in another order.  (Used without a workspace; the bucket
// path below is faster: fp32 atomics, in LDS and in global memory, are slow
// on this GPU -- about 2e11 and 2e10 lane operations/s for the device.)
constexpr int kLdsSlots = 4096;
constexpr int kMaxProbe = 32;
constexpr uint32_t kEmpty = 0xffffffffu;

template <int F>
__global__ void __launch_bounds__(256)
hash_bwd_lds_kernel(const float* __restrict__ x, int64_t n_cap, const int32_t* __restrict__ n_dev,
                    const float* __restrict__ params, HashGridDesc d,
                    const float* __restrict__ genc, float* __restrict__ gparams,
                    float* __restrict__ gx) {
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
    for (int l = 0, na_ = active_levels(d); l < na_; ++l) {
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

// ---------------------------------------------------------------------------
// Bucket path (with a workspace).  Two kernels, no fp32 atomics:
//
// 1. hash_bwd_rec_kernel: a workgroup takes kRecThreads (512) consecutive points and, per
//    level, turns their corner contributions into records (table row, F
//    values).  Lanes of a wavefront whose points lie in the same cell (common
//    on coarse and middle levels: a wavefront holds neighbouring samples of
//    a few rays of one patch) are summed first with cross-lane (DPP)
//    reduce-scatters and leave one record per corner; the other lanes leave
//    one record per corner each.
//    Records go to the list of their bucket (kBucketEntries consecutive rows
//    of the level): ranks within the workgroup come from integer LDS atomics,
//    room in each list from one global atomic per (workgroup, level, bucket).
//    Per level: ranks, barrier, offsets (one wavefront's scan, its
//    reservation atomics issued first), barrier, staging in LDS by bucket
//    while those atomics are in flight, barrier, coalesced writes (measured
//    per level and workgroup on the late iterations of a full training run,
//    with the reservation already overlapping the staging but before the
//    per-wavefront level maximum and the DPP scan: hashing and loads 1.1 us,
//    group search 1.6 us, ranks 1.35 us, waiting for the slowest wavefront
//    1.4 us, scan 0.6 us, staging 0.6 us (wavefront 0 with the reservation
//    1.4 us), writes 1.1 us).
//    Records beyond a list's capacity are added with fp32 atomics instead.
//    The lists share one pool; bucket_layout_kernel (before the records)
//    gives each list 5/4 of the records it took in the previous call on the
//    same workspace plus a floor (a quarter of the pool split evenly), at
//    most kMaxList (or the even split if larger); then either scales these
//    down to the pool or shares out what they leave of it evenly (up to the
//    same cap), and uses the even split when the workspace holds no previous
//    call's layout.  So the buckets that take many records get the room, and
//    a pool of at least 5/4 of a call's records holds every list of a call
//    that repeats the previous one's counts: each list's share is then at
//    least (5/4 c) * pool / (5/4 records + pool / 4) >= c.
//    The largest |value| of each level is kept (integer max on the bits of a
//    non-negative float: per wavefront by DPP, then one LDS atomic per
//    wavefront and one global atomic per workgroup).
// 2. hash_bwd_bucket_sum_kernel: one workgroup per bucket sums its records in
//    LDS as 64-bit fixed point (integer LDS atomics; scale 2^K with
//    K = 61 - ceil(log2(count * max|value|)), so no sum can overflow and each
//    record is rounded to count * max|value| * 2^-62 at worst, far below fp32
//    resolution of anything but values ~2^-40 of the level's largest), then
//    adds the sums to its rows, which no other workgroup touches.
// Integer sums do not depend on the order, so the result is reproducible
// (when no list overflows).
constexpr int kBucketShift = 13;
constexpr int kBucketEntries = 1 << kBucketShift;
constexpr int kMaxBucketsPerLevel = 512;
// Same-cell lane groups of kMinGroup or more lanes are summed across lanes
// (group_sums; measured on the late iterations of a full training run, ms
// per hash-grid backward: a minimum of 3 / 4 / 5 / 6 / 8 lanes -> 1.07 /
// 1.05 / 1.036 / 1.03 / 1.045: smaller groups cost more in cross-lane sums
// than they save in records; with the cheaper ranks and staging of the
// current record kernel 5 / 6 / 8 -> 0.897 / 0.901 / 0.901 ms, no difference); the search for groups (one leader per step,
// lowest pending lane first) stops after kMaxSingles leaders in a row with a
// smaller group (3 / 6 / 10 -> 1.05 / 1.03 / 1.044 ms), and is skipped on
// levels where no lane shares the cell of the lane before it (the fine
// levels).
#ifndef SN_MIN_GROUP
#define SN_MIN_GROUP 6
#endif
#ifndef SN_MAX_SINGLES
#define SN_MAX_SINGLES 6
#endif
constexpr int kMinGroup = SN_MIN_GROUP;
constexpr int kMaxSingles = SN_MAX_SINGLES;
#ifndef SN_REC_THREADS
#define SN_REC_THREADS 512
#endif
// points (lanes) per workgroup of the record kernel (the backward per late
// launch of a full training run: 0.90 ms at 512, 1.22 ms at 256)
constexpr int kRecThreads = SN_REC_THREADS;

// Buckets of each level: level l owns global buckets [base[l], base[l+1]).
struct BucketDesc {
  int32_t base[kMaxLevels + 1];
};

struct BucketPool {
  int32_t* hdr;       // [4] layout stamp: kLayoutMagic, n_buckets, total records
  int32_t* cursor;    // [n_buckets] records reserved per bucket (may exceed its cap)
  uint32_t* maxbits;  // [kMaxLevels] bits of the largest |record value| per level
  int32_t* start;     // [n_buckets] first pool index of each bucket's list
  int32_t* bcap;      // [n_buckets] list capacity
  uint16_t* row;      // [total] row within the bucket
  float* val;         // [total * F]
  int32_t total;      // records in the pool
};

constexpr int32_t kLayoutMagic = 0x4c594f54;
// Largest list unless the pool's even split is larger: the sum kernel takes
// a list in one workgroup, whose integer LDS atomics (~5e9 lane ops/s per CU)
// make a list of n records last at least n * F / 5e9 s; records beyond it
// take the atomic path.  2^18: the coarsest levels' one or two buckets take
// ~230-270 k records per late call of a 1.2 M-point backward (2^17 sent 14 %
// of them to the atomics; the longest list, ~100 us in the sum kernel, runs
// beside the other buckets' workgroups).
constexpr int64_t kMaxList = 262144;
constexpr int64_t kMaxPoolRecords = 1ll << 30;  // keeps the layout's products in int64
constexpr int kLayoutThreads = 1024;
constexpr int kMaxBuckets = kMaxLevels * kMaxBucketsPerLevel;

// One workgroup: list capacities from the previous call's counts (cursor),
// start offsets (exclusive scan), counts and level maxima back to zero.  Any
// workspace content gives a valid layout (starts + caps within the pool);
// only its speed depends on the content.
__global__ void __launch_bounds__(kLayoutThreads)
bucket_layout_kernel(BucketPool pool, int32_t nb) {
  constexpr int kPer = (kMaxBuckets + kLayoutThreads - 1) / kLayoutThreads;
  constexpr int kWaves = kLayoutThreads / 64;
  __shared__ long long wsum[kWaves];
  __shared__ int valid;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t total = pool.total;
  if (tid == 0)
    valid = pool.hdr[0] == kLayoutMagic && pool.hdr[1] == nb && pool.hdr[2] == pool.total;
  __syncthreads();
  const int per = (nb + kLayoutThreads - 1) / kLayoutThreads;  // <= kPer
  const int64_t floor_share = max((int64_t)1, total / (4 * (int64_t)nb));
  const int64_t even = total / nb;
  int64_t want[kPer];
  long long s = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int b = tid * per + q;
    int64_t w = 0;
    if (q < per && b < nb) {
      if (valid) {
        const int64_t prev = min((int64_t)max(pool.cursor[b], 0), total);
        w = prev + prev / 4 + floor_share;
      } else {
        w = even;
      }
      w = min(w, max(even, kMaxList));
    }
    want[q] = w;
    s += w;
  }
  // block sum
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) wsum[wv] = s;
  __syncthreads();
  long long tot = 0;
#pragma unroll
  for (int k = 0; k < kWaves; ++k) tot += wsum[k];
  __syncthreads();
  // scale down to the pool, by the 32.32 fixed-point factor total / tot
  // (w <= 2.25 total < 2^31.2, total <= 2^30: w * factor < 2^64; the
  // rounding keeps the sum within the pool), or share what the wants leave
  // of it evenly (up to the cap)
  const int64_t cap = max(even, kMaxList);
  const bool scale = tot > total;
  const uint64_t factor = scale ? ((uint64_t)total << 32) / (uint64_t)tot : 0;
  const int64_t extra = scale ? 0 : (total - tot) / nb;
  long long mine = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    if (scale) want[q] = (int64_t)(((uint64_t)want[q] * factor) >> 32);
    else want[q] = want[q] > 0 ? min(want[q] + extra, cap) : 0;  // want <= cap already
    mine += want[q];
  }
  // exclusive scan of the per-thread sums
  long long incl = mine;
  for (int off = 1; off < 64; off <<= 1) {
    const long long y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  long long run = incl - mine;
  for (int k = 0; k < wv; ++k) run += wsum[k];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int b = tid * per + q;
    if (q < per && b < nb) {
      pool.start[b] = (int32_t)run;
      pool.bcap[b] = (int32_t)want[q];
      pool.cursor[b] = 0;
      run += want[q];
    }
  }
  if (tid < kMaxLevels) pool.maxbits[tid] = 0u;
  if (tid == 0) {
    pool.hdr[0] = kLayoutMagic;
    pool.hdr[1] = nb;
    pool.hdr[2] = pool.total;
    pool.hdr[3] = 0;
  }
}

// DPP row shift / broadcast (gfx9 encodings: row_shr:n = 0x110 + n,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143); lanes without a source read 0.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, kRowMask, 0xf,
                                                    true));
}

template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ uint32_t dpp_u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, kRowMask, 0xf, true);
}

// One butterfly step of a reduce-scatter inside each 16-lane row: a[0..2H)
// -> a[0..H), the lane keeping the upper half when `upper`, plus the copy of
// that half held by its partner (the lane whose `upper` differs, reached by
// the DPP pattern kCtrl: row_mirror, row_half_mirror or a quad permutation).
template <int H, int kCtrl>
__device__ __forceinline__ void rs_step(float* a, bool upper) {
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const float keep = upper ? a[k + H] : a[k];
    const float send = upper ? a[k] : a[k + H];
    a[k] = keep + dpp_f<kCtrl>(send);
  }
}

// Wavefront sums of the N = 8F values a[] (a group's corner contributions,
// zero outside the group), all returned to lane ld's r[c][f] (a[2c + f] for
// F = 2, a[c] for F = 1): a reduce-scatter inside each row leaves lane i
// with value i % N summed over its row (or half-row), the rows' partial
// sums are added across lanes (row rotation by 8 for N = 8, a swizzle
// across 16, a permute across 32), and lane ld reads the N totals.  About
// 3N + 2 N lane operations instead of N full-wavefront reductions.
template <int F>
__device__ __forceinline__ void group_sums(float (&a)[8 * F], int lane, int ld, float (&r)[8][F]) {
  if constexpr (F == 2) rs_step<8, 0x140>(a, lane & 8);  // row_mirror: lane i <-> 15 - i
  rs_step<4, 0x141>(a, lane & 4);                        // row_half_mirror: i <-> 7 - i
  rs_step<2, 0x4E>(a, lane & 2);                         // quad_perm [2,3,0,1]
  rs_step<1, 0xB1>(a, lane & 1);                         // quad_perm [1,0,3,2]
  float v = a[0];
  if constexpr (F == 1) v += dpp_f<0x128>(v);  // row_ror:8
  v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));  // lane ^ 16
  v += __shfl_xor(v, 32, 64);
#pragma unroll
  for (int j = 0; j < 8 * F; ++j) {
    const float t = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
    if (lane == ld) r[j / F][j % F] = t;
  }
}

// Records of one level wait in LDS, ordered by bucket, before they are
// written: a bucket's records from one workgroup are consecutive in its list,
// so the writes of consecutive staging entries are consecutive in memory.
struct StagedRecord {
  uint32_t key;  // row within the level
  int32_t rank;  // position among the workgroup's records of its bucket
  float v[2];
};

// kAd: the records also carry the table's part of the double backward
// through the analytic x-gradient (the term hash_bwd_bwd_kernel adds with
// fp32 atomics): with the x-gradient's cotangent v = gv[i] and the cotangent
// genc2 of the encoding's derivative along v, corner c of level l adds
// scale_l (v . d w_c / d frac) genc2[i, l, f] to row idx_c, feature f.
template <int F, bool kAd = false>
__global__ void __launch_bounds__(kRecThreads, 4)  // 4 wavefronts per SIMD: 2 workgroups per CU
hash_bwd_rec_kernel(const float* __restrict__ x, int64_t n_cap, const int32_t* __restrict__ n_dev,
                    const float* __restrict__ params, HashGridDesc d,
                    const float* __restrict__ genc, float* __restrict__ gparams,
                    float* __restrict__ gx, BucketDesc bd, BucketPool pool,
                    const float* __restrict__ gv = nullptr,
                    const float* __restrict__ genc2 = nullptr) {
  __shared__ int32_t hist[kMaxBucketsPerLevel];   // records per bucket (this level)
  __shared__ int32_t lofs[kMaxBucketsPerLevel];   // their first staging slot
  // per bucket: (list room left at the workgroup's share, pool index of the
  // share's first record)
  __shared__ int2 rres[kMaxBucketsPerLevel];
  __shared__ StagedRecord stage[kRecThreads * 8];
  __shared__ uint32_t lmax;
  __shared__ int32_t n_staged;
  const int tid = threadIdx.x, lane = tid & 63;
  for (int b = tid; b < kMaxBucketsPerLevel; b += kRecThreads) hist[b] = 0;
  if (tid == 0) lmax = 0u;
  __syncthreads();
  const int64_t n = dev_count(n_cap, n_dev);
  const int64_t n_chunks = ceil_div(n, kRecThreads);
  for (int64_t ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {  // uniform per workgroup
    const int64_t i = ch * kRecThreads + tid;
    const bool act = i < n;
    float p[3] = {0.f, 0.f, 0.f};
    if (act) {
#pragma unroll
      for (int k = 0; k < 3; ++k) p[k] = x[3 * i + k];
    }
    const float* g = genc + (act ? i : 0) * (int64_t)(d.n_levels * F);
    float dx[3] = {0.f, 0.f, 0.f};
    float v[3] = {0.f, 0.f, 0.f};
    if (kAd && act) {
#pragma unroll
      for (int k = 0; k < 3; ++k) v[k] = gv[3 * i + k];
    }
    // dL/denc of 4 levels per load when the rows allow 16-byte loads (F = 2):
    // a point's row is read in whole 32-byte pieces instead of 8 bytes per
    // level, which the L2 (4 MB per XCD against ~4 MB of rows in flight per
    // XCD) does not keep between levels
    const bool vec_g = F == 2 && (d.n_levels & 3) == 0 &&
                       (reinterpret_cast<uintptr_t>(genc) & 15) == 0;
    float4 gq0 = make_float4(0.f, 0.f, 0.f, 0.f), gq1 = gq0;
    for (int l = 0, na_ = active_levels(d); l < na_; ++l) {
      float gl[F];
      if (vec_g) {
        if ((l & 3) == 0 && act) {
          const float4* g4 = reinterpret_cast<const float4*>(g + l * F);
          gq0 = g4[0];
          gq1 = g4[1];
        }
        const int k = l & 3;
        const float a0 = k == 0 ? gq0.x : k == 1 ? gq0.z : k == 2 ? gq1.x : gq1.z;
        const float a1 = k == 0 ? gq0.y : k == 1 ? gq0.w : k == 2 ? gq1.y : gq1.w;
        gl[0] = a0;
        gl[F - 1] = a1;
      } else {
#pragma unroll
        for (int f = 0; f < F; ++f) gl[f] = act ? g[l * F + f] : 0.f;
      }
      uint32_t idx[8], cell[3];
      float w[8], fr[3];
      hg_level(p, d, l, idx, w, fr, cell);
      // records of this lane: its own corners, a group's sums (group leader) or none
      float rv[8][F];
#pragma unroll
      for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int f = 0; f < F; ++f) rv[c][f] = w[c] * gl[f];
      if constexpr (kAd) {
        float g2[F];
        const float* g2p = genc2 + (act ? i : 0) * (int64_t)(d.n_levels * F) + l * F;
#pragma unroll
        for (int f = 0; f < F; ++f) g2[f] = act ? g2p[f] : 0.f;
        float dw[8][3];
        hg_dweights(fr, dw);
        const float sc = d.scale[l];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const float a = sc * (v[0] * dw[c][0] + v[1] * dw[c][1] + v[2] * dw[c][2]);
#pragma unroll
          for (int f = 0; f < F; ++f) rv[c][f] += a * g2[f];
        }
      }
      bool has = act;
      uint64_t pend = __ballot(act);
      {
        // the previous lane's cell by a DPP wave shift (wave_shr:1), not an
        // LDS permute; lane 0 reads 0 and is excluded below
        const uint32_t p0 = dpp_u<0x138>(cell[0]), p1 = dpp_u<0x138>(cell[1]),
                       p2 = dpp_u<0x138>(cell[2]);
        const bool same_prev = lane > 0 && ((pend >> (lane - 1)) & 1ull) && act &&
                               cell[0] == p0 && cell[1] == p1 && cell[2] == p2;
        if (__ballot(same_prev) == 0ull) pend = 0ull;  // uniform
      }
      int singles = 0;
      while (pend) {  // uniform
        const int ld = __ffsll((unsigned long long)pend) - 1;
        const uint32_t c0 = __builtin_amdgcn_readlane(cell[0], ld);
        const uint32_t c1 = __builtin_amdgcn_readlane(cell[1], ld);
        const uint32_t c2 = __builtin_amdgcn_readlane(cell[2], ld);
        const bool in = ((pend >> lane) & 1ull) && cell[0] == c0 && cell[1] == c1 && cell[2] == c2;
        const uint64_t m = __ballot(in);
        if (__popcll(m) >= kMinGroup) {
          float a[8 * F];
#pragma unroll
          for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int f = 0; f < F; ++f) a[c * F + f] = in ? rv[c][f] : 0.f;
          group_sums<F>(a, lane, ld, rv);
          if (in && lane != ld) has = false;
          singles = 0;
        } else if (++singles >= kMaxSingles) {
          break;
        }
        pend &= ~m;
      }
      // ranks within the workgroup's share of each bucket
      const int gb0 = bd.base[l], nb = bd.base[l + 1] - gb0;
      int32_t rank[8];
      // largest |value| as bits: the bits of |x| order like |x| for finite
      // values and inf, and every NaN's bits are above inf's, so a NaN or inf
      // reaches the level maximum and the sums fall back to fp32 (fmaxf would
      // drop a NaN)
      uint32_t vmaxb = 0u;
      if (has) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          rank[c] = atomicAdd(&hist[idx[c] >> kBucketShift], 1);
#pragma unroll
          for (int f = 0; f < F; ++f) vmaxb = max(vmaxb, __float_as_uint(fabsf(rv[c][f])));
        }
      }
      // the wavefront's maximum (prefix maxima within rows, then across rows:
      // lane 63 ends with it), one LDS atomic per wavefront instead of one
      // per lane on the same word (record kernel 592 -> 536 us per late launch)
      vmaxb = max(vmaxb, dpp_u<0x111>(vmaxb));
      vmaxb = max(vmaxb, dpp_u<0x112>(vmaxb));
      vmaxb = max(vmaxb, dpp_u<0x114>(vmaxb));
      vmaxb = max(vmaxb, dpp_u<0x118>(vmaxb));
      vmaxb = max(vmaxb, dpp_u<0x142, 0xa>(vmaxb));
      vmaxb = max(vmaxb, dpp_u<0x143, 0xc>(vmaxb));
      if (lane == 63 && vmaxb) atomicMax(&lmax, vmaxb);
      if (gx && act) {
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
      __syncthreads();
      // wavefront 0 alone: staging offsets (exclusive scan of the counts);
      // with one bucket per lane (nb <= 64) it first issues the lane's list
      // reservation (one global atomic per non-empty bucket: position
      // relative to the list), whose result it needs only after the scan,
      // the barrier below and its own staging: the barrier waits for LDS
      // operations only, so the round trip overlaps all three
      int32_t res_c = 0, res_r0 = 0;
      if (tid < 64) {
        const int per = (nb + 63) / 64;  // buckets per lane, in order
        int sum = 0;
        for (int q = 0; q < per; ++q) {
          const int b = tid * per + q;
          if (b < nb) sum += hist[b];
        }
        if (nb <= 64) {
          res_c = sum;
          if (res_c) res_r0 = atomicAdd(&pool.cursor[gb0 + tid], res_c);
        }
        // inclusive scan across the wavefront on DPP lane moves (row shifts,
        // then the rows' totals broadcast upwards), no LDS round trips
        uint32_t incl = (uint32_t)sum;
        incl += dpp_u<0x111>(incl);
        incl += dpp_u<0x112>(incl);
        incl += dpp_u<0x114>(incl);
        incl += dpp_u<0x118>(incl);
        incl += dpp_u<0x142, 0xa>(incl);
        incl += dpp_u<0x143, 0xc>(incl);
        int run = (int)incl - sum;
        for (int q = 0; q < per; ++q) {
          const int b = tid * per + q;
          if (b < nb) {
            lofs[b] = run;
            run += hist[b];
          }
        }
        if (tid == 63) n_staged = (int)incl;
      }
      if (tid == 0) {
        atomicMax(&pool.maxbits[l], lmax);
        lmax = 0u;
      }
      __syncthreads();
      // every wavefront stages its records (the staging needs only the
      // offsets); wavefront 0 then turns the reservations into (room, pool
      // index while in room) per bucket -- with several buckets per lane it
      // takes them here -- and sets the counts back to zero
      auto stage_own = [&]() {
        if (has) {
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            StagedRecord& e = stage[lofs[idx[c] >> kBucketShift] + rank[c]];
            e.key = idx[c];
            e.rank = rank[c];
#pragma unroll
            for (int f = 0; f < F; ++f) e.v[f] = rv[c][f];
          }
        }
      };
      if (tid < 64) {
        if (nb <= 64) {
          stage_own();
          if (res_c) {
            const int32_t room = pool.bcap[gb0 + tid] - res_r0;
            rres[tid] = make_int2(room, room > 0 ? pool.start[gb0 + tid] + res_r0 : 0);
          }
          if (tid < nb) hist[tid] = 0;
        } else {
          stage_own();
          const int per = (nb + 63) / 64;
          for (int q = 0; q < per; ++q) {
            const int b = tid * per + q;
            if (b < nb) {
              const int32_t c = hist[b];
              if (c) {
                const int32_t r0 = atomicAdd(&pool.cursor[gb0 + b], c);
                const int32_t room = pool.bcap[gb0 + b] - r0;
                rres[b] = make_int2(room, room > 0 ? pool.start[gb0 + b] + r0 : 0);
              }
              hist[b] = 0;
            }
          }
        }
      } else {
        stage_own();
      }
      __syncthreads();
      {
        float* tg = gparams + (int64_t)d.offset[l] * F;
        const int ns = n_staged;
        auto put = [&](const StagedRecord& e, int2 rr) {
          if (e.rank < rr.x) {
            const int64_t at = rr.y + e.rank;
            pool.row[at] = (uint16_t)(e.key & (kBucketEntries - 1));
#pragma unroll
            for (int f = 0; f < F; ++f) pool.val[at * F + f] = e.v[f];
          } else {
#pragma unroll
            for (int f = 0; f < F; ++f) atomicAdd(&tg[(int64_t)e.key * F + f], e.v[f]);
          }
        };
        // two records per lane and step, both read from LDS before either is
        // written out (record kernel 527 -> 516 us per late launch)
        int q = tid;
        for (; q + kRecThreads < ns; q += 2 * kRecThreads) {
          const StagedRecord e0 = stage[q], e1 = stage[q + kRecThreads];
          const int2 r0 = rres[e0.key >> kBucketShift], r1 = rres[e1.key >> kBucketShift];
          put(e0, r0);
          put(e1, r1);
        }
        if (q < ns) {
          const StagedRecord e = stage[q];
          put(e, rres[e.key >> kBucketShift]);
        }
      }
    }
    if (gx && act) {
#pragma unroll
      for (int k = 0; k < 3; ++k) gx[3 * i + k] = dx[k];
    }
  }
}

// round(v * 2^K) as a 64-bit integer (round half to even, the result of
// llrint on the exact product), from the bits of v: v = m * 2^(e - 150), so
// v * 2^K = m * 2^(e - 150 + K), a shift of the 24-bit significand.  The
// caller guarantees |v| * 2^K < 2^62 and finite v.  Integer ops only.
__device__ __forceinline__ long long fixed_of(float v, int K) {
  const uint32_t b = __float_as_uint(v);
  const int e = (int)((b >> 23) & 0xffu);
  uint32_t m = b & 0x7fffffu;
  int sh;
  if (e) {
    m |= 0x800000u;
    sh = e - 150 + K;
  } else {
    sh = -149 + K;  // subnormal: m * 2^-149
  }
  unsigned long long q;
  if (sh >= 0) {
    q = (unsigned long long)m << sh;
  } else if (sh < -25) {
    q = 0ull;  // below half of the last unit
  } else {
    const int r = -sh;
    const uint32_t half = 1u << (r - 1);
    const uint32_t rem = m & ((1u << r) - 1u);
    uint32_t t = m >> r;
    if (rem > half || (rem == half && (t & 1u))) ++t;
    q = t;
  }
  return (b >> 31) ? -(long long)q : (long long)q;
}

// 2^e with count * vmax < 2^e (e from the fp32 product; rounding can only
// raise it, which keeps the bound)
__device__ __forceinline__ int bound_exp(float vmax, int32_t count) {
  int e;
  frexpf(vmax * (float)count, &e);
  return e;
}

template <int F>
__global__ void __launch_bounds__(1024)
hash_bwd_bucket_sum_kernel(HashGridDesc d, BucketDesc bd, BucketPool pool,
                           float* __restrict__ gparams) {
  __shared__ unsigned long long acc[kBucketEntries * F];
  const int gb = blockIdx.x, tid = threadIdx.x;
  int l = 0;
  while (l + 1 < d.n_levels && bd.base[l + 1] <= gb) ++l;
  const int32_t nrec = min(pool.cursor[gb], pool.bcap[gb]);
  const float vmax = __uint_as_float(pool.maxbits[l]);
  if (nrec <= 0 || vmax == 0.f) return;  // uniform per workgroup
  const int64_t at0 = pool.start[gb];
  const int64_t e0 = (int64_t)(gb - bd.base[l]) * kBucketEntries;
  const int64_t rows = min((int64_t)kBucketEntries, (int64_t)d.hsize[l] - e0);
  float* tg = gparams + ((int64_t)d.offset[l] + e0) * F;
  if (!(vmax <= 3.402823466e38f)) {
    // an inf or NaN among the level's values: fp32 sums, so that they propagate
    float* accf = reinterpret_cast<float*>(acc);
    for (int j = tid; j < kBucketEntries * F; j += 1024) accf[j] = 0.f;
    __syncthreads();
    for (int r = tid; r < nrec; r += 1024) {
      const int row = pool.row[at0 + r];
#pragma unroll
      for (int f = 0; f < F; ++f) atomicAdd(&accf[row * F + f], pool.val[(at0 + r) * F + f]);
    }
    __syncthreads();
    for (int64_t j = tid; j < rows * F; j += 1024) tg[j] += accf[j];
    return;
  }
  const int K = 61 - bound_exp(vmax, nrec);  // count * max|value| < 2^(61 - K)
  for (int j = tid; j < kBucketEntries * F; j += 1024) acc[j] = 0ull;
  __syncthreads();
  // kU records per lane in flight: their loads are issued before the first
  // LDS atomic (one workgroup per CU, so the loads' latency is hidden here;
  // 8 against 4: 254 -> 251.5 us per late launch, the LDS atomics being the
  // rest of it)
#ifndef SN_SUM_UNROLL
#define SN_SUM_UNROLL 8
#endif
  constexpr int kU = SN_SUM_UNROLL;
  int r = tid;
  for (; r + (kU - 1) * 1024 < nrec; r += kU * 1024) {
    int row[kU];
    float v[kU][F];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      row[u] = pool.row[at0 + r + u * 1024];
#pragma unroll
      for (int f = 0; f < F; ++f) v[u][f] = pool.val[(at0 + r + u * 1024) * F + f];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int f = 0; f < F; ++f)
        atomicAdd(&acc[row[u] * F + f], (unsigned long long)fixed_of(v[u][f], K));
  }
  for (; r < nrec; r += 1024) {
    const int row = pool.row[at0 + r];
#pragma unroll
    for (int f = 0; f < F; ++f)
      atomicAdd(&acc[row * F + f], (unsigned long long)fixed_of(pool.val[(at0 + r) * F + f], K));
  }
  __syncthreads();
  // sum * 2^-K: the int64 -> fp32 conversion rounds once, the scaling is exact
  for (int64_t j = tid; j < rows * F; j += 1024) {
    const long long s = (long long)acc[j];
    if (s) tg[j] += ldexpf((float)s, -K);
  }
}

// Buckets of every level (kMaxBucketsPerLevel per level at most); returns the
// total or -1 when a level has more.
static int64_t make_bucket_desc(const HashGridDesc& d, BucketDesc* bd) {
  int64_t tot = 0;
  for (int l = 0; l < d.n_levels; ++l) {
    bd->base[l] = (int32_t)tot;
    const int64_t nb = ((int64_t)d.hsize[l] + kBucketEntries - 1) / kBucketEntries;
    if (nb > kMaxBucketsPerLevel) return -1;
    tot += nb;
  }
  bd->base[d.n_levels] = (int32_t)tot;
  return tot;
}

// workspace: hdr[4] | cursor[nb] | maxbits[kMaxLevels] | start[nb] | bcap[nb] |
//            row[total] | val[total*F]
static size_t bucket_head_bytes(int64_t n_buckets) {
  return align_up(sizeof(int32_t) * (4 + 3 * n_buckets + kMaxLevels), 256);
}
static size_t bucket_ws_bytes(int64_t n_buckets, int F, int64_t total) {
  return bucket_head_bytes(n_buckets) + align_up(sizeof(uint16_t) * total, 256) +
         align_up(sizeof(float) * F * total, 256);
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
    if (l < active_levels(d)) {
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
                        const int32_t* n_active_dev,
                        float* enc, void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_fwd", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  d.n_active_dev = n_active_dev;
  SN_REQUIRE(n_cap >= 0, "sn_hashgrid_fwd: n < 0");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && params && enc, "sn_hashgrid_fwd: null pointer");
  if (n_features == 2 && reinterpret_cast<uintptr_t>(params) % 8 == 0 &&
      reinterpret_cast<uintptr_t>(enc) % 8 == 0) {
    hash_fwd_kernel<2, float, true><<<grid_for(n_cap, 256, kHashBlocks), 256, 0,
                                      as_stream(stream)>>>(x, n_cap, n_dev, params, d, enc);
    return check_launch("sn_hashgrid_fwd");
  }
  SN_DISPATCH_F(n_features, hash_fwd_kernel<kF><<<grid_for(n_cap, 256, kHashBlocks), 256, 0,
                                                  as_stream(stream)>>>(x, n_cap, n_dev, params,
                                                                       d, enc));
  return check_launch("sn_hashgrid_fwd");
}

int sn_hashgrid_fwd16_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                          const void* params16, int32_t n_levels, int32_t n_features,
                          int32_t log2_hashmap_size, int32_t base_resolution,
                          float per_level_scale, int32_t n_active_levels,
                        const int32_t* n_active_dev, float* enc,
                          void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_fwd16", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  d.n_active_dev = n_active_dev;
  SN_REQUIRE(n_cap >= 0, "sn_hashgrid_fwd16: n < 0");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && params16 && enc, "sn_hashgrid_fwd16: null pointer");
  const _Float16* p16 = reinterpret_cast<const _Float16*>(params16);
  if (n_features == 2 && reinterpret_cast<uintptr_t>(p16) % 4 == 0 &&
      reinterpret_cast<uintptr_t>(enc) % 8 == 0) {
    hash_fwd_kernel<2, _Float16, true><<<grid_for(n_cap, 256, kHashBlocks), 256, 0,
                                         as_stream(stream)>>>(x, n_cap, n_dev, p16, d, enc);
    return check_launch("sn_hashgrid_fwd16");
  }
  SN_DISPATCH_F(n_features, (hash_fwd_kernel<kF, _Float16><<<grid_for(n_cap, 256, kHashBlocks), 256,
                                                              0, as_stream(stream)>>>(
                                x, n_cap, n_dev, p16, d, enc)));
  return check_launch("sn_hashgrid_fwd16");
}

int sn_hashgrid_dx_dev(const float* x, int64_t n_cap, const int32_t* n_dev, const float* params,
                       const void* params16, int32_t n_levels, int32_t n_features,
                       int32_t log2_hashmap_size, int32_t base_resolution,
                       float per_level_scale, int32_t n_active_levels,
                       const int32_t* n_active_dev, const float* dsdf_dh0, int32_t ld,
                       float* grad, void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_dx", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  d.n_active_dev = n_active_dev;
  SN_REQUIRE(n_cap >= 0, "sn_hashgrid_dx: n < 0");
  SN_REQUIRE(ld >= 3 + n_levels * n_features, "sn_hashgrid_dx: row stride %d < 3 + %d", ld,
             n_levels * n_features);
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && (params || params16) && dsdf_dh0 && grad, "sn_hashgrid_dx: null pointer");
  const unsigned blocks = grid_for(n_cap, 256, kHashBlocks);
  hipStream_t s = as_stream(stream);
  if (n_features == 2 && n_levels % 4 == 0) {
    if (params16)
      hash_dx2_kernel<_Float16><<<blocks, 256, 0, s>>>(
          x, n_cap, n_dev, reinterpret_cast<const _Float16*>(params16), d, dsdf_dh0, ld, grad);
    else
      hash_dx2_kernel<float><<<blocks, 256, 0, s>>>(x, n_cap, n_dev, params, d, dsdf_dh0, ld,
                                                    grad);
    return check_launch("sn_hashgrid_dx");
  }
  if (params16) {
    const _Float16* p16 = reinterpret_cast<const _Float16*>(params16);
    SN_DISPATCH_F(n_features, (hash_dx_kernel<kF, _Float16><<<blocks, 256, 0, s>>>(
                                  x, n_cap, n_dev, p16, d, dsdf_dh0, ld, grad)));
  } else {
    SN_DISPATCH_F(n_features, (hash_dx_kernel<kF, float><<<blocks, 256, 0, s>>>(
                                  x, n_cap, n_dev, params, d, dsdf_dh0, ld, grad)));
  }
  return check_launch("sn_hashgrid_dx");
}

int sn_hashgrid_fwd(const float* x, int64_t n, const float* params, int32_t n_levels,
                    int32_t n_features, int32_t log2_hashmap_size, int32_t base_resolution,
                    float per_level_scale, int32_t n_active_levels, float* enc, void* stream) {
  return sn_hashgrid_fwd_dev(x, n, nullptr, params, n_levels, n_features, log2_hashmap_size,
                             base_resolution, per_level_scale, n_active_levels, nullptr, enc,
                             stream);
}

int64_t sn_hashgrid_bwd_workspace_bytes(int32_t n_levels, int32_t n_features,
                                        int32_t log2_hashmap_size, int32_t base_resolution,
                                        float per_level_scale, int32_t records_per_bucket) {
  HashGridDesc d;
  BucketDesc bd;
  if (records_per_bucket <= 0 || n_features <= 0 || n_features > 2 ||
      make_hashgrid_desc(n_levels, log2_hashmap_size, base_resolution, per_level_scale,
                         n_levels, &d) < 0)
    return -1;
  const int64_t nb = make_bucket_desc(d, &bd);
  if (nb < 0) return -1;
  const int64_t want = (int64_t)records_per_bucket * nb;
  const int64_t total = want < kMaxPoolRecords ? want : kMaxPoolRecords;
  return (int64_t)bucket_ws_bytes(nb, n_features, total);
}

}  // extern "C"

// The backward to the table (and x): bucket records with a workspace (with
// gv / genc2 also the double backward's table term), else the atomic paths;
// returns 1 in *bucketed when the records carried gv / genc2.
static int hashgrid_bwd_impl(const float* x, int64_t n_cap, const int32_t* n_dev,
                             const float* params, HashGridDesc d, int32_t n_features,
                             const float* grad_enc, const float* gv, const float* genc2,
                             float* grad_params, float* grad_x, void* workspace,
                             size_t workspace_bytes, hipStream_t s, int* bucketed) {
  int rc = SN_OK;
  *bucketed = 0;
  const unsigned grid = grid_for(n_cap, 256, kHashBlocks);
  if (g_aggregate == 2 && grad_params && n_features <= 2) {
    // bucket records when a workspace is given (and the 32-bit record
    // counters cannot overflow), else atomics from the workgroup merge
    BucketDesc bd{};
    const int64_t nb = make_bucket_desc(d, &bd);
    const bool buckets = workspace && nb > 0 && n_cap < (1ll << 28) - 4096 && d.n_active > 0;
    if (buckets) {
      const size_t head = bucket_head_bytes(nb);
      const size_t per = sizeof(uint16_t) + sizeof(float) * n_features;
      int64_t total =
          workspace_bytes > head + 512 ? (int64_t)((workspace_bytes - head - 512) / per) : 0;
      if (total > kMaxPoolRecords) total = kMaxPoolRecords;
      while (total > 0 && bucket_ws_bytes(nb, n_features, total) > workspace_bytes) --total;
      if (total < nb) {
        set_error("sn_hashgrid_bwd: workspace of %zu bytes holds less than one record per bucket",
                  workspace_bytes);
        return SN_ESPACE;
      }
      char* w = static_cast<char*>(workspace);
      BucketPool pool;
      pool.hdr = reinterpret_cast<int32_t*>(w);
      pool.cursor = pool.hdr + 4;
      pool.maxbits = reinterpret_cast<uint32_t*>(pool.cursor + nb);
      pool.start = pool.cursor + nb + kMaxLevels;
      pool.bcap = pool.start + nb;
      pool.row = reinterpret_cast<uint16_t*>(w + head);
      pool.val = reinterpret_cast<float*>(w + head + align_up(sizeof(uint16_t) * total, 256));
      pool.total = (int32_t)total;
      const int32_t nb_active = bd.base[d.n_active];
      bucket_layout_kernel<<<1, kLayoutThreads, 0, s>>>(pool, (int32_t)nb);
      rc = check_launch("sn_hashgrid_bwd");
      if (rc) return rc;
      const unsigned rgrid = grid_for(n_cap, kRecThreads, kHashBlocks);
      const bool ad = gv && genc2;
      if (n_features == 1 && ad)
        hash_bwd_rec_kernel<1, true><<<rgrid, kRecThreads, 0, s>>>(
            x, n_cap, n_dev, params, d, grad_enc, grad_params, grad_x, bd, pool, gv, genc2);
      else if (n_features == 1)
        hash_bwd_rec_kernel<1><<<rgrid, kRecThreads, 0, s>>>(x, n_cap, n_dev, params, d,
                                                             grad_enc, grad_params, grad_x, bd,
                                                             pool);
      else if (ad)
        hash_bwd_rec_kernel<2, true><<<rgrid, kRecThreads, 0, s>>>(
            x, n_cap, n_dev, params, d, grad_enc, grad_params, grad_x, bd, pool, gv, genc2);
      else
        hash_bwd_rec_kernel<2><<<rgrid, kRecThreads, 0, s>>>(x, n_cap, n_dev, params, d,
                                                             grad_enc, grad_params, grad_x, bd,
                                                             pool);
      *bucketed = ad ? 1 : 0;
      rc = check_launch("sn_hashgrid_bwd");
      if (rc) return rc;
      if (n_features == 1)
        hash_bwd_bucket_sum_kernel<1><<<nb_active, 1024, 0, s>>>(d, bd, pool, grad_params);
      else
        hash_bwd_bucket_sum_kernel<2><<<nb_active, 1024, 0, s>>>(d, bd, pool, grad_params);
      return check_launch("sn_hashgrid_bwd");
    }
    // (one launch over all levels: measured faster than a launch per level)
    if (n_features == 1)
      hash_bwd_lds_kernel<1><<<grid, 256, 0, s>>>(x, n_cap, n_dev, params, d, grad_enc,
                                                  grad_params, grad_x);
    else
      hash_bwd_lds_kernel<2><<<grid, 256, 0, s>>>(x, n_cap, n_dev, params, d, grad_enc,
                                                  grad_params, grad_x);
    return check_launch("sn_hashgrid_bwd");
  }
  const int agg = g_aggregate >= 1 ? 1 : 0;
  SN_DISPATCH_F(n_features,
                hash_bwd_kernel<kF><<<grid, 256, 0, s>>>(x, n_cap, n_dev, params, d, grad_enc,
                                                         grad_params, grad_x, agg));
  return check_launch("sn_hashgrid_bwd");
}

extern "C" {

int sn_hashgrid_bwd_ws_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                           const float* params, int32_t n_levels, int32_t n_features,
                           int32_t log2_hashmap_size, int32_t base_resolution,
                           float per_level_scale, int32_t n_active_levels,
                           const int32_t* n_active_dev, const float* grad_enc,
                           float* grad_params, float* grad_x, void* workspace,
                           size_t workspace_bytes, void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_bwd", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  d.n_active_dev = n_active_dev;
  SN_REQUIRE(n_cap >= 0, "sn_hashgrid_bwd: n < 0");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && params && grad_enc, "sn_hashgrid_bwd: null input");
  SN_REQUIRE(grad_params || grad_x, "sn_hashgrid_bwd: no output");
  int bucketed = 0;
  return hashgrid_bwd_impl(x, n_cap, n_dev, params, d, n_features, grad_enc, nullptr, nullptr,
                           grad_params, grad_x, workspace, workspace_bytes, as_stream(stream),
                           &bucketed);
}

int sn_hashgrid_bwd_ad_ws_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                              const float* params, int32_t n_levels, int32_t n_features,
                              int32_t log2_hashmap_size, int32_t base_resolution,
                              float per_level_scale, int32_t n_active_levels,
                              const int32_t* n_active_dev, const float* grad_enc,
                              const float* g_gx, const float* grad_enc2, float* grad_params,
                              void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_bwd_ad", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  d.n_active_dev = n_active_dev;
  SN_REQUIRE(n_cap >= 0, "sn_hashgrid_bwd_ad: n < 0");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && params && grad_enc && g_gx && grad_enc2 && grad_params,
             "sn_hashgrid_bwd_ad: null argument");
  hipStream_t s = as_stream(stream);
  int bucketed = 0;
  rc = hashgrid_bwd_impl(x, n_cap, n_dev, params, d, n_features, grad_enc, g_gx, grad_enc2,
                         grad_params, nullptr, workspace, workspace_bytes, s, &bucketed);
  if (rc || bucketed) return rc;
  // no bucket path (no workspace, F > 2): the double backward's table term
  // with fp32 atomics
  SN_DISPATCH_F(n_features, hash_bwd_bwd_kernel<kF><<<grid_for(n_cap, 256, kHashBlocks), 256, 0,
                                                      s>>>(x, n_cap, n_dev, params, d,
                                                           grad_enc2, g_gx, nullptr,
                                                           grad_params, nullptr));
  return check_launch("sn_hashgrid_bwd_ad");
}

int sn_hashgrid_bwd_dev(const float* x, int64_t n_cap, const int32_t* n_dev, const float* params,
                        int32_t n_levels, int32_t n_features, int32_t log2_hashmap_size,
                        int32_t base_resolution, float per_level_scale, int32_t n_active_levels,
                        const int32_t* n_active_dev,
                        const float* grad_enc, float* grad_params, float* grad_x, void* stream) {
  return sn_hashgrid_bwd_ws_dev(x, n_cap, n_dev, params, n_levels, n_features, log2_hashmap_size,
                                base_resolution, per_level_scale, n_active_levels, n_active_dev,
                                grad_enc,
                                grad_params, grad_x, nullptr, 0, stream);
}

int sn_hashgrid_bwd(const float* x, int64_t n, const float* params, int32_t n_levels,
                    int32_t n_features, int32_t log2_hashmap_size, int32_t base_resolution,
                    float per_level_scale, int32_t n_active_levels, const float* grad_enc,
                    float* grad_params, float* grad_x, void* stream) {
  return sn_hashgrid_bwd_dev(x, n, nullptr, params, n_levels, n_features, log2_hashmap_size,
                             base_resolution, per_level_scale, n_active_levels, nullptr, grad_enc,
                             grad_params, grad_x, stream);
}

int sn_hashgrid_bwd_bwd_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                            const float* params, int32_t n_levels, int32_t n_features,
                            int32_t log2_hashmap_size, int32_t base_resolution,
                            float per_level_scale, int32_t n_active_levels,
                        const int32_t* n_active_dev,
                            const float* grad_enc, const float* g_gx, float* grad_grad_enc,
                            float* grad_params, float* grad_x, void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_bwd_bwd", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  d.n_active_dev = n_active_dev;
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
                                 base_resolution, per_level_scale, n_active_levels, nullptr,
                                 grad_enc,
                                 g_gx, grad_grad_enc, grad_params, grad_x, stream);
}

int sn_hashgrid_tangent_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                            const float* params, const void* params16, int32_t n_levels,
                            int32_t n_features, int32_t log2_hashmap_size,
                            int32_t base_resolution, float per_level_scale,
                            int32_t n_active_levels, const int32_t* n_active_dev,
                            const float* g_gx, float* grad_grad_enc, void* stream) {
  clear_error();
  HashGridDesc d;
  int rc = hg_setup("sn_hashgrid_tangent", n_levels, n_features, log2_hashmap_size,
                    base_resolution, per_level_scale, n_active_levels, &d);
  if (rc) return rc;
  d.n_active_dev = n_active_dev;
  SN_REQUIRE(n_cap >= 0, "sn_hashgrid_tangent: n < 0");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && (params || params16) && g_gx && grad_grad_enc,
             "sn_hashgrid_tangent: null argument");
  hipStream_t s = as_stream(stream);
  const bool paired = n_features == 2 && n_levels % 4 == 0 &&
                      (reinterpret_cast<uintptr_t>(grad_grad_enc) & 15) == 0;
  if (!paired) {
    SN_REQUIRE(params, "sn_hashgrid_tangent: this config needs the fp32 table");
    SN_DISPATCH_F(n_features, hash_bwd_bwd_kernel<kF><<<grid_for(n_cap, 256, kHashBlocks), 256,
                                                        0, s>>>(x, n_cap, n_dev, params, d,
                                                                nullptr, g_gx, grad_grad_enc,
                                                                nullptr, nullptr));
    return check_launch("sn_hashgrid_tangent");
  }
  if (params16)
    hash_tangent_kernel<_Float16><<<grid_for(n_cap, 256, kHashBlocks), 256, 0, s>>>(
        x, n_cap, n_dev, static_cast<const _Float16*>(params16), d, g_gx, grad_grad_enc);
  else
    hash_tangent_kernel<float><<<grid_for(n_cap, 256, kHashBlocks), 256, 0, s>>>(
        x, n_cap, n_dev, params, d, g_gx, grad_grad_enc);
  return check_launch("sn_hashgrid_tangent");
}

}  // extern "C"
