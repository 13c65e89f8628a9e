// The SDF MLP of models/fields.py:85-99 on the fp16/bf16 matrix cores with
// fp32 accumulation, activations kept in registers between layers.
//
// Layout.  Every layer is computed transposed, Z^T = W * H^T: units are the
// MFMA rows (M) and points the columns (N), 16 points per column block.  With
// v_mfma_f32_16x16x32 a lane (r = l % 16, q = l / 16) receives column r of
// the B operand (8 K-slots) and returns rows 4q..4q+3 of column r of D.  So a
// lane ends a layer holding point r's units {16 mb + 4q + e}, and those are
// exactly 8 K-slots of the next layer's B operand if the next layer's K axis
// is ordered
//     slot (kb, q, j) -> unit 16 (2 kb + j / 4) + 4 q + j % 4 .
// A dot product does not depend on the order of K as long as A and B agree,
// so the weights (the A operand) are stored in that order once per launch and
// no activation goes through LDS in the forward pass.
//
// Precision.  Weights are rounded once to the MFMA type (a fixed, smooth
// perturbation of the network, below 2^-11 relative for fp16); activations
// and gradients are split into hi + lo parts (x = hi + lo exactly to ~22 bits
// for fp16, ~16 bits for bf16) and both parts are multiplied, so no per-point
// rounding noise enters the SDF at fp16 resolution.  The forward pass runs on
// fp16; the backward on bf16, whose range holds the small loss gradients
// without scaling, with single-part operands (gradients need ~2^-9 per term,
// not the forward's smoothness), recomputing the forward from hi parts only.
//
// Weight gradients need the points on the K axis, i.e. the transpose of the
// register layout: each group's gradient and activation blocks are written to
// a workgroup LDS tile (bf16, rounded to nearest) and read back as 16x16x16
// MFMA operands.  Per-workgroup partial sums are reduced in a fixed order by
// mlp_reduce_rows, so results do not depend on scheduling.
//
// h0 = [x, enc] (input_concat) is held internally as [enc | x | 1 | 0 ...]
// in KB0 blocks of 32: the constant 1 carries the first bias through the
// MFMA (and its gradient out of the weight-gradient product).  The packed
// parameter layout is the fp32 kernels' one:
//   W0[H][3+E] b0[H] { W1[H][H] b1[H] } wo[H] bo
#include "sn_common.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

namespace sn {
namespace m16 {

constexpr int kWaves = 4;
constexpr float kSpThr = 20.f;

__device__ __forceinline__ floatx4 mma_h(half8 a, half8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mma_b(bf16x8 a, bf16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mma_b4(bf16x4 a, bf16x4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void split_h(float v, _Float16& hi, _Float16& lo) {
  hi = (_Float16)v;
  lo = (_Float16)(v - (float)hi);
}
__device__ __forceinline__ void split_b(float v, __bf16& hi, __bf16& lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

// Softplus(beta) with torch's threshold (value z above it) and its derivative
// sigma(beta z) = t / (1 + t), t = e^(beta z).  Branch-free (both sides are
// computed and one selected, so a wavefront never splits) on the hardware
// exp2 / log2 / reciprocal: the backward's recompute, whose values enter bf16
// products only.  A NaN z gives NaN value and derivative.
struct Sp {
  float h, d;
};
__device__ __forceinline__ Sp softplus(float z, float beta, float inv_beta) {
  constexpr float kLog2e = 1.4426950408889634f, kLn2 = 0.6931471805599453f;
  const float bz = z * beta;
  const float t = __builtin_amdgcn_exp2f(bz * kLog2e);
  const float u = 1.f + t;
  const bool lin = bz > kSpThr;
  Sp s;
  s.h = lin ? z : __builtin_amdgcn_logf(u) * (kLn2 * inv_beta);
  s.d = lin ? 1.f : t * __builtin_amdgcn_rcpf(u);
  return s;
}
__device__ __forceinline__ float softplus_h(float z, float beta, float inv_beta) {
  const float bz = z * beta;
  return bz > kSpThr ? z : __logf(1.f + __expf(bz)) * inv_beta;
}

// unit held in slot (kb, q, j) of a hidden-layer K axis
__host__ __device__ constexpr int hid_unit(int kb, int q, int j) {
  return 16 * (2 * kb + (j >> 2)) + 4 * q + (j & 3);
}

template <int H, int NH, int KB0>
struct Cfg {
  static constexpr int MB = H / 16;    // 16-unit blocks of a hidden layer
  static constexpr int KBH = H / 32;   // 32-unit K blocks of a hidden layer
  static constexpr int MD0 = 2 * KB0;  // 16-dim blocks of h0
  // LDS fragments (64 lanes x 16 bytes each)
  static constexpr int nW0 = MB * KB0, nW1 = (NH == 2 ? MB * KBH : 0);
  static constexpr int nW1T = (NH == 2 ? MB * KBH : 0), nW0T = MD0 * KBH;
  static constexpr int kVecFloats = 3 * H + 4;  // b0, b1, wo, bo
  static constexpr size_t fwd_lds = (size_t)(nW0 + nW1) * 64 * 16 + kVecFloats * 4;
  // two block tiles of 64 rows x (64 points + 4) bf16
  static constexpr size_t tile_bytes = 2 * 64 * (16 * kWaves + 4) * 2;
  static constexpr size_t bwd_lds =
      (size_t)(nW0 + nW1 + nW1T + nW0T) * 64 * 16 + kVecFloats * 4 + tile_bytes;
  // forward with the gradient d sdf / d h0: all fragments, no tiles
  static constexpr size_t fwdg_lds =
      (size_t)(nW0 + nW1 + nW1T + nW0T) * 64 * 16 + kVecFloats * 4;
  // forward-over-reverse backward: four block tiles (cotangents and tangents)
  static constexpr size_t bwdad_lds =
      (size_t)(nW0 + nW1 + nW1T + nW0T) * 64 * 16 + kVecFloats * 4 + 2 * tile_bytes;
  __host__ __device__ static int n_params(int d) {
    return H * d + H + (NH == 2 ? H * H + H : 0) + H + 1;
  }
  __host__ __device__ static int gB0(int d) { return H * d; }
  __host__ __device__ static int gW1(int d) { return H * d + H; }
  __host__ __device__ static int gB1(int d) { return H * d + H + H * H; }
  __host__ __device__ static int gWo(int d) { return H * d + H + (NH == 2 ? H * H + H : 0); }
  __host__ __device__ static int gBo(int d) { return gWo(d) + H; }
};

// packed W0 column of internal h0 dim c ([enc | x | 1 | pad]); -2 for the
// bias dim, -1 for padding
__device__ __forceinline__ int h0_col(int c, int E) {
  return c < E ? 3 + c : (c < E + 3 ? c - E : (c == E + 3 ? -2 : -1));
}

// Builds the A-operand fragments and bias vectors in LDS; with kBwd also the
// transposed weights (for products W^T g), in type TT (bf16 for the training
// backward, fp16 for the forward's gradient).
template <int H, int NH, int KB0, bool kBwd, typename TT = __bf16>
__device__ void build_frags(char* smem, const float* __restrict__ mlp, int E) {
  using C = Cfg<H, NH, KB0>;
  const int d = 3 + E;
  _Float16* fW0 = reinterpret_cast<_Float16*>(smem);
  _Float16* fW1 = fW0 + C::nW0 * 512;
  TT* fW1T = reinterpret_cast<TT*>(fW1 + C::nW1 * 512);
  TT* fW0T = fW1T + (kBwd ? C::nW1T * 512 : 0);
  float* vec = reinterpret_cast<float*>(fW0T + (kBwd ? C::nW0T * 512 : 0));
  const float* W0 = mlp;
  const float* W1 = mlp + C::gW1(d);
  // element e of a fragment: lane = e / 8 (r = lane % 16, q = lane / 16), slot j = e % 8
  for (int t = threadIdx.x; t < C::nW0 * 512; t += blockDim.x) {
    const int f = t / 512, e = t % 512, lane = e >> 3, j = e & 7, r = lane & 15, q = lane >> 4;
    const int mb = f / KB0, kb = f % KB0;
    const int col = h0_col(32 * kb + 8 * q + j, E), o = 16 * mb + r;
    fW0[t] = (_Float16)(col >= 0 ? W0[o * d + col] : (col == -2 ? mlp[C::gB0(d) + o] : 0.f));
  }
  for (int t = threadIdx.x; t < C::nW1 * 512; t += blockDim.x) {
    const int f = t / 512, e = t % 512, lane = e >> 3, j = e & 7, r = lane & 15, q = lane >> 4;
    const int mb = f / C::KBH, kb = f % C::KBH;
    fW1[t] = (_Float16)W1[(16 * mb + r) * H + hid_unit(kb, q, j)];
  }
  if (kBwd) {
    for (int t = threadIdx.x; t < C::nW1T * 512; t += blockDim.x) {
      const int f = t / 512, e = t % 512, lane = e >> 3, j = e & 7, r = lane & 15, q = lane >> 4;
      const int mi = f / C::KBH, kb = f % C::KBH;
      fW1T[t] = (TT)W1[hid_unit(kb, q, j) * H + 16 * mi + r];
    }
    for (int t = threadIdx.x; t < C::nW0T * 512; t += blockDim.x) {
      const int f = t / 512, e = t % 512, lane = e >> 3, j = e & 7, r = lane & 15, q = lane >> 4;
      const int md = f / C::KBH, kb = f % C::KBH;
      const int col = h0_col(16 * md + r, E);
      fW0T[t] = (TT)(col >= 0 ? W0[hid_unit(kb, q, j) * d + col] : 0.f);
    }
  }
  for (int t = threadIdx.x; t < H; t += blockDim.x) {
    vec[t] = mlp[C::gB0(d) + t];
    vec[H + t] = NH == 2 ? mlp[C::gB1(d) + t] : 0.f;
    vec[2 * H + t] = mlp[C::gWo(d) + t];
  }
  if (threadIdx.x == 0) vec[3 * H] = mlp[C::gBo(d)];
}

// The 8 h0 values of K block kb for point p (lane slots 8q..8q+7).  With
// vec4 (E % 4 == 0, 16-byte rows) the slots come in two 4-dim groups, each
// either 4 encoding dims (one 16-byte load), the group [x, 1] that follows
// them, or padding; otherwise one guarded load per slot.
// `one` is the value of the constant slot: 1 for h0, 0 for a tangent of h0.
__device__ __forceinline__ void load_h0(const float* __restrict__ x, const float* __restrict__ enc,
                                        int E, bool vec4, int64_t p, bool ok, int kb, int q,
                                        float (&v)[8], float one = 1.f) {
  const int c0 = 32 * kb + 8 * q;
  if (vec4) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int c4 = c0 + 4 * h;
      float4 a = {0.f, 0.f, 0.f, 0.f};
      if (ok && c4 < E) {
        a = *reinterpret_cast<const float4*>(enc + p * E + c4);
      } else if (ok && c4 == E) {
        a = float4{x[p * 3], x[p * 3 + 1], x[p * 3 + 2], one};
      }
      v[4 * h] = a.x;
      v[4 * h + 1] = a.y;
      v[4 * h + 2] = a.z;
      v[4 * h + 3] = a.w;
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = c0 + j;
    const float* src = c < E ? enc + p * E + c : x + p * 3 + (c - E);
    float t = (ok && c == E + 3) ? one : 0.f;
    if (ok && c < E + 3) t = *src;
    v[j] = t;
  }
}

// B operand of a hidden layer from the previous layer's D blocks: slot j of K
// block kb is unit 16 (2 kb + j/4) + 4q + j%4 = block 2kb + j/4, element j%4.
template <int MB>
__device__ __forceinline__ void hid_operand_h(const float (&h)[MB][4], int kb, half8& hi, half8& lo) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    _Float16 a, b;
    split_h(h[2 * kb + (j >> 2)][j & 3], a, b);
    hi[j] = a;
    lo[j] = b;
  }
}
template <int MB>
__device__ __forceinline__ void hid_operand_b(const float (&h)[MB][4], int kb, bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    __bf16 a, b;
    split_b(h[2 * kb + (j >> 2)][j & 3], a, b);
    hi[j] = a;
    lo[j] = b;
  }
}

__device__ __forceinline__ float sum_q(float v) {  // over the 4 lanes of a point (same r)
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}
__device__ __forceinline__ float sum_r(float v) {  // over the 16 points of a q group
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  v += __shfl_xor(v, 8, 64);
  return v;
}

// Forward of one 16-point group: softplus values/derivatives of the first
// hidden layer kept (F), those of the last hidden layer handed unit by unit
// to last(mb, e, value, derivative) as they are made, so they never occupy
// registers together.  Shared by the forward and backward kernels.
template <int H, int NH, int KB0>
struct Fwd {
  static constexpr int MB = H / 16;
  float h1[MB][4], d1[MB][4];  // softplus(z0), softplus'(z0) (NH == 2)
};

// kLo: also multiply the lo parts of the activations (the forward whose values
// are used); the backward's recompute needs only the hi parts.
template <int H, int NH, int KB0, bool kDeriv, bool kLo, typename Last>
__device__ __forceinline__ void forward_group(const half8* fW0, const half8* fW1, const float* vec,
                                              const float (&h0)[KB0][8], float beta, float inv_beta,
                                              Fwd<H, NH, KB0>& F, Last&& last) {
  using C = Cfg<H, NH, KB0>;
  const int l = threadIdx.x & 63, q = l >> 4;
  half8 bh[KB0], bl[KB0];
#pragma unroll
  for (int kb = 0; kb < KB0; ++kb)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      _Float16 a, b;
      split_h(h0[kb][j], a, b);
      bh[kb][j] = a;
      bl[kb][j] = b;
    }
#pragma unroll
  for (int mb = 0; mb < C::MB; ++mb) {
    floatx4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB0; ++kb) {
      const half8 w = fW0[(mb * KB0 + kb) * 64 + l];
      if (kLo) z = mma_h(w, bl[kb], z);
      z = mma_h(w, bh[kb], z);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float zz = z[e];  // b0 enters through the constant h0 dim
      Sp sp;
      if (kDeriv) sp = softplus(zz, beta, inv_beta);
      else sp = Sp{softplus_h(zz, beta, inv_beta), 0.f};
      if (NH == 2) {
        F.h1[mb][e] = sp.h;
        F.d1[mb][e] = sp.d;
      } else {
        last(mb, e, sp.h, sp.d);
      }
    }
  }
  if (NH == 2) {
    half8 ch[C::KBH], cl[C::KBH];
#pragma unroll
    for (int kb = 0; kb < C::KBH; ++kb) hid_operand_h<C::MB>(F.h1, kb, ch[kb], cl[kb]);
#pragma unroll
    for (int mb = 0; mb < C::MB; ++mb) {
      floatx4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < C::KBH; ++kb) {
        const half8 w = fW1[(mb * C::KBH + kb) * 64 + l];
        if (kLo) z = mma_h(w, cl[kb], z);
        z = mma_h(w, ch[kb], z);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float zz = z[e] + vec[H + 16 * mb + 4 * q + e];
        if (kDeriv) {
          const Sp sp = softplus(zz, beta, inv_beta);
          last(mb, e, sp.h, sp.d);
        } else {
          last(mb, e, softplus_h(zz, beta, inv_beta), 0.f);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// forward: sdf [n]
// ---------------------------------------------------------------------------
template <int H, int NH, int KB0>
__global__ void __launch_bounds__(64 * kWaves)
mlp16_fwd_kernel(const float* __restrict__ x, const float* __restrict__ enc, int E, int vec4,
                 int64_t n_cap,
                 const int32_t* __restrict__ n_dev, const float* __restrict__ mlp, float beta,
                 float* __restrict__ sdf_out) {
  using C = Cfg<H, NH, KB0>;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  build_frags<H, NH, KB0, false>(smem, mlp, E);
  __syncthreads();
  const half8* fW0 = reinterpret_cast<const half8*>(smem);
  const half8* fW1 = fW0 + C::nW0 * 64;
  const float* vec = reinterpret_cast<const float*>(fW1 + C::nW1 * 64);
  const float inv_beta = 1.f / beta;
  const int64_t n = dev_count(n_cap, n_dev);
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const int64_t n_groups = ceil_div(n, 16);
  const int64_t waves = (int64_t)gridDim.x * kWaves;
  for (int64_t g = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6); g < n_groups; g += waves) {
    // re-read the weight fragments from LDS every group rather than pinning
    // them in registers: ~80 VGPRs instead of ~200, three times the waves
    asm volatile("" ::: "memory");
    const int64_t p = g * 16 + r;
    const bool ok = p < n;
    float h0[KB0][8];
#pragma unroll
    for (int kb = 0; kb < KB0; ++kb) load_h0(x, enc, E, vec4, p, ok, kb, q, h0[kb]);
    Fwd<H, NH, KB0> F;
    float s = 0.f;
    forward_group<H, NH, KB0, false, true>(
        fW0, fW1, vec, h0, beta, inv_beta, F,
        [&](int mb, int e, float h, float) { s += vec[2 * H + 16 * mb + 4 * q + e] * h; });
    s = sum_q(s) + vec[3 * H];
    if (q == 0 && ok) sdf_out[p] = s;
  }
}

// ---------------------------------------------------------------------------
// forward with the analytic gradient: sdf [n] and d sdf / d h0 [n, 3 + E]
// (packed column order [x, enc], the fp32 kernel's), for the rendered normal
// grad_x sdf = dsdf/dx + sum_l dsdf/denc_l denc_l/dx (fields.py:107-119).  The
// forward as in mlp16_fwd_kernel (fp16 weights, hi + lo activations); the
// backprop of 1 through the same fp16 weights (transposed fragments), with
// the cotangents split into hi + lo fp16 parts, so the gradient is that of
// the network the sdf value comes from to the same resolution.
// ---------------------------------------------------------------------------
template <int H, int NH, int KB0>
__global__ void __launch_bounds__(64 * kWaves)
mlp16_fwd_grad_kernel(const float* __restrict__ x, const float* __restrict__ enc, int E, int vec4,
                      int64_t n_cap, const int32_t* __restrict__ n_dev,
                      const float* __restrict__ mlp, float beta, float* __restrict__ sdf_out,
                      float* __restrict__ dh0) {
  using C = Cfg<H, NH, KB0>;
  constexpr int MB = C::MB, KBH = C::KBH, MD0 = C::MD0;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  build_frags<H, NH, KB0, true, _Float16>(smem, mlp, E);
  __syncthreads();
  const half8* fW0 = reinterpret_cast<const half8*>(smem);
  const half8* fW1 = fW0 + C::nW0 * 64;
  const half8* fW1T = fW1 + C::nW1 * 64;
  const half8* fW0T = fW1T + C::nW1T * 64;
  const float* vec = reinterpret_cast<const float*>(fW0T + C::nW0T * 64);
  const float inv_beta = 1.f / beta;
  const int64_t n = dev_count(n_cap, n_dev);
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const int d = 3 + E;
  const int md_used = (E + 4 + 15) / 16;
  const int64_t n_groups = ceil_div(n, 16);
  const int64_t waves = (int64_t)gridDim.x * kWaves;
  for (int64_t g = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6); g < n_groups; g += waves) {
    asm volatile("" ::: "memory");  // fragments re-read from LDS each group
    const int64_t p = g * 16 + r;
    const bool ok = p < n;
    int qs = q, Es = E, v4s = vec4;
    asm volatile("" : "+v"(qs), "+s"(Es), "+s"(v4s));
    float h0[KB0][8];
#pragma unroll
    for (int kb = 0; kb < KB0; ++kb) load_h0(x, enc, Es, v4s, p, ok, kb, qs, h0[kb]);
    Fwd<H, NH, KB0> F;
    float s = 0.f, gz[MB][4];
    forward_group<H, NH, KB0, true, true>(
        fW0, fW1, vec, h0, beta, inv_beta, F, [&](int mb, int e, float h, float dd) {
          const float wo = vec[2 * H + 16 * mb + 4 * q + e];
          s += wo * h;
          gz[mb][e] = wo * dd;  // d sdf / d z of the last hidden layer
        });
    s = sum_q(s) + vec[3 * H];
    if (q == 0 && ok) sdf_out[p] = s;
    if (NH == 2) {
      half8 gh[KBH], gl[KBH];
#pragma unroll
      for (int kb = 0; kb < KBH; ++kb) hid_operand_h<MB>(gz, kb, gh[kb], gl[kb]);
#pragma unroll
      for (int mi = 0; mi < MB; ++mi) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KBH; ++kb) {
          const half8 w = fW1T[(mi * KBH + kb) * 64 + l];
          acc = mma_h(w, gl[kb], acc);
          acc = mma_h(w, gh[kb], acc);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) gz[mi][e] = acc[e] * F.d1[mi][e];
      }
    }
    half8 gh[KBH], gl[KBH];
#pragma unroll
    for (int kb = 0; kb < KBH; ++kb) hid_operand_h<MB>(gz, kb, gh[kb], gl[kb]);
#pragma unroll
    for (int md = 0; md < MD0; ++md) {
      if (md >= md_used) break;
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KBH; ++kb) {
        const half8 w = fW0T[(md * KBH + kb) * 64 + l];
        acc = mma_h(w, gl[kb], acc);
        acc = mma_h(w, gh[kb], acc);
      }
      // internal dims 16 md + 4q + e of point p, in the packed column order
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int col = h0_col(16 * md + 4 * qs + e, Es);
        if (ok && col >= 0) dh0[p * d + col] = acc[e];
      }
    }
  }
}

// Block tiles: rows of bf16 over the 64 points of a block step (wavefront w
// owns columns 16w..16w+15).  Row pitch 68 bf16 (34 dwords): the writes of a
// wavefront (4 rows 4 apart, 16 columns each) fall on 32 distinct banks, and
// a 16-lane quarter's 8-byte operand reads (16 rows) on 32 distinct banks.
constexpr int kBlkPts = 16 * kWaves;
constexpr int kTileP = kBlkPts + 4;

// Writes D-layout blocks (this wavefront's point r, rows 16 mb + 4q + e).
template <int MB>
__device__ __forceinline__ void tile_store(__bf16* t, const float (&v)[MB][4]) {
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const int col = 16 * (threadIdx.x >> 6) + r;
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int e = 0; e < 4; ++e) t[(16 * mb + 4 * q + e) * kTileP + col] = (__bf16)v[mb][e];
}
// 16x16x16 operand: tile row `row`, points 16k + 4q .. +3 (one 8-byte read).
__device__ __forceinline__ bf16x4 tile_operand(const __bf16* t, int row, int k) {
  const int q = (threadIdx.x & 63) >> 4;
  return *reinterpret_cast<const bf16x4*>(t + row * kTileP + 16 * k + 4 * q);
}
// acc += A[16a .. 16a+15][pts] * B[16b .. 16b+15][pts]^T over the step's points
// (bf16 products: a gradient sum, ~2^-9 relative per term).
__device__ __forceinline__ floatx4 tile_mma(const __bf16* tA, int a, const __bf16* tB, int b,
                                            floatx4 acc) {
  const int r = threadIdx.x & 15;
#pragma unroll
  for (int k = 0; k < kWaves; ++k)
    acc = mma_b4(tile_operand(tA, 16 * a + r, k), tile_operand(tB, 16 * b + r, k), acc);
  return acc;
}

// The workgroup's partial row of the weight gradient (packed layout): weight
// blocks from the wavefronts that own them, bias sums combined over the
// wavefronts in a fixed order through LDS (red: >= (2 kWaves H + kWaves) floats
// of LDS no longer in use).  Called by every lane of the workgroup.
template <int H, int NH, int KB0, int T1, int T0>
__device__ __forceinline__ void write_partial16(float* __restrict__ out, int d, int E, int wid,
                                                float* red, const floatx4 (&acc1)[T1],
                                                const floatx4 (&acc0)[T0],
                                                const float (&db1)[H / 16][4],
                                                const float (&dwo)[H / 16][4], float dbo) {
  using C = Cfg<H, NH, KB0>;
  constexpr int MB = C::MB, MD0 = C::MD0;
  constexpr int NB1 = NH == 2 ? MB * MB : 0, NB0 = MB * MD0;
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  // the workgroup's partial row: weight blocks from their owners, bias sums
  // combined over the wavefronts in a fixed order through LDS
#pragma unroll
  for (int t = 0; t < T0; ++t) {
    const int b = wid + kWaves * t;
    if (b >= NB0) continue;
    const int ob = b / MD0, md = b % MD0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int col = h0_col(16 * md + r, E), o = 16 * ob + 4 * q + e;
      if (col >= 0) out[o * d + col] = acc0[t][e];
      else if (col == -2) out[C::gB0(d) + o] = acc0[t][e];
    }
  }
  if (NH == 2) {
#pragma unroll
    for (int t = 0; t < T1; ++t) {
      const int b = wid + kWaves * t;
      if (b >= NB1) continue;
      const int ob = b / MB, ib = b % MB;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        out[C::gW1(d) + (16 * ob + 4 * q + e) * H + 16 * ib + r] = acc1[t][e];
    }
  }
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int u = 16 * mb + 4 * q + e;
      const float s1 = sum_r(db1[mb][e]), so = sum_r(dwo[mb][e]);
      if (r == 0) {
        red[(wid * 2 + 0) * H + u] = s1;
        red[(wid * 2 + 1) * H + u] = so;
      }
    }
  float sbo = sum_r(dbo);
  sbo += __shfl_xor(sbo, 16, 64);
  sbo += __shfl_xor(sbo, 32, 64);
  if (l == 0) red[kWaves * 2 * H + wid] = sbo;
  __syncthreads();
  for (int t = threadIdx.x; t < 2 * H + 1; t += blockDim.x) {
    float v = 0.f;
    if (t < 2 * H) {
      for (int w = 0; w < kWaves; ++w) v += red[w * 2 * H + t];
      const int k = t / H, u = t % H;
      if (k == 0) { if (NH == 2) out[C::gB1(d) + u] = v; }
      else out[C::gWo(d) + u] = v;
    } else {
      for (int w = 0; w < kWaves; ++w) v += red[kWaves * 2 * H + w];
      out[C::gBo(d)] = v;
    }
  }

}

// ---------------------------------------------------------------------------
// backward of sum g_sdf * sdf: one weight-gradient partial row per workgroup
// (packed layout), g_enc [n, E] and g_x [n, 3] (nullable).
//
// A workgroup steps over 64 points at a time (16 per wavefront).  Each
// wavefront recomputes its group's forward pass and backpropagates it in
// registers; the step's gradient and activation blocks go to the block tiles
// and the weight-gradient blocks are split over the wavefronts (block b is
// accumulated by wavefront b % 4 over all 64 points), so each wavefront holds
// a quarter of the accumulators.
// ---------------------------------------------------------------------------
template <int H, int NH, int KB0>
__global__ void __launch_bounds__(64 * kWaves, 2)
mlp16_bwd_kernel(const float* __restrict__ x, const float* __restrict__ enc, int E, int vec4,
                 int64_t n_cap, const int32_t* __restrict__ n_dev, const float* __restrict__ mlp,
                 float beta, const float* __restrict__ g_sdf, float* __restrict__ g_enc,
                 float* __restrict__ g_x, float* __restrict__ partials) {
  using C = Cfg<H, NH, KB0>;
  constexpr int MB = C::MB, KBH = C::KBH, MD0 = C::MD0;
  constexpr int NB1 = NH == 2 ? MB * MB : 0, NB0 = MB * MD0;
  constexpr int T1 = NH == 2 ? (NB1 + kWaves - 1) / kWaves : 1, T0 = (NB0 + kWaves - 1) / kWaves;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  build_frags<H, NH, KB0, true>(smem, mlp, E);
  __syncthreads();
  const half8* fW0 = reinterpret_cast<const half8*>(smem);
  const half8* fW1 = fW0 + C::nW0 * 64;
  const bf16x8* fW1T = reinterpret_cast<const bf16x8*>(fW1 + C::nW1 * 64);
  const bf16x8* fW0T = fW1T + C::nW1T * 64;
  const float* vec = reinterpret_cast<const float*>(fW0T + C::nW0T * 64);
  __bf16* tileA = reinterpret_cast<__bf16*>(smem + C::bwd_lds - C::tile_bytes);
  __bf16* tileB = tileA + 64 * kTileP;
  // wavefront index as a scalar: the block-ownership tests below are uniform
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const float inv_beta = 1.f / beta;
  const int64_t n = dev_count(n_cap, n_dev);
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const int d = 3 + E;
  const int md_used = (E + 4 + 15) / 16;  // h0 dim blocks holding data

  floatx4 acc1[T1], acc0[T0];
#pragma unroll
  for (int t = 0; t < T1; ++t) acc1[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < T0; ++t) acc0[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  float db1[MB][4], dwo[MB][4], dbo = 0.f;
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int e = 0; e < 4; ++e) db1[a][e] = dwo[a][e] = 0.f;

  const int64_t n_steps = ceil_div(n, (int64_t)kBlkPts);
  for (int64_t step = blockIdx.x; step < n_steps; step += gridDim.x) {
    const int64_t p = (step * kWaves + wid) * 16 + r;
    const bool ok = p < n;
    // the lane's quarter index, opaque to the compiler in every step: the
    // per-slot boundary tests of the h0 loads and gradient stores derive
    // from it and are then evaluated where used, instead of being hoisted
    // out of the loop as dozens of long-lived execution masks (which spill)
    int qs = q, Es = E, v4s = vec4;
    asm volatile("" : "+v"(qs), "+s"(Es), "+s"(v4s));
    float h0[KB0][8];
#pragma unroll
    for (int kb = 0; kb < KB0; ++kb) load_h0(x, enc, Es, v4s, p, ok, kb, qs, h0[kb]);
    const float gs = ok ? g_sdf[p] : 0.f;
    Fwd<H, NH, KB0> F;
    if (q == 0) dbo += gs;
    float gz[MB][4];
    forward_group<H, NH, KB0, true, false>(
        fW0, fW1, vec, h0, beta, inv_beta, F, [&](int mb, int e, float hL, float dL) {
          dwo[mb][e] += gs * hL;
          gz[mb][e] = gs * vec[2 * H + 16 * mb + 4 * q + e] * dL;
        });
    if (NH == 2) {
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int e = 0; e < 4; ++e) db1[mb][e] += gz[mb][e];
      tile_store<MB>(tileA, gz);
      tile_store<MB>(tileB, F.h1);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < T1; ++t) {
        const int b = wid + kWaves * t;
        if (b < NB1) acc1[t] = tile_mma(tileA, b / MB, tileB, b % MB, acc1[t]);
      }
      // g_h1 = W1^T gz1 (own group, registers), then through softplus'(z0)
      bf16x8 gh[KBH], gl[KBH];
#pragma unroll
      for (int kb = 0; kb < KBH; ++kb) hid_operand_b<MB>(gz, kb, gh[kb], gl[kb]);
#pragma unroll
      for (int mi = 0; mi < MB; ++mi) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KBH; ++kb) {
          const bf16x8 w = fW1T[(mi * KBH + kb) * 64 + l];
          acc = mma_b(w, gh[kb], acc);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) gz[mi][e] = acc[e] * F.d1[mi][e];
      }
      __syncthreads();  // every wavefront is done reading the tiles
    }
    tile_store<MB>(tileA, gz);
#pragma unroll
    for (int kb = 0; kb < KB0; ++kb)
#pragma unroll
      for (int j = 0; j < 8; ++j)
        tileB[(32 * kb + 8 * q + j) * kTileP + 16 * wid + r] = (__bf16)h0[kb][j];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < T0; ++t) {
      const int b = wid + kWaves * t;
      if (b < NB0 && (b % MD0) < md_used) acc0[t] = tile_mma(tileA, b / MD0, tileB, b % MD0, acc0[t]);
    }
    if (g_enc || g_x) {
      bf16x8 gh[KBH], gl[KBH];
#pragma unroll
      for (int kb = 0; kb < KBH; ++kb) hid_operand_b<MB>(gz, kb, gh[kb], gl[kb]);
#pragma unroll
      for (int md = 0; md < MD0; ++md) {
        if (md >= md_used) break;
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KBH; ++kb) {
          const bf16x8 w = fW0T[(md * KBH + kb) * 64 + l];
          acc = mma_b(w, gh[kb], acc);
        }
        const int c0 = 16 * md + 4 * qs;  // internal dims c0..c0+3 of point p
        if (v4s) {  // 4-dim groups: encoding dims, or [x, 1]
          if (ok && c0 < Es) {
            if (g_enc)
              *reinterpret_cast<float4*>(g_enc + p * Es + c0) =
                  float4{acc[0], acc[1], acc[2], acc[3]};
          } else if (ok && c0 == Es && g_x) {
            g_x[p * 3] = acc[0];
            g_x[p * 3 + 1] = acc[1];
            g_x[p * 3 + 2] = acc[2];
          }
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int c = c0 + e;
            float* dst = c < Es ? g_enc + p * Es + c : g_x + p * 3 + (c - Es);
            const bool want = c < Es ? g_enc != nullptr : (c < Es + 3 && g_x != nullptr);
            if (ok && want) *dst = acc[e];
          }
        }
      }
    }
    __syncthreads();  // tiles are rewritten by the next step
  }

  write_partial16<H, NH, KB0>(partials + (int64_t)blockIdx.x * C::n_params(d), d, E, wid,
                              reinterpret_cast<float*>(tileA), acc1, acc0, db1, dwo, dbo);
}

// ---------------------------------------------------------------------------
// forward-over-reverse backward of J = sum_p a_p sdf(x_p) + v_p . grad_x sdf(x_p)
// (the analytic gradient trained through its value, fields.py:107-119 with
// create_graph), the bf16 counterpart of mlp_bwd_ad_kernel: tangents along v
// of the inputs hd0 = [encd | v | 0] (encd: the encoding's directional
// derivative, no bias slot) propagated with the forward, then the reverse
// pass of the value (gb) and of the tangent (gt):
//   last layer: gb = a wo s' + wo zd s'',  gt = wo s'   (s'' = beta s'(1 - s'))
//   hidden:     gb_z = (W^T gb) s' + (W^T gt) zd s'',  gt_z = (W^T gt) s'
//   weights:    dW += gb (x) h + gt (x) hd, db += gb, dwo += a h + s' zd, dbo += a
// Outputs dJ/denc and dJ/dencd [n, E] (the encoding's columns) and the
// weight-gradient partial rows.  Same structure as mlp16_bwd_kernel: the
// forward recomputed from hi parts on fp16, cotangents on bf16; the weight
// gradients through four block tiles (gb with h, gt with hd, summed in one
// accumulator per block).
// ---------------------------------------------------------------------------
template <int H, int NH, int KB0>
__global__ void __launch_bounds__(64 * kWaves, 2)  // 2 wavefronts per SIMD (measured 7 % faster with
                                                   // its ~32 spilled registers than 1 without)
mlp16_bwd_ad_kernel(const float* __restrict__ x, const float* __restrict__ enc,
                    const float* __restrict__ v, const float* __restrict__ encd, int E, int vec4,
                    int64_t n_cap, const int32_t* __restrict__ n_dev,
                    const float* __restrict__ mlp, float beta, const float* __restrict__ a_sdf,
                    float* __restrict__ g_enc, float* __restrict__ g_encd,
                    float* __restrict__ partials) {
  using C = Cfg<H, NH, KB0>;
  constexpr int MB = C::MB, KBH = C::KBH, MD0 = C::MD0;
  constexpr int NB1 = NH == 2 ? MB * MB : 0, NB0 = MB * MD0;
  constexpr int T1 = NH == 2 ? (NB1 + kWaves - 1) / kWaves : 1, T0 = (NB0 + kWaves - 1) / kWaves;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  build_frags<H, NH, KB0, true>(smem, mlp, E);
  __syncthreads();
  const half8* fW0 = reinterpret_cast<const half8*>(smem);
  const half8* fW1 = fW0 + C::nW0 * 64;
  const bf16x8* fW1T = reinterpret_cast<const bf16x8*>(fW1 + C::nW1 * 64);
  const bf16x8* fW0T = fW1T + C::nW1T * 64;
  const float* vec = reinterpret_cast<const float*>(fW0T + C::nW0T * 64);
  __bf16* tA = reinterpret_cast<__bf16*>(smem + C::bwdad_lds - 2 * C::tile_bytes);
  __bf16* tB = tA + 64 * kTileP;
  __bf16* tC = tB + 64 * kTileP;
  __bf16* tD = tC + 64 * kTileP;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const float inv_beta = 1.f / beta;
  const int64_t n = dev_count(n_cap, n_dev);
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const int d = 3 + E;
  const int md_used = (E + 4 + 15) / 16;

  floatx4 acc1[T1], acc0[T0];
#pragma unroll
  for (int t = 0; t < T1; ++t) acc1[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < T0; ++t) acc0[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  float db1[MB][4], dwo[MB][4], dbo = 0.f;
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int e = 0; e < 4; ++e) db1[a][e] = dwo[a][e] = 0.f;

  const int64_t n_steps = ceil_div(n, (int64_t)kBlkPts);
  for (int64_t step = blockIdx.x; step < n_steps; step += gridDim.x) {
    const int64_t p = (step * kWaves + wid) * 16 + r;
    const bool ok = p < n;
    int qs = q, Es = E, v4s = vec4;
    asm volatile("" : "+v"(qs), "+s"(Es), "+s"(v4s));
    const float as = ok ? a_sdf[p] : 0.f;
    if (q == 0) dbo += as;
    // layer 0 and its tangent (fp16, hi parts)
    half8 bh[KB0], bdh[KB0];
    {
      float h0[KB0][8], hd0[KB0][8];
#pragma unroll
      for (int kb = 0; kb < KB0; ++kb) {
        load_h0(x, enc, Es, v4s, p, ok, kb, qs, h0[kb]);
        load_h0(v, encd, Es, v4s, p, ok, kb, qs, hd0[kb], 0.f);
      }
#pragma unroll
      for (int kb = 0; kb < KB0; ++kb)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          bh[kb][j] = (_Float16)h0[kb][j];
          bdh[kb][j] = (_Float16)hd0[kb][j];
        }
    }
    float d1[MB][4], zd0[MB][4];  // s'(z0), tangent of z0 (NH == 2)
    float gb[MB][4], gt[MB][4];   // cotangents of the current layer's pre-activation
    if constexpr (NH == 2) {
      float h1[MB][4], hd1[MB][4];
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        floatx4 z = {0.f, 0.f, 0.f, 0.f}, zd = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KB0; ++kb) {
          const half8 w = fW0[(mb * KB0 + kb) * 64 + l];
          z = mma_h(w, bh[kb], z);
          zd = mma_h(w, bdh[kb], zd);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const Sp sp = softplus(z[e], beta, inv_beta);
          h1[mb][e] = sp.h;
          d1[mb][e] = sp.d;
          zd0[mb][e] = zd[e];
          hd1[mb][e] = sp.d * zd[e];
        }
      }
      // layer 1 and its tangent; the activations go to their tiles now
      tile_store<MB>(tB, h1);
      tile_store<MB>(tD, hd1);
      half8 ch[KBH], cl[KBH], cdh[KBH], cdl[KBH];
#pragma unroll
      for (int kb = 0; kb < KBH; ++kb) {
        hid_operand_h<MB>(h1, kb, ch[kb], cl[kb]);
        hid_operand_h<MB>(hd1, kb, cdh[kb], cdl[kb]);
      }
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        floatx4 z = {0.f, 0.f, 0.f, 0.f}, zd = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KBH; ++kb) {
          const half8 w = fW1[(mb * KBH + kb) * 64 + l];
          z = mma_h(w, ch[kb], z);
          zd = mma_h(w, cdh[kb], zd);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int u = 16 * mb + 4 * q + e;
          const Sp sp = softplus(z[e] + vec[H + u], beta, inv_beta);
          const float s2 = beta * sp.d * (1.f - sp.d), wo = vec[2 * H + u];
          dwo[mb][e] += as * sp.h + sp.d * zd[e];
          gb[mb][e] = as * wo * sp.d + wo * zd[e] * s2;
          gt[mb][e] = wo * sp.d;
          db1[mb][e] += gb[mb][e];
        }
      }
      tile_store<MB>(tA, gb);
      tile_store<MB>(tC, gt);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < T1; ++t) {
        const int b = wid + kWaves * t;
        if (b < NB1)
          acc1[t] = tile_mma(tC, b / MB, tD, b % MB, tile_mma(tA, b / MB, tB, b % MB, acc1[t]));
      }
      // to layer 0 (bf16)
      bf16x8 gbh[KBH], gbl[KBH], gth[KBH], gtl[KBH];
#pragma unroll
      for (int kb = 0; kb < KBH; ++kb) {
        hid_operand_b<MB>(gb, kb, gbh[kb], gbl[kb]);
        hid_operand_b<MB>(gt, kb, gth[kb], gtl[kb]);
      }
#pragma unroll
      for (int mi = 0; mi < MB; ++mi) {
        floatx4 ab = {0.f, 0.f, 0.f, 0.f}, at = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KBH; ++kb) {
          const bf16x8 w = fW1T[(mi * KBH + kb) * 64 + l];
          ab = mma_b(w, gbh[kb], ab);
          at = mma_b(w, gth[kb], at);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float s1 = d1[mi][e], s2 = beta * s1 * (1.f - s1);
          gb[mi][e] = ab[e] * s1 + at[e] * zd0[mi][e] * s2;
          gt[mi][e] = at[e] * s1;
        }
      }
      __syncthreads();  // every wavefront is done reading the tiles
    } else {
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        floatx4 z = {0.f, 0.f, 0.f, 0.f}, zd = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KB0; ++kb) {
          const half8 w = fW0[(mb * KB0 + kb) * 64 + l];
          z = mma_h(w, bh[kb], z);
          zd = mma_h(w, bdh[kb], zd);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int u = 16 * mb + 4 * q + e;
          const Sp sp = softplus(z[e], beta, inv_beta);
          const float s2 = beta * sp.d * (1.f - sp.d), wo = vec[2 * H + u];
          dwo[mb][e] += as * sp.h + sp.d * zd[e];
          gb[mb][e] = as * wo * sp.d + wo * zd[e] * s2;
          gt[mb][e] = wo * sp.d;
        }
      }
    }
    // layer 0's weights: gb (x) h0 + gt (x) hd0 (the inputs read again)
    tile_store<MB>(tA, gb);
    tile_store<MB>(tC, gt);
    {
      float h0[KB0][8], hd0[KB0][8];
#pragma unroll
      for (int kb = 0; kb < KB0; ++kb) {
        load_h0(x, enc, Es, v4s, p, ok, kb, qs, h0[kb]);
        load_h0(v, encd, Es, v4s, p, ok, kb, qs, hd0[kb], 0.f);
      }
#pragma unroll
      for (int kb = 0; kb < KB0; ++kb)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          tB[(32 * kb + 8 * q + j) * kTileP + 16 * wid + r] = (__bf16)h0[kb][j];
          tD[(32 * kb + 8 * q + j) * kTileP + 16 * wid + r] = (__bf16)hd0[kb][j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < T0; ++t) {
      const int b = wid + kWaves * t;
      if (b < NB0 && (b % MD0) < md_used)
        acc0[t] = tile_mma(tC, b / MD0, tD, b % MD0, tile_mma(tA, b / MD0, tB, b % MD0, acc0[t]));
    }
    // dJ/denc = W0^T gb, dJ/dencd = W0^T gt (the encoding's columns)
    {
      bf16x8 gbh[KBH], gbl[KBH], gth[KBH], gtl[KBH];
#pragma unroll
      for (int kb = 0; kb < KBH; ++kb) {
        hid_operand_b<MB>(gb, kb, gbh[kb], gbl[kb]);
        hid_operand_b<MB>(gt, kb, gth[kb], gtl[kb]);
      }
#pragma unroll
      for (int md = 0; md < MD0; ++md) {
        if (md >= md_used) break;
        floatx4 ab = {0.f, 0.f, 0.f, 0.f}, at = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KBH; ++kb) {
          const bf16x8 w = fW0T[(md * KBH + kb) * 64 + l];
          ab = mma_b(w, gbh[kb], ab);
          at = mma_b(w, gth[kb], at);
        }
        const int c0 = 16 * md + 4 * qs;  // internal dims c0..c0+3 of point p
        if (v4s) {
          if (ok && c0 < Es) {
            *reinterpret_cast<float4*>(g_enc + p * Es + c0) = float4{ab[0], ab[1], ab[2], ab[3]};
            *reinterpret_cast<float4*>(g_encd + p * Es + c0) = float4{at[0], at[1], at[2], at[3]};
          }
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            if (ok && c0 + e < Es) {
              g_enc[p * Es + c0 + e] = ab[e];
              g_encd[p * Es + c0 + e] = at[e];
            }
          }
        }
      }
    }
    __syncthreads();  // tiles are rewritten by the next step
  }
  write_partial16<H, NH, KB0>(partials + (int64_t)blockIdx.x * C::n_params(d), d, E, wid,
                              reinterpret_cast<float*>(tA), acc1, acc0, db1, dwo, dbo);
}

// Raises a kernel's dynamic LDS limit above the default 64 KB (gfx950 has
// 160 KB), once per kernel; a refusal surfaces as a launch error.
template <typename Kern>
static void allow_lds(Kern k, size_t bytes) {
  static const void* done[16];
  static int n_done = 0;
  const void* f = reinterpret_cast<const void*>(k);
  if (bytes <= 65536) return;
  for (int i = 0; i < n_done; ++i)
    if (done[i] == f) return;
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (n_done < 16) done[n_done++] = f;
}

#define SN_MLP16_DISPATCH(H_, NH_, KB0_, ...)                                           \
  do {                                                                                  \
    if (H_ == 32 && NH_ == 1 && KB0_ == 1) { constexpr int kH = 32, kNH = 1, kKB = 1; __VA_ARGS__; } \
    else if (H_ == 32 && NH_ == 1) { constexpr int kH = 32, kNH = 1, kKB = 2; __VA_ARGS__; }          \
    else if (H_ == 32 && KB0_ == 1) { constexpr int kH = 32, kNH = 2, kKB = 1; __VA_ARGS__; }         \
    else if (H_ == 32) { constexpr int kH = 32, kNH = 2, kKB = 2; __VA_ARGS__; }                      \
    else if (NH_ == 1 && KB0_ == 1) { constexpr int kH = 64, kNH = 1, kKB = 1; __VA_ARGS__; }         \
    else if (NH_ == 1) { constexpr int kH = 64, kNH = 1, kKB = 2; __VA_ARGS__; }                      \
    else if (KB0_ == 1) { constexpr int kH = 64, kNH = 2, kKB = 1; __VA_ARGS__; }                     \
    else { constexpr int kH = 64, kNH = 2, kKB = 2; __VA_ARGS__; }                                    \
  } while (0)

static int check_cfg(const char* who, int E, int H, int NH, int* kb0) {
  SN_REQUIRE(H == 32 || H == 64, "%s: d_hidden must be 32 or 64 (got %d)", who, H);
  SN_REQUIRE(NH == 1 || NH == 2, "%s: n_layers must be 1 or 2 (got %d)", who, NH);
  SN_REQUIRE(E >= 0 && E + 4 <= 64, "%s: encoding width %d must be <= 60", who, E);
  *kb0 = (E + 4 + 31) / 32;
  return SN_OK;
}

// float4 access to rows of E floats
static int vec4_ok(const float* p, int E) {
  return (E % 4 == 0) && ((uintptr_t)p % 16 == 0) ? 1 : 0;
}

constexpr int kFwdBlocksMax = 256 * 4;
constexpr int kBwdBlocksMax = 256 * 2;

static unsigned blocks_for(int64_t n_cap, int max_blocks) {
  int64_t b = ceil_div(ceil_div(n_cap > 0 ? n_cap : 1, 16), kWaves);
  return (unsigned)(b < max_blocks ? b : max_blocks);
}

}  // namespace m16
}  // namespace sn

using namespace sn;
using namespace sn::m16;

extern "C" {

int sn_mlp16_fwd(const float* x, const float* enc, int32_t enc_dims, int64_t n_cap,
                 const int32_t* n_dev, const float* mlp, int32_t d_hidden, int32_t n_layers,
                 float beta, float* sdf, void* stream) {
  clear_error();
  int kb0 = 0;
  int rc = check_cfg("sn_mlp16_fwd", enc_dims, d_hidden, n_layers, &kb0);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0 && beta > 0.f, "sn_mlp16_fwd: bad n / beta");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && mlp && sdf && (enc || enc_dims == 0), "sn_mlp16_fwd: null pointer");
  const unsigned grid = blocks_for(n_cap, kFwdBlocksMax);
  hipStream_t s = as_stream(stream);
  SN_MLP16_DISPATCH(d_hidden, n_layers, kb0, {
    constexpr size_t lds = Cfg<kH, kNH, kKB>::fwd_lds;
    allow_lds(mlp16_fwd_kernel<kH, kNH, kKB>, lds);
    mlp16_fwd_kernel<kH, kNH, kKB><<<grid, 64 * kWaves, lds, s>>>(
        x, enc, enc_dims, vec4_ok(enc, enc_dims), n_cap, n_dev, mlp, beta, sdf);
  });
  return check_launch("sn_mlp16_fwd");
}

int sn_mlp16_fwd_grad(const float* x, const float* enc, int32_t enc_dims, int64_t n_cap,
                      const int32_t* n_dev, const float* mlp, int32_t d_hidden, int32_t n_layers,
                      float beta, float* sdf, float* dh0, void* stream) {
  clear_error();
  int kb0 = 0;
  int rc = check_cfg("sn_mlp16_fwd_grad", enc_dims, d_hidden, n_layers, &kb0);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0 && beta > 0.f, "sn_mlp16_fwd_grad: bad n / beta");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && mlp && sdf && dh0 && (enc || enc_dims == 0), "sn_mlp16_fwd_grad: null pointer");
  const unsigned grid = blocks_for(n_cap, kFwdBlocksMax);
  hipStream_t s = as_stream(stream);
  SN_MLP16_DISPATCH(d_hidden, n_layers, kb0, {
    constexpr size_t lds = Cfg<kH, kNH, kKB>::fwdg_lds;
    allow_lds(mlp16_fwd_grad_kernel<kH, kNH, kKB>, lds);
    mlp16_fwd_grad_kernel<kH, kNH, kKB><<<grid, 64 * kWaves, lds, s>>>(
        x, enc, enc_dims, vec4_ok(enc, enc_dims), n_cap, n_dev, mlp, beta, sdf, dh0);
  });
  return check_launch("sn_mlp16_fwd_grad");
}

size_t sn_mlp16_bwd_workspace_bytes(int64_t n_cap, int32_t enc_dims, int32_t d_hidden,
                                    int32_t n_layers) {
  const int np = sn_mlp_n_params(3 + enc_dims, d_hidden, n_layers);
  if (np < 0) return 0;
  return (size_t)blocks_for(n_cap, kBwdBlocksMax) * np * sizeof(float);
}

int sn_mlp16_bwd(const float* x, const float* enc, int32_t enc_dims, int64_t n_cap,
                 const int32_t* n_dev, const float* mlp, int32_t d_hidden, int32_t n_layers,
                 float beta, const float* g_sdf, float* g_enc, float* g_x, float* g_mlp,
                 int32_t accumulate, void* ws, size_t ws_bytes, void* stream) {
  clear_error();
  int kb0 = 0;
  int rc = check_cfg("sn_mlp16_bwd", enc_dims, d_hidden, n_layers, &kb0);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0 && beta > 0.f, "sn_mlp16_bwd: bad n / beta");
  SN_REQUIRE(mlp && g_mlp, "sn_mlp16_bwd: null pointer");
  const int np = sn_mlp_n_params(3 + enc_dims, d_hidden, n_layers);
  hipStream_t s = as_stream(stream);
  if (n_cap == 0) {
    if (accumulate) return SN_OK;
    return fill_f32(g_mlp, np, 0.f, s);
  }
  SN_REQUIRE(x && g_sdf && (enc || enc_dims == 0), "sn_mlp16_bwd: null input");
  const size_t need = sn_mlp16_bwd_workspace_bytes(n_cap, enc_dims, d_hidden, n_layers);
  if (ws_bytes < need) {
    set_error("sn_mlp16_bwd: workspace %zu < %zu bytes", ws_bytes, need);
    return SN_ESPACE;
  }
  const unsigned grid = blocks_for(n_cap, kBwdBlocksMax);
  float* part = reinterpret_cast<float*>(ws);
  SN_MLP16_DISPATCH(d_hidden, n_layers, kb0, {
    constexpr size_t lds = Cfg<kH, kNH, kKB>::bwd_lds;
    allow_lds(mlp16_bwd_kernel<kH, kNH, kKB>, lds);
    mlp16_bwd_kernel<kH, kNH, kKB><<<grid, 64 * kWaves, lds, s>>>(
        x, enc, enc_dims, vec4_ok(enc, enc_dims) && vec4_ok(g_enc, enc_dims), n_cap, n_dev, mlp,
        beta, g_sdf, g_enc, g_x, part);
  });
  if ((rc = check_launch("sn_mlp16_bwd"))) return rc;
  return mlp_reduce_rows(part, (int64_t)grid, np, g_mlp, accumulate, s);
}

int sn_mlp16_bwd_ad_enc(const float* x, const float* enc, const float* v, const float* enc_dot,
                        int32_t enc_dims, int64_t n_cap, const int32_t* n_dev, const float* mlp,
                        int32_t d_hidden, int32_t n_layers, float beta, const float* g_sdf,
                        float* g_enc, float* g_enc_dot, float* g_mlp, int32_t accumulate,
                        void* ws, size_t ws_bytes, void* stream) {
  clear_error();
  int kb0 = 0;
  int rc = check_cfg("sn_mlp16_bwd_ad_enc", enc_dims, d_hidden, n_layers, &kb0);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0 && beta > 0.f, "sn_mlp16_bwd_ad_enc: bad n / beta");
  SN_REQUIRE(mlp && g_mlp, "sn_mlp16_bwd_ad_enc: null pointer");
  const int np = sn_mlp_n_params(3 + enc_dims, d_hidden, n_layers);
  hipStream_t s = as_stream(stream);
  if (n_cap == 0) {
    if (accumulate) return SN_OK;
    return fill_f32(g_mlp, np, 0.f, s);
  }
  SN_REQUIRE(x && v && g_sdf && g_enc && g_enc_dot && ((enc && enc_dot) || enc_dims == 0),
             "sn_mlp16_bwd_ad_enc: null input");
  const size_t need = sn_mlp16_bwd_workspace_bytes(n_cap, enc_dims, d_hidden, n_layers);
  if (ws_bytes < need) {
    set_error("sn_mlp16_bwd_ad_enc: workspace %zu < %zu bytes", ws_bytes, need);
    return SN_ESPACE;
  }
  const unsigned grid = blocks_for(n_cap, kBwdBlocksMax);
  float* part = reinterpret_cast<float*>(ws);
  const int v4 = vec4_ok(enc, enc_dims) && vec4_ok(enc_dot, enc_dims) &&
                 vec4_ok(g_enc, enc_dims) && vec4_ok(g_enc_dot, enc_dims);
  SN_MLP16_DISPATCH(d_hidden, n_layers, kb0, {
    constexpr size_t lds = Cfg<kH, kNH, kKB>::bwdad_lds;
    allow_lds(mlp16_bwd_ad_kernel<kH, kNH, kKB>, lds);
    mlp16_bwd_ad_kernel<kH, kNH, kKB><<<grid, 64 * kWaves, lds, s>>>(
        x, enc, v, enc_dot, enc_dims, v4, n_cap, n_dev, mlp, beta, g_sdf, g_enc, g_enc_dot,
        part);
  });
  if ((rc = check_launch("sn_mlp16_bwd_ad_enc"))) return rc;
  return mlp_reduce_rows(part, (int64_t)grid, np, g_mlp, accumulate, s);
}

}  // extern "C"
