// The SDF MLP of models/fields.py:85-99 (input_concat: h0 = [x, enc]; NH
// hidden layers of width H with Softplus(beta, threshold 20); linear output of
// width 1) as wavefront tiles on the fp32 matrix cores (v_mfma_f32_16x16x4f32).
//
// A workgroup is 4 wavefronts; each wavefront owns a tile of T points (rows).
// The tile's activations live in LDS (row-major, pitch padded by 4 floats);
// the pre-folded (weight-normed) weights of every layer live in LDS for the
// whole workgroup.  Each GEMM is a set of 16x16 output blocks, each the sum
// over K in steps of 4 of one MFMA with strided LDS operand views
//   A(i,k) = a[i*ai + k*ak],  B(k,j) = b[k*bk + j*bj],
// so one helper serves Act*W^T (forward), G*W (backprop to the inputs) and
// G^T*Act (weight gradients, K = the tile's points).
//
// MFMA 16x16x4 fp32 lane mapping (lane l, r = l % 16, q = l / 16):
//   A operand A[r][q], B operand B[q][r], result D[4q + e][r], e = 0..3.
//
// Weight gradients: each wavefront sums its tiles in registers in a fixed
// order and writes one partial row; sn_mlp_reduce adds the rows in a fixed
// order, so results do not depend on scheduling (no float atomics).
//
// Packed parameter layout (fp32, row-major; the gradient uses the same):
//   W0[H][d_in] b0[H] { W1[H][H] b1[H] } wo[H] bo
#include "sn_common.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace sn {

constexpr int kWaves = 4;
constexpr float kSpThr = 20.f;

__device__ __forceinline__ floatx4 mma_blk(const float* a, int ai, int ak, const float* b,
                                           int bk, int bj, int K, floatx4 acc) {
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll 4
  for (int k = 0; k < K; k += 4)
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r * ai + (k + q) * ak], b[(k + q) * bk + r * bj],
                                               acc, 0, 0, 0);
  return acc;
}

// Softplus(beta) with torch's threshold and its derivative e/(e+1), e = exp(beta z).
__device__ __forceinline__ float sp(float z, float beta) {
  const float bz = z * beta;
  return bz > kSpThr ? z : log1pf(expf(bz)) / beta;
}
__device__ __forceinline__ float sp_d(float z, float beta) {
  const float bz = z * beta;
  if (bz > kSpThr) return 1.f;
  const float e = expf(bz);
  return e / (e + 1.f);
}
// d^2 softplus / dz^2 = beta s (1 - s) (0 above the threshold, where torch's
// derivative is the constant 1).
__device__ __forceinline__ float sp_dd(float z, float beta) {
  const float bz = z * beta;
  if (bz > kSpThr) return 0.f;
  const float s = sp_d(z, beta);
  return beta * s * (1.f - s);
}

__device__ __forceinline__ float sum16(float v) {  // over lanes with equal l / 16
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  v += __shfl_xor(v, 8, 64);
  return v;
}
__device__ __forceinline__ float sum4q(float v) {  // over lanes with equal l % 16
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

template <int H, int NH, int DP>
struct Mlp {
  static constexpr int LDI = DP + 4;
  static constexpr int LDH = H + 4;
  static constexpr int HB = H / 16;
  static constexpr int IB = DP / 16;
  // LDS weight offsets
  static constexpr int oW0 = 0;
  static constexpr int oB0 = oW0 + H * LDI;
  static constexpr int oW1 = oB0 + H;
  static constexpr int oB1 = oW1 + (NH == 2 ? H * LDH : 0);
  static constexpr int oWo = oB1 + (NH == 2 ? H : 0);
  static constexpr int oBo = oWo + H;
  static constexpr int kWFloats = oBo + 4;
  // per-wavefront buffers (floats per row)
  static constexpr int kRow = LDI + 2 * LDH;
  // global packed offsets
  __host__ __device__ static int n_params(int d) {
    return H * d + H + (NH == 2 ? H * H + H : 0) + H + 1;
  }
  __host__ __device__ static int gB0(int d) { return H * d; }
  __host__ __device__ static int gW1(int d) { return H * d + H; }
  __host__ __device__ static int gB1(int d) { return H * d + H + H * H; }
  __host__ __device__ static int gWo(int d) { return H * d + H + (NH == 2 ? H * H + H : 0); }
  __host__ __device__ static int gBo(int d) { return gWo(d) + H; }
};

template <int H, int NH, int DP>
__device__ void load_weights(float* sw, const float* __restrict__ mlp, int d) {
  using M = Mlp<H, NH, DP>;
  for (int t = threadIdx.x; t < H * M::LDI; t += blockDim.x) {
    const int o = t / M::LDI, i = t % M::LDI;
    sw[M::oW0 + t] = (i < d) ? mlp[o * d + i] : 0.f;
  }
  for (int t = threadIdx.x; t < H; t += blockDim.x) {
    sw[M::oB0 + t] = mlp[M::gB0(d) + t];
    sw[M::oWo + t] = mlp[M::gWo(d) + t];
    if (NH == 2) sw[M::oB1 + t] = mlp[M::gB1(d) + t];
  }
  if (NH == 2)
    for (int t = threadIdx.x; t < H * M::LDH; t += blockDim.x) {
      const int o = t / M::LDH, i = t % M::LDH;
      sw[M::oW1 + t] = (i < H) ? mlp[M::gW1(d) + o * H + i] : 0.f;
    }
  if (threadIdx.x == 0) sw[M::oBo] = mlp[M::gBo(d)];
}

// Rows [row0, row0+T) of [a3 | b] into an LDS tile (zero past n and past 3+E).
template <int DP, int T>
__device__ void load_rows(float* s, const float* __restrict__ a3, const float* __restrict__ b,
                          int E, int64_t row0, int64_t n) {
  constexpr int LDI = DP + 4;
  const int lane = threadIdx.x & 63;
  for (int t = lane; t < T * DP; t += 64) {
    const int i = t / DP, c = t % DP;
    const int64_t p = row0 + i;
    float v = 0.f;
    if (p < n) {
      if (c < 3) v = a3[p * 3 + c];
      else if (c < 3 + E) v = b[p * E + (c - 3)];
    }
    s[i * LDI + c] = v;
  }
}

// z = A*W^T (+ bias) for all (rb, cb) blocks of a tile.
template <int T, int CB, bool kBias>
__device__ __forceinline__ void gemm_xwT(const float* sA, int lda, const float* sW, int ldw,
                                         const float* bias, int K, floatx4 (&z)[T / 16][CB]) {
  const int r = (threadIdx.x & 63) & 15;
#pragma unroll
  for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mma_blk(sA + rb * 16 * lda, lda, 1, sW + cb * 16 * ldw, 1, ldw, K, acc);
      if (kBias) {
        const float b = bias[cb * 16 + r];
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e] += b;
      }
      z[rb][cb] = acc;
    }
}

// y = G*W, G [T x K] in LDS (pitch ldg), W row-major [K][ldw]; CB output blocks.
template <int T, int CB>
__device__ __forceinline__ void gemm_gw(const float* sG, int ldg, const float* sW, int ldw,
                                        int K, floatx4 (&y)[T / 16][CB]) {
#pragma unroll
  for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      y[rb][cb] = mma_blk(sG + rb * 16 * ldg, ldg, 1, sW + cb * 16, ldw, 1, K, acc);
    }
}

// acc += G^T * A  (G [T x 16*OB] pitch ldg, A [T x 16*IB] pitch lda): weight gradient.
template <int T, int OB, int IB>
__device__ __forceinline__ void gemm_gtA(const float* sG, int ldg, const float* sA, int lda,
                                         floatx4 (&acc)[OB][IB]) {
#pragma unroll
  for (int ob = 0; ob < OB; ++ob)
#pragma unroll
    for (int ib = 0; ib < IB; ++ib)
      acc[ob][ib] = mma_blk(sG + ob * 16, 1, ldg, sA + ib * 16, lda, 1, T, acc[ob][ib]);
}

template <int T, int CB>
__device__ __forceinline__ void store_frags(float* s, int ld, const floatx4 (&v)[T / 16][CB]) {
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
  for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int e = 0; e < 4; ++e) s[(rb * 16 + 4 * q + e) * ld + cb * 16 + r] = v[rb][cb][e];
}

// Column sums of fragments, added into per-lane partials (column cb*16 + r).
template <int T, int CB>
__device__ __forceinline__ void colsum_add(const floatx4 (&v)[T / 16][CB], float (&acc)[CB]) {
#pragma unroll
  for (int cb = 0; cb < CB; ++cb) {
    float s = 0.f;
#pragma unroll
    for (int rb = 0; rb < T / 16; ++rb) s += v[rb][cb][0] + v[rb][cb][1] + v[rb][cb][2] + v[rb][cb][3];
    acc[cb] += s;
  }
}

// Writes one wavefront's gradient partial row in the packed layout.
template <int H, int NH, int DP>
__device__ void write_partial(float* __restrict__ out, int d, const floatx4 (&dW0)[H / 16][DP / 16],
                              const floatx4 (&dW1)[H / 16][H / 16], const float (&db0)[H / 16],
                              const float (&db1)[H / 16], const float (&dwo)[H / 16], float dbo) {
  using M = Mlp<H, NH, DP>;
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
  for (int ob = 0; ob < M::HB; ++ob)
#pragma unroll
    for (int ib = 0; ib < M::IB; ++ib)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int o = ob * 16 + 4 * q + e, i = ib * 16 + r;
        if (i < d) out[o * d + i] = dW0[ob][ib][e];
      }
  if (NH == 2) {
#pragma unroll
    for (int ob = 0; ob < M::HB; ++ob)
#pragma unroll
      for (int ib = 0; ib < M::HB; ++ib)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          out[M::gW1(d) + (ob * 16 + 4 * q + e) * H + ib * 16 + r] = dW1[ob][ib][e];
  }
#pragma unroll
  for (int cb = 0; cb < M::HB; ++cb) {
    const float s0 = sum4q(db0[cb]), s1 = sum4q(db1[cb]), so = sum4q(dwo[cb]);
    if (q == 0) {
      out[M::gB0(d) + cb * 16 + r] = s0;
      if (NH == 2) out[M::gB1(d) + cb * 16 + r] = s1;
      out[M::gWo(d) + cb * 16 + r] = so;
    }
  }
  const float sbo = sum4q(dbo);
  if (l == 0) out[M::gBo(d)] = sbo;
}

__device__ __forceinline__ int64_t count_of(const int32_t* n_dev, int64_t n_cap) {
  if (!n_dev) return n_cap;
  const int64_t v = *n_dev;
  return v < 0 ? 0 : (v < n_cap ? v : n_cap);
}

// ----------------------------------------------------------------------------
// Forward: sdf [n]; with dh0 != null also dsdf/dh0 [n, d] (backprop of 1).
// ----------------------------------------------------------------------------
template <int H, int NH, int DP, int T>
__global__ void __launch_bounds__(64 * kWaves)
mlp_fwd_kernel(const float* __restrict__ x, const float* __restrict__ enc, int E, int64_t n_cap,
               const int32_t* __restrict__ n_dev, const float* __restrict__ mlp, float beta,
               float* __restrict__ sdf_out, float* __restrict__ dh0) {
  using M = Mlp<H, NH, DP>;
  extern __shared__ float smem[];
  float* sw = smem;
  const int wid = threadIdx.x >> 6;
  float* bufI = sw + M::kWFloats + wid * T * M::kRow;
  float* bufA = bufI + T * M::LDI;
  float* bufB = bufA + T * M::LDH;
  const int d = 3 + E;
  load_weights<H, NH, DP>(sw, mlp, d);
  const int64_t n = count_of(n_dev, n_cap);
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const int64_t n_tiles = ceil_div(n, (int64_t)T * kWaves);
  __syncthreads();
  for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const int64_t row0 = (tile * kWaves + wid) * T;
    load_rows<DP, T>(bufI, x, enc, E, row0, n);
    __syncthreads();
    floatx4 z[T / 16][M::HB];
    gemm_xwT<T, M::HB, true>(bufI, M::LDI, sw + M::oW0, M::LDI, sw + M::oB0, DP, z);
    floatx4 z1[T / 16][M::HB];
    if (NH == 2) {
#pragma unroll
      for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
        for (int cb = 0; cb < M::HB; ++cb) {
          z1[rb][cb] = z[rb][cb];
#pragma unroll
          for (int e = 0; e < 4; ++e) z[rb][cb][e] = sp(z[rb][cb][e], beta);
        }
      store_frags<T, M::HB>(bufA, M::LDH, z);
      __syncthreads();
      gemm_xwT<T, M::HB, true>(bufA, M::LDH, sw + M::oW1, M::LDH, sw + M::oB1, H, z);
    }
    // output layer
#pragma unroll
    for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float p = 0.f;
#pragma unroll
        for (int cb = 0; cb < M::HB; ++cb) p += sp(z[rb][cb][e], beta) * sw[M::oWo + cb * 16 + r];
        p = sum16(p) + sw[M::oBo];
        const int64_t row = row0 + rb * 16 + 4 * q + e;
        if (r == 0 && row < n) sdf_out[row] = p;
      }
    if (dh0) {
      floatx4 g[T / 16][M::HB];
#pragma unroll
      for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
        for (int cb = 0; cb < M::HB; ++cb)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            g[rb][cb][e] = sw[M::oWo + cb * 16 + r] * sp_d(z[rb][cb][e], beta);
      store_frags<T, M::HB>(bufB, M::LDH, g);
      __syncthreads();
      if (NH == 2) {
        gemm_gw<T, M::HB>(bufB, M::LDH, sw + M::oW1, M::LDH, H, g);
#pragma unroll
        for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
          for (int cb = 0; cb < M::HB; ++cb)
#pragma unroll
            for (int e = 0; e < 4; ++e) g[rb][cb][e] *= sp_d(z1[rb][cb][e], beta);
        __syncthreads();
        store_frags<T, M::HB>(bufB, M::LDH, g);
        __syncthreads();
      }
      floatx4 gh0[T / 16][M::IB];
      gemm_gw<T, M::IB>(bufB, M::LDH, sw + M::oW0, M::LDI, H, gh0);
#pragma unroll
      for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
        for (int ib = 0; ib < M::IB; ++ib)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int64_t row = row0 + rb * 16 + 4 * q + e;
            const int c = ib * 16 + r;
            if (row < n && c < d) dh0[row * d + c] = gh0[rb][ib][e];
          }
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------
// Backward (sdf only): given dL/dsdf, the weight-gradient partials (one row
// per wavefront) and dL/denc [n, E] (+ dL/dx [n,3] if requested).
// ----------------------------------------------------------------------------
template <int H, int NH, int DP, int T>
__global__ void __launch_bounds__(64 * kWaves)
mlp_bwd_kernel(const float* __restrict__ x, const float* __restrict__ enc, int E, int64_t n_cap,
               const int32_t* __restrict__ n_dev, const float* __restrict__ mlp, float beta,
               const float* __restrict__ g_sdf, float* __restrict__ g_enc,
               float* __restrict__ g_x, float* __restrict__ partials) {
  using M = Mlp<H, NH, DP>;
  extern __shared__ float smem[];
  float* sw = smem;
  const int wid = threadIdx.x >> 6;
  float* bufI = sw + M::kWFloats + wid * T * M::kRow;
  float* bufA = bufI + T * M::LDI;
  float* bufB = bufA + T * M::LDH;
  const int d = 3 + E;
  load_weights<H, NH, DP>(sw, mlp, d);
  const int64_t n = count_of(n_dev, n_cap);
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const int64_t n_tiles = ceil_div(n, (int64_t)T * kWaves);
  floatx4 dW0[M::HB][M::IB], dW1[M::HB][M::HB];
#pragma unroll
  for (int a = 0; a < M::HB; ++a) {
#pragma unroll
    for (int b = 0; b < M::IB; ++b) dW0[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < M::HB; ++b) dW1[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
  float db0[M::HB], db1[M::HB], dwo[M::HB], dbo = 0.f;
#pragma unroll
  for (int a = 0; a < M::HB; ++a) db0[a] = db1[a] = dwo[a] = 0.f;
  __syncthreads();
  for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const int64_t row0 = (tile * kWaves + wid) * T;
    load_rows<DP, T>(bufI, x, enc, E, row0, n);
    __syncthreads();
    floatx4 z[T / 16][M::HB], z1[T / 16][M::HB];
    gemm_xwT<T, M::HB, true>(bufI, M::LDI, sw + M::oW0, M::LDI, sw + M::oB0, DP, z);
    if (NH == 2) {
      floatx4 h[T / 16][M::HB];
#pragma unroll
      for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
        for (int cb = 0; cb < M::HB; ++cb) {
          z1[rb][cb] = z[rb][cb];
#pragma unroll
          for (int e = 0; e < 4; ++e) h[rb][cb][e] = sp(z[rb][cb][e], beta);
        }
      store_frags<T, M::HB>(bufA, M::LDH, h);
      __syncthreads();
      gemm_xwT<T, M::HB, true>(bufA, M::LDH, sw + M::oW1, M::LDH, sw + M::oB1, H, z);
    }
    // output layer: g_zL = gs * wo * s'(zL); dwo += gs * hL; dbo += gs
    float gs[T / 16][4];
#pragma unroll
    for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = row0 + rb * 16 + 4 * q + e;
        gs[rb][e] = row < n ? g_sdf[row] : 0.f;
        dbo += (r == 0) ? gs[rb][e] : 0.f;
      }
    floatx4 g[T / 16][M::HB];
#pragma unroll
    for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
      for (int cb = 0; cb < M::HB; ++cb) {
        const float w = sw[M::oWo + cb * 16 + r];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          dwo[cb] += gs[rb][e] * sp(z[rb][cb][e], beta);
          g[rb][cb][e] = gs[rb][e] * w * sp_d(z[rb][cb][e], beta);
        }
      }
    store_frags<T, M::HB>(bufB, M::LDH, g);
    __syncthreads();
    if (NH == 2) {
      colsum_add<T, M::HB>(g, db1);
      gemm_gtA<T, M::HB, M::HB>(bufB, M::LDH, bufA, M::LDH, dW1);
      gemm_gw<T, M::HB>(bufB, M::LDH, sw + M::oW1, M::LDH, H, g);
#pragma unroll
      for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
        for (int cb = 0; cb < M::HB; ++cb)
#pragma unroll
          for (int e = 0; e < 4; ++e) g[rb][cb][e] *= sp_d(z1[rb][cb][e], beta);
      __syncthreads();
      store_frags<T, M::HB>(bufB, M::LDH, g);
      __syncthreads();
    }
    colsum_add<T, M::HB>(g, db0);
    gemm_gtA<T, M::HB, M::IB>(bufB, M::LDH, bufI, M::LDI, dW0);
    if (g_enc || g_x) {
      floatx4 gh0[T / 16][M::IB];
      gemm_gw<T, M::IB>(bufB, M::LDH, sw + M::oW0, M::LDI, H, gh0);
#pragma unroll
      for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
        for (int ib = 0; ib < M::IB; ++ib)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int64_t row = row0 + rb * 16 + 4 * q + e;
            const int c = ib * 16 + r;
            if (row < n) {
              if (c < 3) {
                if (g_x) g_x[row * 3 + c] = gh0[rb][ib][e];
              } else if (c < d && g_enc) {
                g_enc[row * E + (c - 3)] = gh0[rb][ib][e];
              }
            }
          }
    }
    __syncthreads();
  }
  // (dbo was counted once per row by the r == 0 lanes)
  write_partial<H, NH, DP>(partials + (int64_t)(blockIdx.x * kWaves + wid) * M::n_params(d), d,
                           dW0, dW1, db0, db1, dwo, dbo);
}

// ----------------------------------------------------------------------------
// Backward of J = sum_p a_p sdf(x_p) + v_p . grad_x sdf(x_p) (the analytic
// gradient trained through its value, fields.py:107-119 with create_graph):
// forward-over-reverse.  Tangents along v: hd0 = [v, encd] where encd is the
// directional derivative of enc along v (from the hash-grid kernel).  Outputs
// the weight-gradient partials, gb_h0 = dJ/dh0 and gt_h0 = dJ/dhd0 (both [n, d]).
// ----------------------------------------------------------------------------
template <int H, int NH, int DP, int T>
__global__ void __launch_bounds__(64 * kWaves)
mlp_bwd_ad_kernel(const float* __restrict__ x, const float* __restrict__ enc, const float* __restrict__ v,
                  const float* __restrict__ encd, int E, int64_t n_cap,
                  const int32_t* __restrict__ n_dev, const float* __restrict__ mlp, float beta,
                  const float* __restrict__ a_sdf, float* __restrict__ gb_h0,
                  float* __restrict__ gt_h0, float* __restrict__ partials, int out_ld,
                  int out_c0) {
  using M = Mlp<H, NH, DP>;
  extern __shared__ float smem[];
  float* sw = smem;
  const int wid = threadIdx.x >> 6;
  float* bufI = sw + M::kWFloats + wid * 2 * T * M::kRow;
  float* bufA = bufI + T * M::LDI;
  float* bufB = bufA + T * M::LDH;
  float* tbufI = bufB + T * M::LDH;
  float* tbufA = tbufI + T * M::LDI;
  float* tbufB = tbufA + T * M::LDH;
  const int d = 3 + E;
  load_weights<H, NH, DP>(sw, mlp, d);
  const int64_t n = count_of(n_dev, n_cap);
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const int64_t n_tiles = ceil_div(n, (int64_t)T * kWaves);
  floatx4 dW0[M::HB][M::IB], dW1[M::HB][M::HB];
#pragma unroll
  for (int a = 0; a < M::HB; ++a) {
#pragma unroll
    for (int b = 0; b < M::IB; ++b) dW0[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < M::HB; ++b) dW1[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
  float db0[M::HB], db1[M::HB], dwo[M::HB], dbo = 0.f;
#pragma unroll
  for (int a = 0; a < M::HB; ++a) db0[a] = db1[a] = dwo[a] = 0.f;
  __syncthreads();
  for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const int64_t row0 = (tile * kWaves + wid) * T;
    load_rows<DP, T>(bufI, x, enc, E, row0, n);
    load_rows<DP, T>(tbufI, v, encd, E, row0, n);
    __syncthreads();
    floatx4 z[T / 16][M::HB], zd[T / 16][M::HB], z1[T / 16][M::HB], z1d[T / 16][M::HB];
    gemm_xwT<T, M::HB, true>(bufI, M::LDI, sw + M::oW0, M::LDI, sw + M::oB0, DP, z);
    gemm_xwT<T, M::HB, false>(tbufI, M::LDI, sw + M::oW0, M::LDI, nullptr, DP, zd);
    if (NH == 2) {
      floatx4 h[T / 16][M::HB], hd[T / 16][M::HB];
#pragma unroll
      for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
        for (int cb = 0; cb < M::HB; ++cb) {
          z1[rb][cb] = z[rb][cb];
          z1d[rb][cb] = zd[rb][cb];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            h[rb][cb][e] = sp(z[rb][cb][e], beta);
            hd[rb][cb][e] = sp_d(z[rb][cb][e], beta) * zd[rb][cb][e];
          }
        }
      store_frags<T, M::HB>(bufA, M::LDH, h);
      store_frags<T, M::HB>(tbufA, M::LDH, hd);
      __syncthreads();
      gemm_xwT<T, M::HB, true>(bufA, M::LDH, sw + M::oW1, M::LDH, sw + M::oB1, H, z);
      gemm_xwT<T, M::HB, false>(tbufA, M::LDH, sw + M::oW1, M::LDH, nullptr, H, zd);
    }
    // output: J = a*(wo.hL + bo) + wo.hdL;  gb_zL = a wo s' + wo zdL s'';  gt_zL = wo s'
    float as[T / 16][4];
#pragma unroll
    for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = row0 + rb * 16 + 4 * q + e;
        as[rb][e] = row < n ? a_sdf[row] : 0.f;
        dbo += (r == 0) ? as[rb][e] : 0.f;
      }
    floatx4 gb[T / 16][M::HB], gt[T / 16][M::HB];
#pragma unroll
    for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
      for (int cb = 0; cb < M::HB; ++cb) {
        const float w = sw[M::oWo + cb * 16 + r];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float zz = z[rb][cb][e], s1 = sp_d(zz, beta);
          dwo[cb] += as[rb][e] * sp(zz, beta) + s1 * zd[rb][cb][e];
          gb[rb][cb][e] = as[rb][e] * w * s1 + w * zd[rb][cb][e] * sp_dd(zz, beta);
          gt[rb][cb][e] = w * s1;
        }
      }
    store_frags<T, M::HB>(bufB, M::LDH, gb);
    store_frags<T, M::HB>(tbufB, M::LDH, gt);
    __syncthreads();
    if (NH == 2) {
      colsum_add<T, M::HB>(gb, db1);
      gemm_gtA<T, M::HB, M::HB>(bufB, M::LDH, bufA, M::LDH, dW1);
      gemm_gtA<T, M::HB, M::HB>(tbufB, M::LDH, tbufA, M::LDH, dW1);
      gemm_gw<T, M::HB>(bufB, M::LDH, sw + M::oW1, M::LDH, H, gb);
      gemm_gw<T, M::HB>(tbufB, M::LDH, sw + M::oW1, M::LDH, H, gt);
#pragma unroll
      for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
        for (int cb = 0; cb < M::HB; ++cb)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float zz = z1[rb][cb][e], s1 = sp_d(zz, beta);
            gb[rb][cb][e] = gb[rb][cb][e] * s1 + gt[rb][cb][e] * z1d[rb][cb][e] * sp_dd(zz, beta);
            gt[rb][cb][e] = gt[rb][cb][e] * s1;
          }
      __syncthreads();
      store_frags<T, M::HB>(bufB, M::LDH, gb);
      store_frags<T, M::HB>(tbufB, M::LDH, gt);
      __syncthreads();
    }
    colsum_add<T, M::HB>(gb, db0);
    gemm_gtA<T, M::HB, M::IB>(bufB, M::LDH, bufI, M::LDI, dW0);
    gemm_gtA<T, M::HB, M::IB>(tbufB, M::LDH, tbufI, M::LDI, dW0);
    floatx4 o1[T / 16][M::IB], o2[T / 16][M::IB];
    gemm_gw<T, M::IB>(bufB, M::LDH, sw + M::oW0, M::LDI, H, o1);
    gemm_gw<T, M::IB>(tbufB, M::LDH, sw + M::oW0, M::LDI, H, o2);
#pragma unroll
    for (int rb = 0; rb < T / 16; ++rb)
#pragma unroll
      for (int ib = 0; ib < M::IB; ++ib)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t row = row0 + rb * 16 + 4 * q + e;
          const int c = ib * 16 + r;
          if (row < n && c < d && c >= out_c0) {
            if (gb_h0) gb_h0[row * out_ld + (c - out_c0)] = o1[rb][ib][e];
            if (gt_h0) gt_h0[row * out_ld + (c - out_c0)] = o2[rb][ib][e];
          }
        }
    __syncthreads();
  }
  write_partial<H, NH, DP>(partials + (int64_t)(blockIdx.x * kWaves + wid) * M::n_params(d), d,
                           dW0, dW1, db0, db1, dwo, dbo);
}

// Column sums of the partial rows in a fixed order: a workgroup owns 64
// columns; its 16 row groups each sum every 16th row, then the 16 group sums
// are added in order.  Independent of the launch, so results are reproducible.
constexpr int kRedCols = 64, kRedGroups = 16;

__global__ void __launch_bounds__(kRedCols * kRedGroups)
reduce_rows_kernel(const float* __restrict__ partials, int64_t rows, int cols, float* __restrict__ out,
                   int accumulate) {
  __shared__ float part[kRedGroups][kRedCols];
  const int cl = threadIdx.x % kRedCols, g = threadIdx.x / kRedCols;
  const int c = blockIdx.x * kRedCols + cl;
  float s = 0.f;
  if (c < cols) {
#pragma unroll 8
    for (int64_t r = g; r < rows; r += kRedGroups) s += partials[r * cols + c];
  }
  part[g][cl] = s;
  __syncthreads();
  if (g == 0 && c < cols) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < kRedGroups; ++k) t += part[k][cl];
    out[c] = accumulate ? out[c] + t : t;
  }
}

int mlp_reduce_rows(const float* part, int64_t rows, int cols, float* out, int accumulate,
                    hipStream_t s) {
  reduce_rows_kernel<<<(unsigned)ceil_div(cols, kRedCols), kRedCols * kRedGroups, 0, s>>>(
      part, rows, cols, out, accumulate);
  return check_launch("mlp_reduce_rows");
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
constexpr int kGridMax = 256;

struct MlpCfg {
  int H, NH, DP, d;
};

static int mlp_cfg(const char* who, int E, int H, int NH, MlpCfg* c) {
  SN_REQUIRE(H == 32 || H == 64, "%s: d_hidden must be 32 or 64 (got %d)", who, H);
  SN_REQUIRE(NH == 1 || NH == 2, "%s: n_layers must be 1 or 2 (got %d)", who, NH);
  SN_REQUIRE(E >= 0 && 3 + E <= 64, "%s: input width 3+%d must be <= 64", who, E);
  c->H = H;
  c->NH = NH;
  c->d = 3 + E;
  c->DP = c->d <= 32 ? 32 : (c->d <= 48 ? 48 : 64);
  return SN_OK;
}

template <int H, int NH, int DP>
constexpr size_t smem_bytes(int T, int nbuf) {
  return sizeof(float) * (Mlp<H, NH, DP>::kWFloats + (size_t)kWaves * nbuf * T * Mlp<H, NH, DP>::kRow);
}

#define SN_MLP_DISPATCH(c, ...)                                                        \
  do {                                                                                 \
    if (c.H == 32 && c.NH == 1 && c.DP == 32) { constexpr int kH = 32, kNH = 1, kDP = 32; __VA_ARGS__; } \
    else if (c.H == 32 && c.NH == 1 && c.DP == 48) { constexpr int kH = 32, kNH = 1, kDP = 48; __VA_ARGS__; } \
    else if (c.H == 32 && c.NH == 1 && c.DP == 64) { constexpr int kH = 32, kNH = 1, kDP = 64; __VA_ARGS__; } \
    else if (c.H == 32 && c.NH == 2 && c.DP == 32) { constexpr int kH = 32, kNH = 2, kDP = 32; __VA_ARGS__; } \
    else if (c.H == 32 && c.NH == 2 && c.DP == 48) { constexpr int kH = 32, kNH = 2, kDP = 48; __VA_ARGS__; } \
    else if (c.H == 32 && c.NH == 2 && c.DP == 64) { constexpr int kH = 32, kNH = 2, kDP = 64; __VA_ARGS__; } \
    else if (c.H == 64 && c.NH == 1 && c.DP == 32) { constexpr int kH = 64, kNH = 1, kDP = 32; __VA_ARGS__; } \
    else if (c.H == 64 && c.NH == 1 && c.DP == 48) { constexpr int kH = 64, kNH = 1, kDP = 48; __VA_ARGS__; } \
    else if (c.H == 64 && c.NH == 1 && c.DP == 64) { constexpr int kH = 64, kNH = 1, kDP = 64; __VA_ARGS__; } \
    else if (c.H == 64 && c.NH == 2 && c.DP == 32) { constexpr int kH = 64, kNH = 2, kDP = 32; __VA_ARGS__; } \
    else if (c.H == 64 && c.NH == 2 && c.DP == 48) { constexpr int kH = 64, kNH = 2, kDP = 48; __VA_ARGS__; } \
    else { constexpr int kH = 64, kNH = 2, kDP = 64; __VA_ARGS__; }                     \
  } while (0)

// Blocks use more than the default 64 KB of dynamic LDS (gfx950 has 160 KB).
template <typename Kern>
static void allow_smem(Kern k, size_t bytes) {
  // once per kernel (kernels of one signature share this instantiation, so
  // the cache is keyed by the kernel); a refusal surfaces as a launch error
  static const void* done[64];
  static size_t done_bytes[64];
  static int n_done = 0;
  const void* f = reinterpret_cast<const void*>(k);
  for (int i = 0; i < n_done; ++i)
    if (done[i] == f && done_bytes[i] >= bytes) return;
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (n_done < 64) {
    done[n_done] = f;
    done_bytes[n_done++] = bytes;
  }
}

static unsigned grid_blocks(int64_t n_cap, int T) {
  int64_t b = ceil_div(n_cap > 0 ? n_cap : 1, (int64_t)T * kWaves);
  return (unsigned)(b < kGridMax ? b : kGridMax);
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_mlp_n_params(int32_t d_in, int32_t d_hidden, int32_t n_layers) {
  if (d_in <= 0 || d_hidden <= 0 || n_layers < 1 || n_layers > 2) return -1;
  return d_hidden * d_in + d_hidden + (n_layers == 2 ? d_hidden * d_hidden + d_hidden : 0) +
         d_hidden + 1;
}

size_t sn_mlp_bwd_workspace_bytes(int64_t n_cap, int32_t enc_dims, int32_t d_hidden,
                                  int32_t n_layers, int32_t analytic) {
  int np = sn_mlp_n_params(3 + enc_dims, d_hidden, n_layers);
  if (np < 0) return 0;
  const int T = analytic ? 16 : 32;
  return (size_t)grid_blocks(n_cap, T) * kWaves * np * sizeof(float);
}

int sn_mlp_fwd(const float* x, const float* enc, int32_t enc_dims, int64_t n_cap,
               const int32_t* n_dev, const float* mlp, int32_t d_hidden, int32_t n_layers,
               float beta, float* sdf, float* dsdf_dh0, void* stream) {
  clear_error();
  MlpCfg c;
  int rc = mlp_cfg("sn_mlp_fwd", enc_dims, d_hidden, n_layers, &c);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0, "sn_mlp_fwd: n < 0");
  if (n_cap == 0) return SN_OK;
  SN_REQUIRE(x && mlp && sdf && (enc || enc_dims == 0), "sn_mlp_fwd: null pointer");
  constexpr int T = 32;
  hipStream_t s = as_stream(stream);
  SN_MLP_DISPATCH(c, {
    const size_t sm = smem_bytes<kH, kNH, kDP>(T, 1);
    allow_smem(mlp_fwd_kernel<kH, kNH, kDP, T>, sm);
    mlp_fwd_kernel<kH, kNH, kDP, T><<<grid_blocks(n_cap, T), 64 * kWaves, sm, s>>>(
        x, enc, enc_dims, n_cap, n_dev, mlp, beta, sdf, dsdf_dh0);
  });
  return check_launch("sn_mlp_fwd");
}

int sn_mlp_bwd(const float* x, const float* enc, int32_t enc_dims, int64_t n_cap,
               const int32_t* n_dev, const float* mlp, int32_t d_hidden, int32_t n_layers,
               float beta, const float* g_sdf, float* g_enc, float* g_x, float* g_mlp,
               int32_t accumulate, void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  MlpCfg c;
  int rc = mlp_cfg("sn_mlp_bwd", enc_dims, d_hidden, n_layers, &c);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0, "sn_mlp_bwd: n < 0");
  SN_REQUIRE(mlp && g_mlp, "sn_mlp_bwd: null pointer");
  const size_t need = sn_mlp_bwd_workspace_bytes(n_cap, enc_dims, d_hidden, n_layers, 0);
  if (workspace_bytes < need) {
    set_error("sn_mlp_bwd: workspace %zu < %zu bytes", workspace_bytes, need);
    return SN_ESPACE;
  }
  constexpr int T = 32;
  hipStream_t s = as_stream(stream);
  const int np = sn_mlp_n_params(c.d, d_hidden, n_layers);
  const unsigned grid = grid_blocks(n_cap, T);
  float* part = reinterpret_cast<float*>(workspace);
  SN_MLP_DISPATCH(c, {
    const size_t sm = smem_bytes<kH, kNH, kDP>(T, 1);
    allow_smem(mlp_bwd_kernel<kH, kNH, kDP, T>, sm);
    mlp_bwd_kernel<kH, kNH, kDP, T><<<grid, 64 * kWaves, sm, s>>>(
        x, enc, enc_dims, n_cap, n_dev, mlp, beta, g_sdf, g_enc, g_x, part);
  });
  if ((rc = mlp_reduce_rows(part, (int64_t)grid * kWaves, np, g_mlp, accumulate, s))) return rc;
  return check_launch("sn_mlp_bwd");
}

static int mlp_bwd_ad(const float* x, const float* enc, const float* v, const float* enc_dot,
                      int32_t enc_dims, int64_t n_cap, const int32_t* n_dev, const float* mlp,
                      int32_t d_hidden, int32_t n_layers, float beta, const float* g_sdf,
                      float* g_h0, float* g_h0_dot, float* g_mlp, int32_t accumulate,
                      void* workspace, size_t workspace_bytes, void* stream, int out_ld,
                      int out_c0) {
  MlpCfg c;
  int rc = mlp_cfg("sn_mlp_bwd_ad", enc_dims, d_hidden, n_layers, &c);
  if (rc) return rc;
  SN_REQUIRE(n_cap >= 0, "sn_mlp_bwd_ad: n < 0");
  SN_REQUIRE(mlp && g_mlp && v && (enc_dot || enc_dims == 0), "sn_mlp_bwd_ad: null pointer");
  const size_t need = sn_mlp_bwd_workspace_bytes(n_cap, enc_dims, d_hidden, n_layers, 1);
  if (workspace_bytes < need) {
    set_error("sn_mlp_bwd_ad: workspace %zu < %zu bytes", workspace_bytes, need);
    return SN_ESPACE;
  }
  constexpr int T = 16;
  hipStream_t s = as_stream(stream);
  const int np = sn_mlp_n_params(c.d, d_hidden, n_layers);
  const unsigned grid = grid_blocks(n_cap, T);
  float* part = reinterpret_cast<float*>(workspace);
  SN_MLP_DISPATCH(c, {
    const size_t sm = smem_bytes<kH, kNH, kDP>(T, 2);
    allow_smem(mlp_bwd_ad_kernel<kH, kNH, kDP, T>, sm);
    mlp_bwd_ad_kernel<kH, kNH, kDP, T><<<grid, 64 * kWaves, sm, s>>>(
        x, enc, v, enc_dot, enc_dims, n_cap, n_dev, mlp, beta, g_sdf, g_h0, g_h0_dot, part,
        out_ld, out_c0);
  });
  if ((rc = mlp_reduce_rows(part, (int64_t)grid * kWaves, np, g_mlp, accumulate, s))) return rc;
  return check_launch("sn_mlp_bwd_ad");
}

int sn_mlp_bwd_ad(const float* x, const float* enc, const float* v, const float* enc_dot,
                  int32_t enc_dims, int64_t n_cap, const int32_t* n_dev, const float* mlp,
                  int32_t d_hidden, int32_t n_layers, float beta, const float* g_sdf,
                  float* g_h0, float* g_h0_dot, float* g_mlp, int32_t accumulate,
                  void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  return mlp_bwd_ad(x, enc, v, enc_dot, enc_dims, n_cap, n_dev, mlp, d_hidden, n_layers, beta,
                    g_sdf, g_h0, g_h0_dot, g_mlp, accumulate, workspace, workspace_bytes, stream,
                    3 + enc_dims, 0);
}

int sn_mlp_bwd_ad_enc(const float* x, const float* enc, const float* v, const float* enc_dot,
                      int32_t enc_dims, int64_t n_cap, const int32_t* n_dev, const float* mlp,
                      int32_t d_hidden, int32_t n_layers, float beta, const float* g_sdf,
                      float* g_enc, float* g_enc_dot, float* g_mlp, int32_t accumulate,
                      void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  return mlp_bwd_ad(x, enc, v, enc_dot, enc_dims, n_cap, n_dev, mlp, d_hidden, n_layers, beta,
                    g_sdf, g_enc, g_enc_dot, g_mlp, accumulate, workspace, workspace_bytes,
                    stream, enc_dims, 3);
}

}  // extern "C"
