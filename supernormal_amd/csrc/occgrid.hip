// Occupancy-grid update (NA/nerfacc/grid.py:196-239) with contract_inv for the
// AABB case (csrc/contraction.cu:36-61, helpers_contraction.h:23-28).
//
// Deterministic by construction:
//   * scatter-max of the new occupancies goes through an order-preserving
//     float -> uint32 key and integer atomicMax, so duplicate cell indices give
//     the same result in any order (the reference's index_put is racy);
//   * the grid mean is a compensated (Kahan) fp32 sum over fixed 1024-cell
//     chunks, sequential inside a chunk and across chunks, identical to the
//     C oracle (no fp64 on the device).
// Compiled with -ffp-contract=off.
#include "sn_common.h"

namespace sn {

int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);

constexpr int kMeanChunk = 1024;

__device__ __forceinline__ uint32_t f2key(float f) {
  uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
  uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(b);
}

__global__ void __launch_bounds__(256)
cell_points_kernel(const int64_t* __restrict__ idx, int32_t n, const float* __restrict__ jit,
                   int32_t rx, int32_t ry, int32_t rz, const float* __restrict__ roi,
                   float* __restrict__ x) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t c = idx[i];
  const int64_t ryz = (int64_t)ry * rz;
  const int64_t cx = c / ryz;
  const int64_t cy = (c / rz) % ry;
  const int64_t cz = c % rz;
  const float coord[3] = {(float)cx, (float)cy, (float)cz};
  const float res[3] = {(float)rx, (float)ry, (float)rz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float u = (coord[d] + jit[3 * i + d]) / res[d];
    x[3 * i + d] = u * (roi[3 + d] - roi[d]) + roi[d];
  }
}

__global__ void __launch_bounds__(256)
scatter_max_kernel(const int64_t* __restrict__ idx, const float* __restrict__ occ, int32_t n,
                   int32_t n_cells, uint32_t* __restrict__ stage) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t c = idx[i];
  if (c < 0 || c >= n_cells) return;
  atomicMax(&stage[c], f2key(occ[i]));
}

__global__ void __launch_bounds__(256)
ema_apply_kernel(float* __restrict__ occs, int32_t n_cells, const uint32_t* __restrict__ stage,
                 float decay) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n_cells;
       c += (int64_t)gridDim.x * blockDim.x) {
    uint32_t k = stage[c];
    if (k == 0u) continue;
    float a = occs[c] * decay;
    float b = key2f(k);
    // torch.maximum: NaN propagates, otherwise the larger value.
    occs[c] = (a != a || b != b) ? __uint_as_float(0x7fc00000u) : fmaxf(a, b);
  }
}

// Compensated (Kahan) fp32 summation in a fixed order: every 1024-cell chunk
// sequentially, then the chunk sums sequentially; the order and the
// operations (no contraction) are the oracle's, so the mean is bit-exact.
__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
  const float y = x - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

__global__ void __launch_bounds__(256)
chunk_sum_kernel(const float* __restrict__ occs, int32_t n_cells, float* __restrict__ part) {
  const int64_t ch = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t beg = ch * kMeanChunk;
  if (beg >= n_cells) return;
  int64_t end = beg + kMeanChunk;
  if (end > n_cells) end = n_cells;
  float s = 0.f, c = 0.f;
  if (end - beg == kMeanChunk && (reinterpret_cast<uintptr_t>(occs + beg) & 15) == 0) {
    // the same sequential order, 16 bytes per load
    const float4* v4 = reinterpret_cast<const float4*>(occs + beg);
#pragma unroll 8
    for (int k = 0; k < kMeanChunk / 4; ++k) {
      const float4 v = v4[k];
      kahan_add(s, c, v.x);
      kahan_add(s, c, v.y);
      kahan_add(s, c, v.z);
      kahan_add(s, c, v.w);
    }
  } else {
    for (int64_t k = beg; k < end; ++k) kahan_add(s, c, occs[k]);
  }
  part[ch] = s;
}

// One workgroup: tiles of the partial sums are staged in LDS by all lanes,
// then lane 0 adds them in order (the loads no longer wait one by one).
constexpr int kThreThreads = 256, kThreTile = 4096;
__global__ void __launch_bounds__(kThreThreads)
thre_kernel(const float* __restrict__ part, int64_t n_chunks, int32_t n_cells, float thre,
            float* __restrict__ thre_out) {
  __shared__ float tile[kThreTile];
  float s = 0.f, c = 0.f;
  for (int64_t t0 = 0; t0 < n_chunks; t0 += kThreTile) {  // uniform
    const int64_t m = n_chunks - t0 < kThreTile ? n_chunks - t0 : kThreTile;
    for (int64_t i = threadIdx.x; i < m; i += kThreThreads) tile[i] = part[t0 + i];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int64_t i = 0; i < m; ++i) kahan_add(s, c, tile[i]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float mean = s / (float)n_cells;
    *thre_out = fminf(mean, thre);
  }
}

__global__ void __launch_bounds__(256)
binarize_kernel(const float* __restrict__ occs, int32_t n_cells,
                const float* __restrict__ thre_in, uint8_t* __restrict__ binary) {
  const float t = *thre_in;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n_cells;
       c += (int64_t)gridDim.x * blockDim.x)
    binary[c] = occs[c] > t ? 1 : 0;
}

__global__ void __launch_bounds__(256)
zero_u32_kernel(uint32_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_occ_cell_points(const int64_t* cell_idx, int32_t n, const float* jitter, int32_t rx,
                       int32_t ry, int32_t rz, const float* roi_aabb, float* x, void* stream) {
  clear_error();
  SN_REQUIRE(n >= 0 && rx > 0 && ry > 0 && rz > 0, "sn_occ_cell_points: bad sizes");
  if (n == 0) return SN_OK;
  SN_REQUIRE(cell_idx && jitter && roi_aabb && x, "sn_occ_cell_points: null pointer");
  cell_points_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, as_stream(stream)>>>(
      cell_idx, n, jitter, rx, ry, rz, roi_aabb, x);
  return check_launch("sn_occ_cell_points");
}

size_t sn_occ_workspace_bytes(int32_t n_cells) {
  size_t a = align_up((size_t)(n_cells > 0 ? n_cells : 1) * sizeof(uint32_t), 256);
  size_t chunks = (size_t)ceil_div(n_cells > 0 ? n_cells : 1, kMeanChunk);
  size_t b = align_up(chunks * sizeof(float), 256) + 256;
  return a > b ? a : b;
}

int sn_occ_ema_max(float* occs, int32_t n_cells, const int64_t* cell_idx, const float* occ,
                   int32_t n, float decay, void* ws, size_t ws_bytes, void* stream) {
  clear_error();
  SN_REQUIRE(n_cells > 0 && n >= 0, "sn_occ_ema_max: bad sizes");
  SN_REQUIRE(occs && (n == 0 || (cell_idx && occ)), "sn_occ_ema_max: null pointer");
  if (ws_bytes < sn_occ_workspace_bytes(n_cells)) {
    set_error("sn_occ_ema_max: workspace %zu < %zu", ws_bytes, sn_occ_workspace_bytes(n_cells));
    return SN_ESPACE;
  }
  if (n == 0) return SN_OK;
  hipStream_t s = as_stream(stream);
  uint32_t* stage = reinterpret_cast<uint32_t*>(ws);
  zero_u32_kernel<<<grid_for(n_cells, 256), 256, 0, s>>>(stage, n_cells);
  scatter_max_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(cell_idx, occ, n, n_cells,
                                                                 stage);
  ema_apply_kernel<<<grid_for(n_cells, 256), 256, 0, s>>>(occs, n_cells, stage, decay);
  return check_launch("sn_occ_ema_max");
}

int sn_occ_binarize(const float* occs, int32_t n_cells, float thre, uint8_t* binary,
                    float* thre_out, void* ws, size_t ws_bytes, void* stream) {
  clear_error();
  SN_REQUIRE(n_cells > 0, "sn_occ_binarize: n_cells <= 0");
  SN_REQUIRE(occs && binary, "sn_occ_binarize: null pointer");
  if (ws_bytes < sn_occ_workspace_bytes(n_cells)) {
    set_error("sn_occ_binarize: workspace %zu < %zu", ws_bytes,
              sn_occ_workspace_bytes(n_cells));
    return SN_ESPACE;
  }
  hipStream_t s = as_stream(stream);
  const int64_t n_chunks = ceil_div(n_cells, kMeanChunk);
  float* part = reinterpret_cast<float*>(ws);
  float* t_dev = thre_out;
  if (!t_dev) t_dev = reinterpret_cast<float*>((char*)ws + align_up(n_chunks * sizeof(float), 256));
  chunk_sum_kernel<<<(unsigned)ceil_div(n_chunks, 256), 256, 0, s>>>(occs, n_cells, part);
  thre_kernel<<<1, kThreThreads, 0, s>>>(part, n_chunks, n_cells, thre, t_dev);
  binarize_kernel<<<grid_for(n_cells, 256), 256, 0, s>>>(occs, n_cells, t_dev, binary);
  return check_launch("sn_occ_binarize");
}

}  // extern "C"
