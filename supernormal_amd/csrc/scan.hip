// Device-wide exclusive scans: pack_info, counts -> packed_info, stream
// compaction of marching samples, occupied-cell lists, and the small fill
// kernels used instead of hipMemsetAsync (which misbehaves inside captured
// graphs).
//
// Scan layout: a workgroup of 256 lanes owns a tile of 1024 consecutive
// items (4 per lane); pass 1 writes one total per tile, pass 2 scans the tile
// totals in one workgroup, pass 3 rescans each tile and adds its offset.  All
// integer, so the result is exact and independent of the launch shape.
//
// Every kernel takes a capacity n_cap (which sizes the launch) and an
// optional device count n_dev (the number of valid items, <= n_cap), so the
// scans run without a host round trip inside a sync-free iteration.
#include "sn_common.h"

namespace sn {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int64_t eff_n(int64_t n_cap, const int32_t* n_dev) {
  if (!n_dev) return n_cap;
  const int64_t v = *n_dev;
  return v < 0 ? 0 : (v < n_cap ? v : n_cap);
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void fill_f32_kernel(float* p, int64_t n, float v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return SN_OK;
  fill_i32_kernel<<<grid_for(n, 256), 256, 0, s>>>(p, n, v);
  return check_launch("fill_i32");
}

// p[i] = v for i in [lo * lo_mult, min(hi * hi_mult, cap)) with lo / hi read
// from the device (counts of a sync-free iteration).
__global__ void fill_span_kernel(float* __restrict__ p, int64_t cap,
                                 const int32_t* __restrict__ lo_dev, int32_t lo_mult,
                                 const int32_t* __restrict__ hi_dev, int32_t hi_mult, float v) {
  const int64_t lo = (int64_t)max(*lo_dev, 0) * lo_mult;
  int64_t hi = (int64_t)max(*hi_dev, 0) * hi_mult;
  hi = hi < cap ? hi : cap;
  for (int64_t i = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hi;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

int fill_f32(float* p, int64_t n, float v, hipStream_t s) {
  if (n <= 0) return SN_OK;
  fill_f32_kernel<<<grid_for(n, 256), 256, 0, s>>>(p, n, v);
  return check_launch("fill_f32");
}

__device__ inline int32_t wave_incl_scan(int32_t v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int32_t t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan of one value per lane across the workgroup; returns the
// exclusive prefix and writes the workgroup total to *total.
__device__ inline int32_t block_excl_scan(int32_t v, int32_t* total) {
  __shared__ int32_t wave_tot[kScanBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  int32_t incl = wave_incl_scan(v);
  if (lane == kWave - 1) wave_tot[wid] = incl;
  __syncthreads();
  int32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / kWave; ++w) {
    int32_t t = wave_tot[w];
    if (w < wid) off += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// Item sources.
struct I32Src {
  const int32_t* p;
  __device__ int32_t operator()(int64_t i) const { return p[i]; }
};
struct U8Src {
  const uint8_t* p;
  __device__ int32_t operator()(int64_t i) const { return p[i] ? 1 : 0; }
};

template <typename Src>
__global__ void __launch_bounds__(kScanBlock)
tile_totals_kernel(Src src, int64_t n_cap, const int32_t* n_dev, int32_t* tile_tot) {
  const int64_t n = eff_n(n_cap, n_dev);
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int32_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j)
    if (base + j < n) s += src(base + j);
  int32_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) tile_tot[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of the tile totals in place, grand total to
// tile_tot[n_tiles] and (optionally) *total_out.
__global__ void __launch_bounds__(kScanBlock)
scan_tiles_kernel(int32_t* tile_tot, int64_t n_cap, const int32_t* n_dev, int32_t* total_out) {
  const int64_t n_tiles = ceil_div(n_cap, kScanTile);
  int32_t carry = 0;
  for (int64_t c = 0; c < n_tiles; c += kScanTile) {
    const int64_t base = c + threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      v[j] = (base + j < n_tiles) ? tile_tot[base + j] : 0;
      s += v[j];
    }
    int32_t tot;
    int32_t off = block_excl_scan(s, &tot) + carry;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
      if (base + j < n_tiles) tile_tot[base + j] = off;
      off += v[j];
    }
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    tile_tot[n_tiles] = carry;
    if (total_out) *total_out = carry;
  }
  (void)n_dev;
}

// Pass 3 writers.
struct PackedInfoDst {  // (base, count) rows
  int32_t* p;
  __device__ void operator()(int64_t i, int32_t excl, int32_t v) const {
    p[2 * i] = excl;
    p[2 * i + 1] = v;
  }
};
struct OffsetDst {
  int32_t* p;
  __device__ void operator()(int64_t i, int32_t excl, int32_t) const { p[i] = excl; }
};
// offset of flagged items, -1 elsewhere
struct FlagOffsetDst {
  int32_t* p;
  __device__ void operator()(int64_t i, int32_t excl, int32_t v) const { p[i] = v ? excl : -1; }
};

template <typename Src, typename Dst>
__global__ void __launch_bounds__(kScanBlock)
tile_scan_write_kernel(Src src, int64_t n_cap, const int32_t* n_dev, const int32_t* tile_off,
                       Dst dst) {
  const int64_t n = eff_n(n_cap, n_dev);
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int32_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = (base + j < n) ? src(base + j) : 0;
    s += v[j];
  }
  int32_t tot;
  int32_t off = block_excl_scan(s, &tot) + tile_off[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (base + j < n) dst(base + j, off, v[j]);
    off += v[j];
  }
}

size_t scan_ws_bytes(int64_t n) {
  int64_t tiles = ceil_div(n > 0 ? n : 1, kScanTile);
  return align_up((size_t)(tiles + 1) * sizeof(int32_t), 256);
}

template <typename Src, typename Dst>
int exclusive_scan(Src src, int64_t n_cap, const int32_t* n_dev, Dst dst, int32_t* total,
                   void* ws, size_t ws_bytes, hipStream_t s, const char* what) {
  if (ws_bytes < scan_ws_bytes(n_cap)) {
    set_error("%s: workspace %zu < %zu bytes", what, ws_bytes, scan_ws_bytes(n_cap));
    return SN_ESPACE;
  }
  int32_t* tiles = reinterpret_cast<int32_t*>(ws);
  if (n_cap <= 0) {
    if (total) return fill_i32(total, 1, 0, s);
    return SN_OK;
  }
  const int64_t n_tiles = ceil_div(n_cap, kScanTile);
  tile_totals_kernel<<<(unsigned)n_tiles, kScanBlock, 0, s>>>(src, n_cap, n_dev, tiles);
  scan_tiles_kernel<<<1, kScanBlock, 0, s>>>(tiles, n_cap, n_dev, total);
  tile_scan_write_kernel<<<(unsigned)n_tiles, kScanBlock, 0, s>>>(src, n_cap, n_dev, tiles, dst);
  return check_launch(what);
}

// offsets[i] = number of flagged items before i when flags[i] != 0, else -1;
// *total = number of flagged items (device).  Used by the sync-free iteration.
int scan_flags_dev(const uint8_t* flags, int64_t n_cap, const int32_t* n_dev, int32_t* offsets,
                   int32_t* total, void* ws, size_t ws_bytes, hipStream_t s) {
  return exclusive_scan(U8Src{flags}, n_cap, n_dev, FlagOffsetDst{offsets}, total, ws, ws_bytes,
                        s, "scan_flags_dev");
}

__global__ void count_indices_kernel(const int64_t* idx, int64_t n_cap, const int32_t* n_dev,
                                     int32_t n_rays, int32_t* counts, int32_t* n_bad) {
  const int64_t n = eff_n(n_cap, n_dev);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = idx[i];
    if (r >= 0 && r < n_rays)
      atomicAdd(&counts[r], 1);
    else if (n_bad)
      atomicAdd(n_bad, 1);
  }
}

__global__ void compact_kernel(const uint8_t* keep, int64_t n_cap, const int32_t* n_dev,
                               const int32_t* offs, const int64_t* ri, const float* t0,
                               const float* t1, int64_t* o_ri, float* o_t0, float* o_t1) {
  const int64_t n = eff_n(n_cap, n_dev);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (keep[i]) {
      int32_t o = offs[i];
      o_ri[o] = ri[i];
      o_t0[o] = t0[i];
      o_t1[o] = t1[i];
    }
  }
}

// packed_out[r] = (first kept sample of ray r, kept count) from the rays'
// pre-compaction packed_in and the exclusive keep offsets.
__global__ void compact_packed_kernel(const int32_t* __restrict__ packed_in, int32_t n_rays,
                                      int64_t n_cap, const int32_t* __restrict__ n_dev,
                                      const int32_t* __restrict__ offs,
                                      const int32_t* __restrict__ total,
                                      int32_t* __restrict__ packed_out) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rays) return;
  const int64_t n = eff_n(n_cap, n_dev);
  const int64_t b = packed_in[2 * r], c = packed_in[2 * r + 1];
  const int32_t t = *total;
  const int32_t s = b < n ? offs[b] : t;
  const int32_t e = b + c < n ? offs[b + c] : t;
  packed_out[2 * r] = s;
  packed_out[2 * r + 1] = e - s;
}

__global__ void nonzero_write_kernel(const uint8_t* keep, int64_t n, const int32_t* offs,
                                     int64_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (keep[i]) out[offs[i]] = i;
}

}  // namespace sn

using namespace sn;

extern "C" {

size_t sn_scan_workspace_bytes(int64_t n) { return scan_ws_bytes(n); }

int sn_fill_span_dev(float* p, int64_t cap, const int32_t* lo_dev, int32_t lo_mult,
                     const int32_t* hi_dev, int32_t hi_mult, float value, void* stream) {
  clear_error();
  SN_REQUIRE(cap >= 0 && lo_mult >= 0 && hi_mult >= 0, "sn_fill_span_dev: negative size");
  if (cap == 0) return SN_OK;
  SN_REQUIRE(p && lo_dev && hi_dev, "sn_fill_span_dev: null pointer");
  fill_span_kernel<<<grid_for(cap, 256), 256, 0, as_stream(stream)>>>(p, cap, lo_dev, lo_mult,
                                                                      hi_dev, hi_mult, value);
  return check_launch("sn_fill_span_dev");
}

int sn_counts_to_packed_info(const int32_t* counts, int32_t n, int32_t* packed_info,
                             int32_t* total, void* ws, size_t ws_bytes, void* stream) {
  clear_error();
  SN_REQUIRE(n >= 0, "sn_counts_to_packed_info: n=%d < 0", n);
  SN_REQUIRE(n == 0 || (counts && packed_info), "sn_counts_to_packed_info: null pointer");
  return exclusive_scan(I32Src{counts}, n, nullptr, PackedInfoDst{packed_info}, total, ws,
                        ws_bytes, as_stream(stream), "sn_counts_to_packed_info");
}

size_t sn_pack_info_workspace_bytes(int32_t n_rays) {
  return align_up((size_t)(n_rays > 0 ? n_rays : 1) * sizeof(int32_t), 256) +
         scan_ws_bytes(n_rays);
}

int sn_pack_info_dev(const int64_t* ray_indices, int64_t n_cap, const int32_t* n_dev,
                     int32_t n_rays, int32_t* packed_info, int32_t* n_bad, void* ws,
                     size_t ws_bytes, void* stream) {
  clear_error();
  SN_REQUIRE(n_rays >= 0 && n_cap >= 0, "sn_pack_info: negative size");
  SN_REQUIRE(n_cap == 0 || ray_indices, "sn_pack_info: null ray_indices");
  SN_REQUIRE(n_rays == 0 || packed_info, "sn_pack_info: null packed_info");
  if (ws_bytes < sn_pack_info_workspace_bytes(n_rays)) {
    set_error("sn_pack_info: workspace %zu < %zu bytes", ws_bytes,
              sn_pack_info_workspace_bytes(n_rays));
    return SN_ESPACE;
  }
  hipStream_t s = as_stream(stream);
  int32_t* counts = reinterpret_cast<int32_t*>(ws);
  size_t cbytes = align_up((size_t)(n_rays > 0 ? n_rays : 1) * sizeof(int32_t), 256);
  int rc = fill_i32(counts, n_rays, 0, s);
  if (rc) return rc;
  if (n_bad && (rc = fill_i32(n_bad, 1, 0, s))) return rc;
  if (n_cap > 0) {
    count_indices_kernel<<<grid_for(n_cap, 256), 256, 0, s>>>(ray_indices, n_cap, n_dev,
                                                              n_rays, counts, n_bad);
    if ((rc = check_launch("sn_pack_info/count"))) return rc;
  }
  return exclusive_scan(I32Src{counts}, n_rays, nullptr, PackedInfoDst{packed_info}, nullptr,
                        (char*)ws + cbytes, ws_bytes - cbytes, s, "sn_pack_info");
}

int sn_pack_info(const int64_t* ray_indices, int64_t n_samples, int32_t n_rays,
                 int32_t* packed_info, int32_t* n_bad, void* ws, size_t ws_bytes,
                 void* stream) {
  return sn_pack_info_dev(ray_indices, n_samples, nullptr, n_rays, packed_info, n_bad, ws,
                          ws_bytes, stream);
}

size_t sn_compact_workspace_bytes(int64_t n) {
  return align_up((size_t)(n > 0 ? n : 1) * sizeof(int32_t), 256) + scan_ws_bytes(n);
}

int sn_compact_samples_dev(const uint8_t* keep, int32_t n_cap, const int32_t* n_dev,
                           const int64_t* ray_indices, const float* t_starts,
                           const float* t_ends, int64_t* out_ray_indices, float* out_t_starts,
                           float* out_t_ends, int32_t* n_out, const int32_t* packed_in,
                           int32_t n_rays, int32_t* packed_out, void* ws, size_t ws_bytes,
                           void* stream) {
  clear_error();
  SN_REQUIRE(n_cap >= 0, "sn_compact_samples: n < 0");
  SN_REQUIRE(n_out, "sn_compact_samples: null n_out");
  if (ws_bytes < sn_compact_workspace_bytes(n_cap)) {
    set_error("sn_compact_samples: workspace %zu < %zu bytes", ws_bytes,
              sn_compact_workspace_bytes(n_cap));
    return SN_ESPACE;
  }
  hipStream_t s = as_stream(stream);
  int32_t* offs = reinterpret_cast<int32_t*>(ws);
  size_t obytes = align_up((size_t)(n_cap > 0 ? n_cap : 1) * sizeof(int32_t), 256);
  int rc = exclusive_scan(U8Src{keep}, n_cap, n_dev, OffsetDst{offs}, n_out,
                          (char*)ws + obytes, ws_bytes - obytes, s, "sn_compact_samples/scan");
  if (rc || n_cap == 0) return rc;
  compact_kernel<<<grid_for(n_cap, 256), 256, 0, s>>>(keep, n_cap, n_dev, offs, ray_indices,
                                                      t_starts, t_ends, out_ray_indices,
                                                      out_t_starts, out_t_ends);
  if (packed_in && packed_out && n_rays > 0)
    compact_packed_kernel<<<(unsigned)ceil_div(n_rays, 256), 256, 0, s>>>(
        packed_in, n_rays, n_cap, n_dev, offs, n_out, packed_out);
  return check_launch("sn_compact_samples");
}

int sn_compact_samples(const uint8_t* keep, int32_t n, const int64_t* ray_indices,
                       const float* t_starts, const float* t_ends, int64_t* out_ray_indices,
                       float* out_t_starts, float* out_t_ends, int32_t* n_out, void* ws,
                       size_t ws_bytes, void* stream) {
  return sn_compact_samples_dev(keep, n, nullptr, ray_indices, t_starts, t_ends, out_ray_indices,
                                out_t_starts, out_t_ends, n_out, nullptr, 0, nullptr, ws,
                                ws_bytes, stream);
}

int sn_occ_nonzero(const uint8_t* binary, int32_t n_cells, int64_t* out_idx, int32_t* count,
                   void* ws, size_t ws_bytes, void* stream) {
  clear_error();
  SN_REQUIRE(n_cells >= 0, "sn_occ_nonzero: n_cells < 0");
  SN_REQUIRE(count && (n_cells == 0 || (binary && out_idx)), "sn_occ_nonzero: null pointer");
  if (ws_bytes < sn_compact_workspace_bytes(n_cells)) {
    set_error("sn_occ_nonzero: workspace %zu < %zu bytes", ws_bytes,
              sn_compact_workspace_bytes(n_cells));
    return SN_ESPACE;
  }
  hipStream_t s = as_stream(stream);
  int32_t* offs = reinterpret_cast<int32_t*>(ws);
  size_t obytes = align_up((size_t)(n_cells > 0 ? n_cells : 1) * sizeof(int32_t), 256);
  int rc = exclusive_scan(U8Src{binary}, n_cells, nullptr, OffsetDst{offs}, count,
                          (char*)ws + obytes, ws_bytes - obytes, s, "sn_occ_nonzero/scan");
  if (rc || n_cells == 0) return rc;
  nonzero_write_kernel<<<grid_for(n_cells, 256), 256, 0, s>>>(binary, n_cells, offs, out_idx);
  return check_launch("sn_occ_nonzero");
}

}  // extern "C"
