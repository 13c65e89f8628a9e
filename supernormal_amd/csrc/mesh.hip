// Surface extraction from a sampled field (renderer.py:9-34 with the mesher
// replaced): marching tetrahedra on an rx x ry x rz grid of values.  Each cube
// is split into 6 tetrahedra around the main diagonal (corners 0 and 6); a
// tetrahedron with one corner on its own side of the iso-level yields one
// triangle, two against two yields a quad (two triangles).  Triangles are
// oriented so that their normal points from the inside corners (value < iso)
// to the outside ones.
//
// Two passes over the cubes: count triangles per cube, exclusive scan
// (sn_counts_to_packed_info), then each cube writes its triangles at its
// base.  Output: triangle soup, 9 floats per triangle (3 world-space
// vertices); within a cube the order is tetrahedra 0..5, then the case order.
#include "sn_common.h"

namespace sn {

__constant__ int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__constant__ int kTet[6][4] = {{0, 5, 1, 6}, {0, 1, 2, 6}, {0, 2, 3, 6},
                               {0, 3, 7, 6}, {0, 7, 4, 6}, {0, 4, 5, 6}};

__device__ __forceinline__ int tet_tris(int code) {
  if (code == 0 || code == 15) return 0;
  const int pc = __popc(code);
  return pc == 2 ? 2 : 1;
}

struct MtGrid {
  int rx, ry, rz;
  float iso;
  float mn[3], step[3];  // world = mn + grid * step
};

__device__ __forceinline__ void cube_values(const float* __restrict__ w, const MtGrid& g,
                                            int i, int j, int k, float (&v)[8]) {
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int64_t x = i + kCorner[c][0], y = j + kCorner[c][1], z = k + kCorner[c][2];
    v[c] = w[(x * g.ry + y) * g.rz + z] - g.iso;
  }
}

__global__ void __launch_bounds__(256)
mt_count_kernel(const float* __restrict__ w, MtGrid g, int32_t* __restrict__ counts) {
  const int64_t nx = g.rx - 1, ny = g.ry - 1, nz = g.rz - 1;
  const int64_t n = nx * ny * nz;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / (ny * nz)), j = (int)((t / nz) % ny), k = (int)(t % nz);
    float v[8];
    cube_values(w, g, i, j, k, v);
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      int code = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a) code |= (v[kTet[q][a]] < 0.f ? 1 : 0) << a;
      cnt += tet_tris(code);
    }
    counts[t] = cnt;
  }
}

__device__ __forceinline__ void edge_point(const float (&tv)[4], const float (&tp)[4][3], int a,
                                           int b, float (&out)[3]) {
  const float s = tv[a] / (tv[a] - tv[b]);
#pragma unroll
  for (int d = 0; d < 3; ++d) out[d] = tp[a][d] + (tp[b][d] - tp[a][d]) * s;
}

__device__ __forceinline__ void emit_tri(float* __restrict__ out, const float (&p0)[3],
                                         const float (&p1)[3], const float (&p2)[3],
                                         const float (&od)[3], const MtGrid& g) {
  const float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
  const float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
  const float nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2],
              nz = e1[0] * e2[1] - e1[1] * e2[0];
  const bool flip = nx * od[0] + ny * od[1] + nz * od[2] < 0.f;
  const float* q[3] = {p0, flip ? p2 : p1, flip ? p1 : p2};
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) out[3 * c + d] = g.mn[d] + q[c][d] * g.step[d];
}

__global__ void __launch_bounds__(256)
mt_emit_kernel(const float* __restrict__ w, MtGrid g, const int32_t* __restrict__ packed,
               float* __restrict__ tris) {
  const int64_t nx = g.rx - 1, ny = g.ry - 1, nz = g.rz - 1;
  const int64_t n = nx * ny * nz;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (packed[2 * t + 1] == 0) continue;
    int64_t o = packed[2 * t];
    const int i = (int)(t / (ny * nz)), j = (int)((t / nz) % ny), k = (int)(t % nz);
    float v[8];
    cube_values(w, g, i, j, k, v);
    for (int q = 0; q < 6; ++q) {
      float tv[4], tp[4][3];
      int code = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int c = kTet[q][a];
        tv[a] = v[c];
        tp[a][0] = (float)(i + kCorner[c][0]);
        tp[a][1] = (float)(j + kCorner[c][1]);
        tp[a][2] = (float)(k + kCorner[c][2]);
        code |= (tv[a] < 0.f ? 1 : 0) << a;
      }
      const int nt = tet_tris(code);
      if (nt == 0) continue;
      // inside -> outside direction: mean of outside corners - mean of inside corners
      float od[3] = {0.f, 0.f, 0.f};
      const int n_in = __popc(code);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        float si = 0.f, so = 0.f;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          if (code & (1 << a)) si += tp[a][d];
          else so += tp[a][d];
        }
        od[d] = so / (float)(4 - n_in) - si / (float)n_in;
      }
      if (nt == 1) {
        // the corner alone on its side
        const int lone = (n_in == 1) ? __ffs(code) - 1 : __ffs(~code & 15) - 1;
        int others[3], m = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
          if (a != lone) others[m++] = a;
        float p0[3], p1[3], p2[3];
        edge_point(tv, tp, lone, others[0], p0);
        edge_point(tv, tp, lone, others[1], p1);
        edge_point(tv, tp, lone, others[2], p2);
        emit_tri(tris + 9 * o, p0, p1, p2, od, g);
        ++o;
      } else {
        // corners a < b on one side, c0 < c1 on the other; a = 0 by symmetry of the pair
        const int pair = (code & 1) ? code : (~code & 15);  // the side holding corner 0
        const int a = 0, b = __ffs(pair & ~1) - 1;
        int cs[2], m = 0;
#pragma unroll
        for (int x = 1; x < 4; ++x)
          if (x != b) cs[m++] = x;
        float p0[3], p1[3], p2[3], p3[3];
        edge_point(tv, tp, a, cs[0], p0);
        edge_point(tv, tp, a, cs[1], p1);
        edge_point(tv, tp, b, cs[1], p2);
        edge_point(tv, tp, b, cs[0], p3);
        emit_tri(tris + 9 * o, p0, p1, p2, od, g);
        emit_tri(tris + 9 * (o + 1), p0, p2, p3, od, g);
        o += 2;
      }
    }
  }
}

static int mt_grid(const char* who, int32_t rx, int32_t ry, int32_t rz, float iso,
                   const float* bmin, const float* bmax, MtGrid* g) {
  SN_REQUIRE(rx >= 2 && ry >= 2 && rz >= 2, "%s: grid must be at least 2^3", who);
  SN_REQUIRE((int64_t)(rx - 1) * (ry - 1) * (rz - 1) < (1ll << 31), "%s: grid too large", who);
  g->rx = rx;
  g->ry = ry;
  g->rz = rz;
  g->iso = iso;
  const int32_t r[3] = {rx, ry, rz};
  for (int d = 0; d < 3; ++d) {
    g->mn[d] = bmin[d];
    g->step[d] = (bmax[d] - bmin[d]) / (float)(r[d] - 1);
  }
  return SN_OK;
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_mt_count(const float* values, int32_t rx, int32_t ry, int32_t rz, float iso,
                int32_t* counts, void* stream) {
  clear_error();
  const float z[3] = {0.f, 0.f, 0.f}, o[3] = {1.f, 1.f, 1.f};
  MtGrid g;
  int rc = mt_grid("sn_mt_count", rx, ry, rz, iso, z, o, &g);
  if (rc) return rc;
  SN_REQUIRE(values && counts, "sn_mt_count: null pointer");
  const int64_t n = (int64_t)(rx - 1) * (ry - 1) * (rz - 1);
  mt_count_kernel<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(values, g, counts);
  return check_launch("sn_mt_count");
}

int sn_mt_emit(const float* values, int32_t rx, int32_t ry, int32_t rz, float iso,
               const float* bound_min, const float* bound_max, const int32_t* packed_info,
               float* triangles, void* stream) {
  clear_error();
  SN_REQUIRE(bound_min && bound_max, "sn_mt_emit: null bounds");
  MtGrid g;
  int rc = mt_grid("sn_mt_emit", rx, ry, rz, iso, bound_min, bound_max, &g);
  if (rc) return rc;
  SN_REQUIRE(values && packed_info && triangles, "sn_mt_emit: null pointer");
  const int64_t n = (int64_t)(rx - 1) * (ry - 1) * (rz - 1);
  mt_emit_kernel<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(values, g, packed_info,
                                                                   triangles);
  return check_launch("sn_mt_emit");
}

}  // extern "C"
