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
#include "mc_tables.h"

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

// ---------------------------------------------------------------------------
// Marching cubes (renderer.py:27-34 calls PyMCubes' marching_cubes): the case
// of a cube is sum((value < iso) << corner) and mc_tables.h gives its
// triangles as edge triples.  A vertex lies on each edge whose corners are on
// different sides, at the linear interpolation of the values, computed from
// the edge's lower grid point towards the higher one, so the two to four
// cubes that share an edge write the bit-identical vertex; its id is
// 3 * (linear index of the lower grid point) + axis (the indexed mesh
// PyMCubes returns has one vertex per such edge).
// ---------------------------------------------------------------------------
// edge -> (lower corner, higher corner, axis) in the corner numbering of mc_tables.h
__constant__ int kMcEdgeLo[12] = {0, 1, 3, 0, 4, 5, 7, 4, 0, 1, 2, 3};
__constant__ int kMcEdgeHi[12] = {1, 2, 2, 3, 5, 6, 6, 7, 4, 5, 6, 7};
__constant__ int kMcEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};

__device__ __forceinline__ int mc_case(const float (&v)[8]) {
  int c = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) c |= (v[q] < 0.f ? 1 : 0) << q;
  return c;
}

__global__ void __launch_bounds__(256)
mc_count_kernel(const float* __restrict__ w, MtGrid g, int32_t* __restrict__ counts) {
  const int64_t nx = g.rx - 1, ny = g.ry - 1, nz = g.rz - 1;
  const int64_t n = nx * ny * nz;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / (ny * nz)), j = (int)((t / nz) % ny), k = (int)(t % nz);
    float v[8];
    cube_values(w, g, i, j, k, v);
    counts[t] = kMcTriCount[mc_case(v)];
  }
}

__global__ void __launch_bounds__(256)
mc_emit_kernel(const float* __restrict__ w, MtGrid g, const int32_t* __restrict__ packed,
               float* __restrict__ tris, int64_t* __restrict__ vert_ids) {
  const int64_t nx = g.rx - 1, ny = g.ry - 1, nz = g.rz - 1;
  const int64_t n = nx * ny * nz;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int cnt = packed[2 * t + 1];
    if (cnt == 0) continue;
    const int64_t o = packed[2 * t];
    const int i = (int)(t / (ny * nz)), j = (int)((t / nz) % ny), k = (int)(t % nz);
    float v[8];
    cube_values(w, g, i, j, k, v);
    const int cs = mc_case(v);
    for (int q = 0; q < cnt; ++q) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int e = kMcTriTable[cs][3 * q + c];
        const int lo = kMcEdgeLo[e], hc = kMcEdgeHi[e], ax = kMcEdgeAxis[e];
        const int li = i + kCorner[lo][0], lj = j + kCorner[lo][1], lk = k + kCorner[lo][2];
        const float s = v[lo] / (v[lo] - v[hc]);
        const int gi[3] = {li, lj, lk};
        float* out = tris + 9 * (o + q) + 3 * c;
#pragma unroll
        for (int d = 0; d < 3; ++d)
          out[d] = g.mn[d] + ((float)gi[d] + (d == ax ? s : 0.f)) * g.step[d];
        if (vert_ids)
          vert_ids[3 * (o + q) + c] =
              3 * (((int64_t)li * g.ry + lj) * g.rz + lk) + ax;
      }
    }
  }
}

// First hit of each ray with a marching-cubes mesh (exp_runner.py:618-630 uses
// trimesh's ray.intersects_location, multiple_hits=False): the ray walks the
// cubes of the grid in order (Amanatides-Woo), testing the triangles each cube
// emitted (packed_info of the counts); the triangles of a cube lie inside it,
// so the nearest hit within the first cube that has one is the first hit.
// Moller-Trumbore with a small tolerance on the barycentrics so that rays
// through a shared triangle edge are not lost.
__global__ void __launch_bounds__(256)
mesh_first_hit_kernel(const float* __restrict__ ro, const float* __restrict__ rd,
                      const float* __restrict__ t_near, const float* __restrict__ t_far,
                      int32_t n_rays, MtGrid g, const int32_t* __restrict__ packed,
                      const float* __restrict__ tris, float* __restrict__ t_hit,
                      int32_t* __restrict__ tri_hit) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rays) return;
  float o[3], d[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    o[c] = ro[3 * r + c];
    d[c] = rd[3 * r + c];
  }
  const int nc[3] = {g.rx - 1, g.ry - 1, g.rz - 1};
  // clip to the grid box
  float t0 = t_near ? t_near[r] : 0.f, t1 = t_far ? t_far[r] : 3.4e38f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float lo = g.mn[c], hi = g.mn[c] + (float)nc[c] * g.step[c];
    if (d[c] != 0.f) {
      float ta = (lo - o[c]) / d[c], tb = (hi - o[c]) / d[c];
      if (ta > tb) { const float x = ta; ta = tb; tb = x; }
      t0 = fmaxf(t0, ta);
      t1 = fminf(t1, tb);
    } else if (o[c] < lo || o[c] > hi) {
      t1 = -1.f;
    }
  }
  float best = 3.4e38f;
  int best_tri = -1;
  if (t0 <= t1) {
    int cell[3], stp[3];
    float tmax[3], tdel[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float p = (o[c] + d[c] * t0 - g.mn[c]) / g.step[c];
      int ci = (int)floorf(p);
      ci = ci < 0 ? 0 : (ci > nc[c] - 1 ? nc[c] - 1 : ci);
      cell[c] = ci;
      if (d[c] > 0.f) {
        stp[c] = 1;
        tmax[c] = (g.mn[c] + (float)(ci + 1) * g.step[c] - o[c]) / d[c];
        tdel[c] = g.step[c] / d[c];
      } else if (d[c] < 0.f) {
        stp[c] = -1;
        tmax[c] = (g.mn[c] + (float)ci * g.step[c] - o[c]) / d[c];
        tdel[c] = -g.step[c] / d[c];
      } else {
        stp[c] = 0;
        tmax[c] = 3.4e38f;
        tdel[c] = 3.4e38f;
      }
    }
    const int max_steps = nc[0] + nc[1] + nc[2] + 3;
    for (int it = 0; it < max_steps; ++it) {
      const int64_t cube = ((int64_t)cell[0] * nc[1] + cell[1]) * nc[2] + cell[2];
      const int base = packed[2 * cube], cnt = packed[2 * cube + 1];
      for (int q = 0; q < cnt; ++q) {
        const float* tr = tris + 9 * ((int64_t)base + q);
        const float e1[3] = {tr[3] - tr[0], tr[4] - tr[1], tr[5] - tr[2]};
        const float e2[3] = {tr[6] - tr[0], tr[7] - tr[1], tr[8] - tr[2]};
        const float pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2],
                             d[0] * e2[1] - d[1] * e2[0]};
        const float det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
        if (fabsf(det) < 1e-20f) continue;
        const float inv = 1.f / det;
        const float tv[3] = {o[0] - tr[0], o[1] - tr[1], o[2] - tr[2]};
        const float u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
        if (u < -1e-5f || u > 1.f + 1e-5f) continue;
        const float qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2],
                             tv[0] * e1[1] - tv[1] * e1[0]};
        const float vv = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
        if (vv < -1e-5f || u + vv > 1.f + 1e-5f) continue;
        const float th = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
        if (th >= t0 && th <= t1 && th < best) {
          best = th;
          best_tri = base + q;
        }
      }
      if (best_tri >= 0) break;
      // next cube along the smallest tmax
      int ax = 0;
      if (tmax[1] < tmax[ax]) ax = 1;
      if (tmax[2] < tmax[ax]) ax = 2;
      if (tmax[ax] > t1) break;
      cell[ax] += stp[ax];
      if (cell[ax] < 0 || cell[ax] >= nc[ax]) break;
      tmax[ax] += tdel[ax];
    }
  }
  t_hit[r] = best_tri >= 0 ? best : __int_as_float(0x7f800000);
  tri_hit[r] = best_tri;
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

int sn_mc_count(const float* values, int32_t rx, int32_t ry, int32_t rz, float iso,
                int32_t* counts, void* stream) {
  clear_error();
  const float z[3] = {0.f, 0.f, 0.f}, o[3] = {1.f, 1.f, 1.f};
  MtGrid g;
  int rc = mt_grid("sn_mc_count", rx, ry, rz, iso, z, o, &g);
  if (rc) return rc;
  SN_REQUIRE(values && counts, "sn_mc_count: null pointer");
  const int64_t n = (int64_t)(rx - 1) * (ry - 1) * (rz - 1);
  mc_count_kernel<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(values, g, counts);
  return check_launch("sn_mc_count");
}

int sn_mc_emit(const float* values, int32_t rx, int32_t ry, int32_t rz, float iso,
               const float* bound_min, const float* bound_max, const int32_t* packed_info,
               float* triangles, int64_t* vertex_ids, void* stream) {
  clear_error();
  SN_REQUIRE(bound_min && bound_max, "sn_mc_emit: null bounds");
  MtGrid g;
  int rc = mt_grid("sn_mc_emit", rx, ry, rz, iso, bound_min, bound_max, &g);
  if (rc) return rc;
  SN_REQUIRE(values && packed_info && triangles, "sn_mc_emit: null pointer");
  const int64_t n = (int64_t)(rx - 1) * (ry - 1) * (rz - 1);
  mc_emit_kernel<<<grid_for(n, 256), 256, 0, as_stream(stream)>>>(values, g, packed_info,
                                                                   triangles, vertex_ids);
  return check_launch("sn_mc_emit");
}

int sn_mesh_first_hit(const float* rays_o, const float* rays_d, const float* t_near,
                      const float* t_far, int32_t n_rays, int32_t rx, int32_t ry, int32_t rz,
                      const float* bound_min, const float* bound_max,
                      const int32_t* packed_info, const float* triangles, float* t_hit,
                      int32_t* tri_hit, void* stream) {
  clear_error();
  SN_REQUIRE(n_rays >= 0, "sn_mesh_first_hit: n_rays < 0");
  SN_REQUIRE(bound_min && bound_max, "sn_mesh_first_hit: null bounds");
  MtGrid g;
  int rc = mt_grid("sn_mesh_first_hit", rx, ry, rz, 0.f, bound_min, bound_max, &g);
  if (rc) return rc;
  if (n_rays == 0) return SN_OK;
  SN_REQUIRE(rays_o && rays_d && packed_info && triangles && t_hit && tri_hit,
             "sn_mesh_first_hit: null pointer");
  mesh_first_hit_kernel<<<(unsigned)ceil_div(n_rays, 256), 256, 0, as_stream(stream)>>>(
      rays_o, rays_d, t_near, t_far, n_rays, g, packed_info, triangles, t_hit, tri_hit);
  return check_launch("sn_mesh_first_hit");
}

}  // extern "C"
