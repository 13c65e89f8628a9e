// One fused forward + backward of SuperNormal's patch rendering and losses
// (models/renderer.py:142-276 and exp_runner.py:169-203) for every patch:
//   NeuS alpha from the SDF at sample starts / ends        (renderer.py:164-179)
//   patch render weights w = alpha T                        (vol_rendering.py:533-576)
//   per-sample normal: directional finite differences 'dfd' (renderer.py:187-223)
//                      or the analytic gradient 'ad'         (renderer.py:225-226)
//   weight_sum, comp_normal accumulation                    (renderer.py:263-267)
//   normal (l2 or l1), eikonal and mask BCE losses          (exp_runner.py:191-203)
// and the exact backward of the total loss to dL/dsdf (starts and the ends of
// samples whose end is not the next sample's start), dL/dgrad ('ad') and
// dL/dinv_s.
//
// Lane mapping: a wavefront holds floor(64 / P) patches; lane = slot * P + k
// with k = row * Pw + col the ray of the patch.  Patches are independent
// (each term of the loss involves one patch; the normalisers mask_sum and
// the eikonal count are inputs), so no cross-wave communication is needed.
// Neighbouring rays' values (finite differences along the patch rows and
// columns) move by lane shuffles.  The sample loop is sequential per ray in
// the reference's order, forward for T and reverse for the weight gradient
// (division-free recurrence, finite at alpha = 1).
//
// Scratch: during the forward pass T[s,k] is parked in g_sdf_s[s,k]; the
// reverse pass reads it back before overwriting that slot.
//
// Per-wave loss partials (normal, mask, eikonal, dL/dinv_s) go to `partials`
// [n_waves, 4] and are summed in a fixed order by the caller's reduction.
#include "sn_common.h"

namespace sn {

struct RenderArgs {
  int32_t n_patches, Ph, Pw, P;
  const float* rays_o;     // [B, P, 3]
  const float* rays_d;     // [B, P, 3]
  const float* plane_n;    // [B, 3]
  const float* vinv;       // [B, P, 9] row-major 3x3
  const float* n_true;     // [B, P, 3]
  const float* mask;       // [B, P]
  const int32_t* packed;   // [B, 2] samples of each patch (base, count)
  const float* t0;         // [S] centre-ray sample starts
  const float* t1;         // [S]
  const int32_t* diff_idx; // [S] row of the sample's end in sdf_e, or -1 (end = next start)
  const float* sdf_s;      // [S, P]
  const float* sdf_e;      // [Nd, P]
  const float* grad_s;     // [S, P, 3] analytic gradient (ad) or null (dfd)
  const float* inv_s;      // device scalar
  const float* mask_sum;   // device scalar (sum of mask + 1e-5)
  const int32_t* n_samples;// device scalar S (eikonal mean over S*P)
  float normal_w, eik_w, mask_w;
  int32_t l1;
  // outputs
  float* g_sdf_s;          // [S, P]
  float* g_sdf_e;          // [Nd, P]
  float* g_grad_s;         // [S, P, 3] (ad)
  float* partials;         // [n_waves, 4]
  float* weight_sum;       // [B, P] (nullable)
  float* comp_normal;      // [B, P, 3] (nullable)
};

__device__ __forceinline__ float sigmoidf_(float x) { return 1.f / (1.f + expf(-x)); }

// per (sample, ray) quantities of the forward model
struct SampleF {
  float fs, fe, ts, te, c, e, r, alpha, g[3], proj[3], dtd, denx, deny;
};

__global__ void __launch_bounds__(64)
render_step_kernel(RenderArgs a) {
  const int lane = threadIdx.x & 63;
  const int P = a.P, Pw = a.Pw, Ph = a.Ph;
  const int per_wave = 64 / P;
  const int slot = lane / P;
  const int k = lane - slot * P;
  const int p = blockIdx.x * per_wave + slot;
  const bool lane_ok = slot < per_wave && p < a.n_patches;
  const int row = k / Pw, col = k - row * Pw;
  const int kc = (Ph / 2) * Pw + Pw / 2;
  const float inv_s = *a.inv_s;
  const float mask_sum = *a.mask_sum;
  const int64_t n_s = *a.n_samples;
  const float n_eik = (float)n_s * (float)P;
  // ends stored right after the S*P starts when no separate buffer is given
  const float* sdf_e = a.sdf_e ? a.sdf_e : a.sdf_s + n_s * P;
  float* g_sdf_e = a.sdf_e ? a.g_sdf_e : a.g_sdf_s + n_s * P;
  const float n_mask = (float)a.n_patches * (float)P;
  const bool ad = a.grad_s != nullptr;

  int base = 0, cnt = 0;
  float o[3] = {0, 0, 0}, d[3] = {0, 0, 1}, vi[9] = {0}, nt[3] = {0, 0, 0}, m = 0.f;
  float dcn = 1.f, dkn = 1.f;
  if (lane_ok) {
    base = a.packed[2 * p];
    cnt = a.packed[2 * p + 1];
    for (int c = 0; c < 3; ++c) {
      o[c] = a.rays_o[(p * P + k) * 3 + c];
      d[c] = a.rays_d[(p * P + k) * 3 + c];
      nt[c] = a.n_true[(p * P + k) * 3 + c];
    }
    for (int c = 0; c < 9; ++c) vi[c] = a.vinv[(p * P + k) * 9 + c];
    m = a.mask[p * P + k];
    const float* n = a.plane_n + 3 * p;
    const float* dc = a.rays_d + (p * P + kc) * 3;
    // t_k = (t_c * (d_c . n)) / (d_k . n)   (renderer.py:146-149)
    dcn = (dc[0] * n[0] + dc[1] * n[1]) + dc[2] * n[2];
    dkn = (d[0] * n[0] + d[1] * n[1]) + d[2] * n[2];
  }
  // loop bound: the longest patch of this wavefront
  int max_cnt = cnt;
  for (int off = 32; off >= 1; off >>= 1) max_cnt = max(max_cnt, __shfl_xor(max_cnt, off, 64));

  // shuffle sources inside this lane's patch (self when absent; lanes past
  // the last whole patch slot only ever read themselves)
  const bool in_slot = slot < per_wave;
  const int L = slot * P;
  const int lx_l = (in_slot && col > 0) ? L + k - 1 : lane;
  const int lx_r = (in_slot && col < Pw - 1) ? L + k + 1 : lane;
  const int ly_u = (in_slot && row > 0) ? L + k - Pw : lane;
  const int ly_d = (in_slot && row < Ph - 1) ? L + k + Pw : lane;
  const int lx_0 = in_slot ? L + row * Pw : lane, lx_w = in_slot ? L + row * Pw + Pw - 1 : lane;
  const int ly_0 = in_slot ? L + col : lane, ly_h = in_slot ? L + (Ph - 1) * Pw + col : lane;

  // evaluates the forward quantities of sample j for this lane (all lanes call it)
  auto eval = [&](int j, bool act) {
    SampleF f;
    const int s = base + j;
    f.fs = act ? a.sdf_s[(int64_t)s * P + k] : 0.f;
    int di = act ? a.diff_idx[s] : 0;
    if (act) f.fe = di >= 0 ? sdf_e[(int64_t)di * P + k] : a.sdf_s[(int64_t)(s + 1) * P + k];
    else f.fe = 0.f;
    const float t0c = act ? a.t0[s] : 0.f, t1c = act ? a.t1[s] : 1.f;
    f.ts = (t0c * dcn) / dkn;
    f.te = (t1c * dcn) / dkn;
    f.c = sigmoidf_(f.fs * inv_s);
    f.e = sigmoidf_(f.fe * inv_s);
    f.r = ((f.c - f.e) + 1e-5f) / (f.c + 1e-5f);
    f.alpha = fminf(fmaxf(f.r, 0.f), 1.f);
    if (ad) {
      for (int c = 0; c < 3; ++c) f.g[c] = act ? a.grad_s[((int64_t)s * P + k) * 3 + c] : 0.f;
      f.proj[0] = f.proj[1] = f.proj[2] = 0.f;
      f.dtd = f.denx = f.deny = 1.f;
    } else {
      // positions of the starts on this ray and its right / lower neighbours
      float x[3];
      for (int c = 0; c < 3; ++c) x[c] = o[c] + d[c] * f.ts;
      float xr[3], xd[3];
      for (int c = 0; c < 3; ++c) {
        xr[c] = __shfl(x[c], lx_r, 64);
        xd[c] = __shfl(x[c], ly_d, 64);
      }
      const float dist_r = sqrtf((xr[0] - x[0]) * (xr[0] - x[0]) + (xr[1] - x[1]) * (xr[1] - x[1]) +
                                 (xr[2] - x[2]) * (xr[2] - x[2]));
      const float dist_d = sqrtf((xd[0] - x[0]) * (xd[0] - x[0]) + (xd[1] - x[1]) * (xd[1] - x[1]) +
                                 (xd[2] - x[2]) * (xd[2] - x[2]));
      const float dist_l = __shfl(dist_r, lx_l, 64);  // distance to the left neighbour
      const float dist_u = __shfl(dist_d, ly_u, 64);
      const float f_l = __shfl(f.fs, lx_l, 64), f_r = __shfl(f.fs, lx_r, 64);
      const float f_u = __shfl(f.fs, ly_u, 64), f_d = __shfl(f.fs, ly_d, 64);
      f.dtd = f.te - f.ts;
      const float df_dt = (f.fe - f.fs) / f.dtd;
      float dfx, dfy;
      if (col == 0) { f.denx = dist_r; dfx = (f_r - f.fs) / f.denx; }
      else if (col == Pw - 1) { f.denx = dist_l; dfx = (f.fs - f_l) / f.denx; }
      else { f.denx = dist_l + dist_r; dfx = (f_r - f_l) / f.denx; }
      if (row == 0) { f.deny = dist_d; dfy = (f_d - f.fs) / f.deny; }
      else if (row == Ph - 1) { f.deny = dist_u; dfy = (f.fs - f_u) / f.deny; }
      else { f.deny = dist_u + dist_d; dfy = (f_d - f_u) / f.deny; }
      f.proj[0] = df_dt;
      f.proj[1] = dfx;
      f.proj[2] = dfy;
      for (int c = 0; c < 3; ++c)
        f.g[c] = vi[3 * c + 0] * df_dt + vi[3 * c + 1] * dfx + vi[3 * c + 2] * dfy;
    }
    return f;
  };

  // ---------------- forward ----------------
  float T = 1.f, ws = 0.f, cn[3] = {0.f, 0.f, 0.f}, eik = 0.f;
  for (int j = 0; j < max_cnt; ++j) {
    const bool act = lane_ok && j < cnt;
    SampleF f = eval(j, act);
    if (act) {
      const float w = f.alpha * T;
      a.g_sdf_s[(int64_t)(base + j) * P + k] = T;  // parked for the reverse pass
      T *= (1.f - f.alpha);
      ws += w;
      for (int c = 0; c < 3; ++c) cn[c] += w * f.g[c];
      const float nrm = sqrtf(f.g[0] * f.g[0] + f.g[1] * f.g[1] + f.g[2] * f.g[2]);
      eik += (nrm - 1.f) * (nrm - 1.f);
    }
  }
  // losses of this ray and their gradients w.r.t. (weight_sum, comp_normal)
  float l_n = 0.f, l_m = 0.f, g_cn[3] = {0.f, 0.f, 0.f}, g_ws = 0.f;
  if (lane_ok) {
    for (int c = 0; c < 3; ++c) {
      const float err = (cn[c] - nt[c]) * m;
      if (a.l1) {
        l_n += fabsf(err);
        g_cn[c] = a.normal_w * (err > 0.f ? 1.f : (err < 0.f ? -1.f : 0.f)) * m / mask_sum;
      } else {
        l_n += err * err;
        g_cn[c] = a.normal_w * 2.f * err * m / mask_sum;
      }
    }
    l_n /= mask_sum;
    const float lo = 1e-5f, hi = 1.f - 1e-5f;
    const float xw = fminf(fmaxf(ws, lo), hi);
    l_m = -(m * fmaxf(logf(xw), -100.f) + (1.f - m) * fmaxf(logf(1.f - xw), -100.f)) / n_mask;
    const bool pass = ws >= lo && ws <= hi;
    g_ws = pass ? a.mask_w * (xw - m) / fmaxf((1.f - xw) * xw, 1e-12f) / n_mask : 0.f;
    if (a.weight_sum) a.weight_sum[p * P + k] = ws;
    if (a.comp_normal)
      for (int c = 0; c < 3; ++c) a.comp_normal[(p * P + k) * 3 + c] = cn[c];
  }
  // ---------------- reverse ----------------
  float Q = 0.f, carry = 0.f, g_inv = 0.f;
  float a_next = 0.f, gw_next = 0.f;  // alpha and dL/dw of sample j+1
  for (int j = max_cnt - 1; j >= 0; --j) {
    const bool act = lane_ok && j < cnt;
    SampleF f = eval(j, act);
    const int s = base + j;
    const float Tj = act ? a.g_sdf_s[(int64_t)s * P + k] : 0.f;
    const float w = f.alpha * Tj;
    // dL/dgrad of this sample
    const float nrm = sqrtf(f.g[0] * f.g[0] + f.g[1] * f.g[1] + f.g[2] * f.g[2]);
    float gg[3];
    const float ek = nrm > 0.f ? a.eik_w * 2.f * (nrm - 1.f) / n_eik / nrm : 0.f;
    for (int c = 0; c < 3; ++c) gg[c] = act ? (w * g_cn[c] + ek * f.g[c]) : 0.f;
    // dL/dw and the weight recurrence
    const float gw = act ? (g_ws + g_cn[0] * f.g[0] + g_cn[1] * f.g[1] + g_cn[2] * f.g[2]) : 0.f;
    Q = (j == cnt - 1) ? 0.f : (gw_next * a_next) + (1.f - a_next) * Q;
    const float g_alpha = act ? Tj * (gw - Q) : 0.f;
    a_next = f.alpha;
    gw_next = gw;
    // alpha = clip(r): r = (c - e + eps) / (c + eps)
    const float gr = (f.r >= 0.f && f.r <= 1.f) ? g_alpha : 0.f;
    const float den = f.c + 1e-5f;
    const float g_c = gr * f.e / (den * den);
    const float g_e = -gr / den;
    float g_fs = g_c * inv_s * f.c * (1.f - f.c);
    float g_fe = g_e * inv_s * f.e * (1.f - f.e);
    g_inv += act ? (g_c * f.fs * f.c * (1.f - f.c) + g_e * f.fe * f.e * (1.f - f.e)) : 0.f;
    if (ad) {
      if (act)
        for (int c = 0; c < 3; ++c) a.g_grad_s[((int64_t)s * P + k) * 3 + c] = gg[c];
    } else {
      // grad = V^-1 proj  =>  dL/dproj = V^-T dL/dgrad
      const float gp0 = vi[0] * gg[0] + vi[3] * gg[1] + vi[6] * gg[2];
      const float gp1 = vi[1] * gg[0] + vi[4] * gg[1] + vi[7] * gg[2];
      const float gp2 = vi[2] * gg[0] + vi[5] * gg[1] + vi[8] * gg[2];
      // df_dt = (fe - fs) / (te - ts)
      g_fe += gp0 / f.dtd;
      g_fs -= gp0 / f.dtd;
      // df_dx, df_dy use neighbouring rays' start values: gather the
      // coefficient each neighbour applies to this ray's f.
      const float ax = act ? gp1 / f.denx : 0.f, ay = act ? gp2 / f.deny : 0.f;
      const float ax_l = __shfl(ax, lx_l, 64), ax_r = __shfl(ax, lx_r, 64);
      const float ax_0 = __shfl(ax, lx_0, 64), ax_w = __shfl(ax, lx_w, 64);
      const float ay_u = __shfl(ay, ly_u, 64), ay_d = __shfl(ay, ly_d, 64);
      const float ay_0 = __shfl(ay, ly_0, 64), ay_h = __shfl(ay, ly_h, 64);
      float gx = 0.f, gy = 0.f;
      // central differences at col+1 (minus) and col-1 (plus)
      if (col + 1 >= 1 && col + 1 <= Pw - 2) gx -= ax_r;
      if (col - 1 >= 1 && col - 1 <= Pw - 2) gx += ax_l;
      // one-sided differences at the patch borders
      if (col == 1) gx += ax_0;
      if (col == 0) gx -= ax_0;
      if (col == Pw - 1) gx += ax_w;
      if (col == Pw - 2) gx -= ax_w;
      if (row + 1 >= 1 && row + 1 <= Ph - 2) gy -= ay_d;
      if (row - 1 >= 1 && row - 1 <= Ph - 2) gy += ay_u;
      if (row == 1) gy += ay_0;
      if (row == 0) gy -= ay_0;
      if (row == Ph - 1) gy += ay_h;
      if (row == Ph - 2) gy -= ay_h;
      g_fs += gx + gy;
    }
    if (act) {
      // the end value: its own SDF row, or the next sample's start
      const int di = a.diff_idx[s];
      if (di >= 0) {
        g_sdf_e[(int64_t)di * P + k] = g_fe;
        if (j + 1 < cnt) a.g_sdf_s[(int64_t)(s + 1) * P + k] = carry;
      } else {
        a.g_sdf_s[(int64_t)(s + 1) * P + k] = carry + g_fe;
      }
      carry = g_fs;
    }
  }
  if (lane_ok && cnt > 0) a.g_sdf_s[(int64_t)base * P + k] = carry;
  // per-wave partials: normal, mask, eikonal, dL/dinv_s
  float v0 = lane_ok ? a.normal_w * l_n : 0.f;
  float v1 = lane_ok ? a.mask_w * l_m : 0.f;
  float v2 = lane_ok ? a.eik_w * eik / n_eik : 0.f;
  float v3 = lane_ok ? g_inv : 0.f;
  for (int off = 32; off >= 1; off >>= 1) {
    v0 += __shfl_xor(v0, off, 64);
    v1 += __shfl_xor(v1, off, 64);
    v2 += __shfl_xor(v2, off, 64);
    v3 += __shfl_xor(v3, off, 64);
  }
  if (lane == 0) {
    float* o4 = a.partials + 4 * blockIdx.x;
    o4[0] = v0;
    o4[1] = v1;
    o4[2] = v2;
    o4[3] = v3;
  }
}

__global__ void reduce4_kernel(const float* __restrict__ part, int n, float* __restrict__ out) {
  const int c = threadIdx.x;
  if (c >= 4) return;
  float s = 0.f;
  for (int i = 0; i < n; ++i) s += part[4 * i + c];
  out[c] = s;
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_render_step_waves(int32_t n_patches, int32_t patch_size) {
  if (patch_size <= 0 || patch_size > 64 || n_patches < 0) return -1;
  return (int)ceil_div(n_patches, 64 / patch_size);
}

int sn_render_step(int32_t n_patches, int32_t patch_h, int32_t patch_w, const float* rays_o,
                   const float* rays_d, const float* plane_n, const float* v_inv,
                   const float* n_true, const float* mask, const int32_t* packed_info,
                   const float* t_starts, const float* t_ends, const int32_t* diff_idx,
                   const float* sdf_starts, const float* sdf_ends, const float* grad_starts,
                   const float* inv_s, const float* mask_sum, const int32_t* n_samples,
                   float normal_w, float eik_w, float mask_w, int32_t loss_l1,
                   float* g_sdf_starts, float* g_sdf_ends, float* g_grad_starts,
                   float* losses, float* partials, float* weight_sum, float* comp_normal,
                   void* stream) {
  clear_error();
  const int P = patch_h * patch_w;
  SN_REQUIRE(n_patches >= 0 && patch_h > 0 && patch_w > 0 && P <= 64,
             "sn_render_step: bad patch shape %dx%d", patch_h, patch_w);
  SN_REQUIRE(patch_h >= 2 && patch_w >= 2, "sn_render_step: patches need >= 2 rows and cols");
  if (n_patches == 0) return SN_OK;
  SN_REQUIRE(rays_o && rays_d && plane_n && n_true && mask && packed_info && t_starts &&
                 t_ends && diff_idx && sdf_starts && inv_s && mask_sum && n_samples &&
                 g_sdf_starts && losses && partials,
             "sn_render_step: null pointer");
  SN_REQUIRE(grad_starts ? (g_grad_starts != nullptr) : (v_inv != nullptr),
             "sn_render_step: 'ad' needs g_grad_starts, 'dfd' needs v_inv");
  RenderArgs a;
  a.n_patches = n_patches;
  a.Ph = patch_h;
  a.Pw = patch_w;
  a.P = P;
  a.rays_o = rays_o;
  a.rays_d = rays_d;
  a.plane_n = plane_n;
  a.vinv = v_inv;
  a.n_true = n_true;
  a.mask = mask;
  a.packed = packed_info;
  a.t0 = t_starts;
  a.t1 = t_ends;
  a.diff_idx = diff_idx;
  a.sdf_s = sdf_starts;
  a.sdf_e = sdf_ends;
  a.grad_s = grad_starts;
  a.inv_s = inv_s;
  a.mask_sum = mask_sum;
  a.n_samples = n_samples;
  a.normal_w = normal_w;
  a.eik_w = eik_w;
  a.mask_w = mask_w;
  a.l1 = loss_l1;
  a.g_sdf_s = g_sdf_starts;
  a.g_sdf_e = g_sdf_ends;
  a.g_grad_s = g_grad_starts;
  a.partials = partials;
  a.weight_sum = weight_sum;
  a.comp_normal = comp_normal;
  const int waves = sn_render_step_waves(n_patches, P);
  hipStream_t s = as_stream(stream);
  render_step_kernel<<<waves, 64, 0, s>>>(a);
  reduce4_kernel<<<1, 64, 0, s>>>(partials, waves, losses);
  return check_launch("sn_render_step");
}

}  // extern "C"
