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
//
// This file is built with -fno-hip-fp32-correctly-rounded-divide-sqrt: its fp32
// divisions and square roots are the hardware reciprocal and square root (a few
// ulp), not the correctly rounded sequences, which were a third of the sample-
// parallel kernel's VALU instructions.
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

// ---------------------------------------------------------------------------
// Sample-parallel form for small patches (Ph x Pw <= 9): one wavefront per
// patch, lane j = sample j of the patch (in chunks of 64), every ray of the
// patch in the lane's registers, so the finite differences across the patch
// are register arithmetic.  Along the samples, the transmittance is an
// exclusive prefix product and the weight backward's recurrence
//   Q_j = gw_{j+1} alpha_{j+1} + (1 - alpha_{j+1}) Q_{j+1}
// a suffix scan of affine maps Q -> b + a Q, both as wavefront scans; chunks
// carry T forwards and Q backwards.  The scans and sums are DPP lane moves
// (VALU, no LDS round trip), the backward walks a chunk's samples in reverse
// lane order so that its suffix scan is a prefix scan.  Per-lane partial sums
// are reduced once per patch.  Same quantities and outputs as
// render_step_kernel, summed in another order.
// ---------------------------------------------------------------------------
// DPP controls (gfx9): row_shr:n = 0x110 + n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143, wave_shr:1 = 0x138, wave_shl:1 = 0x130.  A lane whose
// source lies outside the wavefront, or whose row is masked off, reads `old`.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ float dpp_or(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), kCtrl,
                                                    kRowMask, 0xf, false));
}

// Inclusive scans over the 64 lanes: Kogge-Stone within rows of 16 lanes,
// then the last lane of row 0 / rows 0-1 broadcast into the rows above.
__device__ __forceinline__ float wave_incl_sum(float v) {
  v += dpp_or<0x111>(0.f, v);
  v += dpp_or<0x112>(0.f, v);
  v += dpp_or<0x114>(0.f, v);
  v += dpp_or<0x118>(0.f, v);
  v += dpp_or<0x142, 0xa>(0.f, v);
  v += dpp_or<0x143, 0xc>(0.f, v);
  return v;
}

__device__ __forceinline__ float wave_incl_prod(float v) {
  v *= dpp_or<0x111>(1.f, v);
  v *= dpp_or<0x112>(1.f, v);
  v *= dpp_or<0x114>(1.f, v);
  v *= dpp_or<0x118>(1.f, v);
  v *= dpp_or<0x142, 0xa>(1.f, v);
  v *= dpp_or<0x143, 0xc>(1.f, v);
  return v;
}

__device__ __forceinline__ float lane63(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Sum over the wavefront, the same value in every lane.
__device__ __forceinline__ float wave_sum(float v) { return lane63(wave_incl_sum(v)); }

// Affine maps x -> b + a x: lane l ends with M_l o M_(l-1) o ... o M_0.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ void affine_step(float& a, float& b) {
  const float as = dpp_or<kCtrl, kRowMask>(1.f, a), bs = dpp_or<kCtrl, kRowMask>(0.f, b);
  b = b + a * bs;
  a = a * as;
}

__device__ __forceinline__ void wave_incl_affine(float& a, float& b) {
  affine_step<0x111>(a, b);
  affine_step<0x112>(a, b);
  affine_step<0x114>(a, b);
  affine_step<0x118>(a, b);
  affine_step<0x142, 0xa>(a, b);
  affine_step<0x143, 0xc>(a, b);
}

template <int Ph, int Pw, bool AD>
__global__ void __launch_bounds__(64)
render_patch_kernel(RenderArgs a) {
  constexpr int P = Ph * Pw;
  const int p = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const float inv_s = *a.inv_s;
  const float mask_sum = *a.mask_sum;
  const int64_t n_s = *a.n_samples;
  const float n_eik = (float)n_s * (float)P;
  const float n_mask = (float)a.n_patches * (float)P;
  const float* sdf_e = a.sdf_e ? a.sdf_e : a.sdf_s + n_s * P;
  float* g_sdf_e = a.sdf_e ? a.g_sdf_e : a.g_sdf_s + n_s * P;
  const int base = a.packed[2 * p], cnt = a.packed[2 * p + 1];
  constexpr int kc = (Ph / 2) * Pw + Pw / 2;

  // per-ray constants of the patch
  float o[P][3], d[P][3], dkn[P];
  const float* nn = a.plane_n + 3 * p;
#pragma unroll
  for (int k = 0; k < P; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      o[k][c] = a.rays_o[((int64_t)p * P + k) * 3 + c];
      d[k][c] = a.rays_d[((int64_t)p * P + k) * 3 + c];
    }
    dkn[k] = (d[k][0] * nn[0] + d[k][1] * nn[1]) + d[k][2] * nn[2];
  }
  const float dcn = (d[kc][0] * nn[0] + d[kc][1] * nn[1]) + d[kc][2] * nn[2];

  // per-sample forward quantities of every ray (lane = sample j)
  struct Q {
    float fs[P], fe[P], c[P], e[P], r[P], alpha[P], g[P][3];
    float gp_coef[P][3];  // dfd: 1/dtd, 1/denx, 1/deny of each ray
  };
  auto eval = [&](int j, bool act, Q& q) {
    const int64_t s = (int64_t)base + j;
    int di = -1;
    float t0c = 0.f, t1c = 1.f;
    if (act) {
      di = a.diff_idx[s];
      t0c = a.t0[s];
      t1c = a.t1[s];
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
      float fs = 0.f, fe = 0.f;
      if (act) {
        fs = a.sdf_s[s * P + k];
        fe = di >= 0 ? sdf_e[(int64_t)di * P + k] : a.sdf_s[(s + 1) * P + k];
      }
      q.fs[k] = fs;
      q.fe[k] = fe;
      q.c[k] = sigmoidf_(fs * inv_s);
      q.e[k] = sigmoidf_(fe * inv_s);
      q.r[k] = ((q.c[k] - q.e[k]) + 1e-5f) / (q.c[k] + 1e-5f);
      q.alpha[k] = act ? fminf(fmaxf(q.r[k], 0.f), 1.f) : 0.f;
    }
    if (AD) {
#pragma unroll
      for (int k = 0; k < P; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) q.g[k][c] = act ? a.grad_s[(s * P + k) * 3 + c] : 0.f;
    } else {
      float ts[P], te[P], x[P][3];
#pragma unroll
      for (int k = 0; k < P; ++k) {
        ts[k] = (t0c * dcn) / dkn[k];
        te[k] = (t1c * dcn) / dkn[k];
#pragma unroll
        for (int c = 0; c < 3; ++c) x[k][c] = o[k][c] + d[k][c] * ts[k];
      }
      auto dist = [&](int k1, int k2) {
        const float dx = x[k2][0] - x[k1][0], dy = x[k2][1] - x[k1][1], dz = x[k2][2] - x[k1][2];
        return sqrtf(dx * dx + dy * dy + dz * dz);
      };
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const int row = k / Pw, col = k % Pw;
        const float dtd = te[k] - ts[k];
        const float df_dt = (q.fe[k] - q.fs[k]) / dtd;
        float denx, dfx, deny, dfy;
        if (col == 0) { denx = dist(k, k + 1); dfx = (q.fs[k + 1] - q.fs[k]) / denx; }
        else if (col == Pw - 1) { denx = dist(k - 1, k); dfx = (q.fs[k] - q.fs[k - 1]) / denx; }
        else { denx = dist(k - 1, k) + dist(k, k + 1); dfx = (q.fs[k + 1] - q.fs[k - 1]) / denx; }
        if (row == 0) { deny = dist(k, k + Pw); dfy = (q.fs[k + Pw] - q.fs[k]) / deny; }
        else if (row == Ph - 1) { deny = dist(k - Pw, k); dfy = (q.fs[k] - q.fs[k - Pw]) / deny; }
        else { deny = dist(k - Pw, k) + dist(k, k + Pw); dfy = (q.fs[k + Pw] - q.fs[k - Pw]) / deny; }
        q.gp_coef[k][0] = dtd;
        q.gp_coef[k][1] = denx;
        q.gp_coef[k][2] = deny;
        const float* vi = a.vinv + ((int64_t)p * P + k) * 9;
#pragma unroll
        for (int c = 0; c < 3; ++c)
          q.g[k][c] = act ? (vi[3 * c + 0] * df_dt + vi[3 * c + 1] * dfx + vi[3 * c + 2] * dfy) : 0.f;
      }
    }
  };

  // ---------------- forward: T by prefix products, sums over samples --------
  float Tin[P], ws[P], cn[P][3], eik[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    Tin[k] = 1.f;
    ws[k] = 0.f;
    cn[k][0] = cn[k][1] = cn[k][2] = 0.f;
    eik[k] = 0.f;
  }
  for (int j0 = 0; j0 < cnt; j0 += 64) {
    const int j = j0 + lane;
    const bool act = j < cnt;
    Q q;
    eval(j, act, q);
#pragma unroll
    for (int k = 0; k < P; ++k) {
      // inclusive prefix product of (1 - alpha), then exclusive (wave_shr:1)
      const float incl = wave_incl_prod(1.f - q.alpha[k]);
      const float excl = dpp_or<0x138>(1.f, incl);
      const float T = Tin[k] * excl;
      const float w = q.alpha[k] * T;
      if (act) a.g_sdf_s[((int64_t)base + j) * P + k] = T;  // parked for the backward
      const float nrm = sqrtf(q.g[k][0] * q.g[k][0] + q.g[k][1] * q.g[k][1] + q.g[k][2] * q.g[k][2]);
      // per-lane partial sums, reduced once after the last chunk
      ws[k] += w;
#pragma unroll
      for (int c = 0; c < 3; ++c) cn[k][c] += w * q.g[k][c];
      eik[k] += act ? (nrm - 1.f) * (nrm - 1.f) : 0.f;
      Tin[k] *= lane63(incl);
    }
  }
#pragma unroll
  for (int k = 0; k < P; ++k) {
    ws[k] = wave_sum(ws[k]);
#pragma unroll
    for (int c = 0; c < 3; ++c) cn[k][c] = wave_sum(cn[k][c]);
    eik[k] = wave_sum(eik[k]);
  }
  // ---------------- losses of every ray and their gradients -----------------
  float g_cn[P][3], g_ws[P];
  float l_n = 0.f, l_m = 0.f, l_e = 0.f;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const float m = a.mask[(int64_t)p * P + k];
    float ln = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float nt = a.n_true[((int64_t)p * P + k) * 3 + c];
      const float err = (cn[k][c] - nt) * m;
      if (a.l1) {
        ln += fabsf(err);
        g_cn[k][c] = a.normal_w * (err > 0.f ? 1.f : (err < 0.f ? -1.f : 0.f)) * m / mask_sum;
      } else {
        ln += err * err;
        g_cn[k][c] = a.normal_w * 2.f * err * m / mask_sum;
      }
    }
    l_n += ln / mask_sum;
    const float lo = 1e-5f, hi = 1.f - 1e-5f;
    const float xw = fminf(fmaxf(ws[k], lo), hi);
    l_m += -(m * fmaxf(logf(xw), -100.f) + (1.f - m) * fmaxf(logf(1.f - xw), -100.f)) / n_mask;
    const bool pass = ws[k] >= lo && ws[k] <= hi;
    g_ws[k] = pass ? a.mask_w * (xw - m) / fmaxf((1.f - xw) * xw, 1e-12f) / n_mask : 0.f;
    l_e += eik[k] / n_eik;
    if (lane == 0) {
      if (a.weight_sum) a.weight_sum[(int64_t)p * P + k] = ws[k];
      if (a.comp_normal)
#pragma unroll
        for (int c = 0; c < 3; ++c) a.comp_normal[((int64_t)p * P + k) * 3 + c] = cn[k][c];
    }
  }
  // ---------------- backward, chunks in reverse ------------------------------
  float Qin[P];
#pragma unroll
  for (int k = 0; k < P; ++k) Qin[k] = 0.f;
  float g_inv = 0.f;
  const int n_chunks = (cnt + 63) / 64;
  for (int ch = n_chunks - 1; ch >= 0; --ch) {
    // reverse lane order: lane 0 holds the chunk's last sample
    const int j = ch * 64 + 63 - lane;
    const bool act = j < cnt;
    const int64_t s = (int64_t)base + j;
    Q q;
    eval(j, act, q);
    const int di = act ? a.diff_idx[s] : 0;
    float G[P];  // dL/dfs of every ray of this sample (own terms + neighbours')
    float GE[P]; // dL/dfe
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const float Tj = act ? a.g_sdf_s[s * P + k] : 0.f;
      const float w = q.alpha[k] * Tj;
      const float nrm = sqrtf(q.g[k][0] * q.g[k][0] + q.g[k][1] * q.g[k][1] + q.g[k][2] * q.g[k][2]);
      const float ek = nrm > 0.f ? a.eik_w * 2.f * (nrm - 1.f) / n_eik / nrm : 0.f;
      float gg[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) gg[c] = act ? (w * g_cn[k][c] + ek * q.g[k][c]) : 0.f;
      const float gw = act ? (g_ws[k] + g_cn[k][0] * q.g[k][0] + g_cn[k][1] * q.g[k][1] +
                              g_cn[k][2] * q.g[k][2]) : 0.f;
      // suffix scan of the maps M_j(Q) = gw_j alpha_j + (1 - alpha_j) Q: in
      // reversed lane order a prefix scan, lane l = S_j (chunk end .. j)
      float ma = act ? 1.f - q.alpha[k] : 1.f, mb = act ? gw * q.alpha[k] : 0.f;
      wave_incl_affine(ma, mb);
      // Q_j = S_{j+1}(Qin): the previous lane's map (lane 0: the identity)
      const float sa = dpp_or<0x138>(1.f, ma), sb = dpp_or<0x138>(0.f, mb);
      const float Qj = sb + sa * Qin[k];
      const float Qnext = lane63(mb + ma * Qin[k]);  // S_{chunk start}(Qin): carried back
      const float g_alpha = act ? Tj * (gw - Qj) : 0.f;
      const float gr = (q.r[k] >= 0.f && q.r[k] <= 1.f) ? g_alpha : 0.f;
      const float den = q.c[k] + 1e-5f;
      const float g_c = gr * q.e[k] / (den * den);
      const float g_e = -gr / den;
      G[k] = g_c * inv_s * q.c[k] * (1.f - q.c[k]);
      GE[k] = g_e * inv_s * q.e[k] * (1.f - q.e[k]);
      g_inv += act ? (g_c * q.fs[k] * q.c[k] * (1.f - q.c[k]) + g_e * q.fe[k] * q.e[k] * (1.f - q.e[k])) : 0.f;
      Qin[k] = Qnext;
      if (AD) {
        if (act)
#pragma unroll
          for (int c = 0; c < 3; ++c) a.g_grad_s[(s * P + k) * 3 + c] = gg[c];
      } else {
        const float* vi = a.vinv + ((int64_t)p * P + k) * 9;
        // grad = V^-1 proj  =>  dL/dproj = V^-T dL/dgrad; kept in gg
        const float gp0 = vi[0] * gg[0] + vi[3] * gg[1] + vi[6] * gg[2];
        const float gp1 = vi[1] * gg[0] + vi[4] * gg[1] + vi[7] * gg[2];
        const float gp2 = vi[2] * gg[0] + vi[5] * gg[1] + vi[8] * gg[2];
        gg[0] = gp0 / q.gp_coef[k][0];
        gg[1] = act ? gp1 / q.gp_coef[k][1] : 0.f;
        gg[2] = act ? gp2 / q.gp_coef[k][2] : 0.f;
        q.gp_coef[k][0] = gg[0];
        q.gp_coef[k][1] = gg[1];
        q.gp_coef[k][2] = gg[2];
      }
    }
    if (!AD) {
      // df_dt = (fe - fs) / dtd; finite differences across the patch
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const int row = k / Pw, col = k % Pw;
        GE[k] += q.gp_coef[k][0];
        G[k] -= q.gp_coef[k][0];
        const float ax = q.gp_coef[k][1], ay = q.gp_coef[k][2];
        if (col == 0) { G[k + 1] += ax; G[k] -= ax; }
        else if (col == Pw - 1) { G[k] += ax; G[k - 1] -= ax; }
        else { G[k + 1] += ax; G[k - 1] -= ax; }
        if (row == 0) { G[k + Pw] += ay; G[k] -= ay; }
        else if (row == Ph - 1) { G[k] += ay; G[k - Pw] -= ay; }
        else { G[k + Pw] += ay; G[k - Pw] -= ay; }
      }
    }
    // ends: own SDF row, or the next sample's start (added to that start);
    // sample j - 1 sits in lane + 1 (wave_shl:1)
    const bool to_next = act && di < 0;
    const float prev_to_next = dpp_or<0x130>(0.f, to_next ? 1.f : 0.f);
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const float from_prev = dpp_or<0x130>(0.f, GE[k]);
      float gs = G[k];
      if (lane < 63 && prev_to_next != 0.f) gs += from_prev;
      if (act) {
        // the chunk's first sample may also receive from the previous chunk,
        // whose lane 63 adds its share later (read-modify-write below)
        a.g_sdf_s[s * P + k] = gs;
        if (di >= 0) g_sdf_e[(int64_t)di * P + k] = GE[k];
      }
    }
    if (lane == 0 && to_next && j + 1 < cnt) {
      // sample j + 1 is the next chunk's first sample, already written
#pragma unroll
      for (int k = 0; k < P; ++k) a.g_sdf_s[(s + 1) * P + k] += GE[k];
    }
  }
  // per-patch partials: normal, mask, eikonal, dL/dinv_s
  const float v3 = wave_sum(g_inv);
  if (lane == 0) {
    float* o4 = a.partials + 4 * (int64_t)p;
    o4[0] = a.normal_w * l_n;
    o4[1] = a.mask_w * l_m;
    o4[2] = a.eik_w * l_e;
    o4[3] = v3;
  }
}

// Fixed-order sum of n rows of 4 partials: 256 lanes stride the rows, then the
// lanes' sums are added in lane order.
__global__ void __launch_bounds__(256)
reduce4_wide_kernel(const float* __restrict__ part, int n, float* __restrict__ out) {
  __shared__ float sh[256][4];
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  for (int i = threadIdx.x; i < n; i += 256)
#pragma unroll
    for (int c = 0; c < 4; ++c) s[c] += part[4 * (int64_t)i + c];
#pragma unroll
  for (int c = 0; c < 4; ++c) sh[threadIdx.x][c] = s[c];
  __syncthreads();
  if (threadIdx.x < 4) {
    float t = 0.f;
    for (int i = 0; i < 256; ++i) t += sh[i][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

static int g_render_impl = 1;  // 1: sample-parallel kernel where the patch shape allows

}  // namespace sn

using namespace sn;

extern "C" {

int sn_render_step_waves(int32_t n_patches, int32_t patch_size) {
  if (patch_size <= 0 || patch_size > 64 || n_patches < 0) return -1;
  return n_patches;  // one partial row per patch at most (sample-parallel kernel)
}

void sn_render_step_set_impl(int32_t impl) { g_render_impl = impl; }

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
  hipStream_t s = as_stream(stream);
  const bool ad = grad_starts != nullptr;
  if (g_render_impl == 1 && patch_h == 3 && patch_w == 3) {
    if (ad) render_patch_kernel<3, 3, true><<<n_patches, 64, 0, s>>>(a);
    else render_patch_kernel<3, 3, false><<<n_patches, 64, 0, s>>>(a);
    reduce4_wide_kernel<<<1, 256, 0, s>>>(partials, n_patches, losses);
  } else if (g_render_impl == 1 && patch_h == 2 && patch_w == 4) {
    if (ad) render_patch_kernel<2, 4, true><<<n_patches, 64, 0, s>>>(a);
    else render_patch_kernel<2, 4, false><<<n_patches, 64, 0, s>>>(a);
    reduce4_wide_kernel<<<1, 256, 0, s>>>(partials, n_patches, losses);
  } else {
    const int waves = (int)ceil_div(n_patches, 64 / P);
    render_step_kernel<<<waves, 64, 0, s>>>(a);
    reduce4_wide_kernel<<<1, 256, 0, s>>>(partials, waves, losses);
  }
  return check_launch("sn_render_step");
}

}  // extern "C"
