/*
 * sn_api.h -- C ABI of libsupernormal_amd.so, the MI355X (gfx950) kernels of
 * SuperNormal's training hot path.
 *
 * Conventions (every entry point):
 *   - All tensor arguments are DEVICE pointers to caller-owned, contiguous
 *     buffers.  Nothing here allocates device memory unless the comment says
 *     so; scratch space is passed in as `workspace` with a size queried from
 *     the matching *_workspace_bytes() function.
 *   - `stream` is a hipStream_t passed as void*; all work is enqueued on it,
 *     nothing synchronises the device.
 *   - Return value: 0 on success, a negative SN_E* code on failure.  The
 *     message of the last failure on the calling host thread is returned by
 *     sn_last_error().  No C++ exception crosses this boundary.
 *   - dtypes follow the reference: packed_info int32 [n,2] (base, count),
 *     ray_indices int64, all data float32.
 *
 * Reference paths are relative to the SuperNormal repository; NA/ is
 * third_parties/nerfacc-0.3.5/nerfacc-0.3.5/ and csrc/ is NA/nerfacc/cuda/csrc/.
 */
#ifndef SN_API_H
#define SN_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SN_OK 0
#define SN_EINVAL (-1)  /* bad argument (shape, size, null pointer)       */
#define SN_EHIP (-2)    /* a HIP runtime call failed                       */
#define SN_ESPACE (-3)  /* workspace or capacity too small                 */

/* Message of the last failure on this host thread ("" if none). */
const char* sn_last_error(void);
/* ABI version of this library (increments on any signature change). */
int sn_abi_version(void);

/* ------------------------------------------------------------------------ */
/* Packing: replaces nerfacc.pack_info (NA/nerfacc/pack.py:46-77).           */
/* counts[r] = #{i : ray_indices[i] == r}; packed_info[r] = (exclusive       */
/* cumsum of counts, count).  Indices outside [0, n_rays) are an error in the */
/* reference (scatter_add index error); here they are ignored and counted in */
/* the optional `n_bad` output (device int32, nullable).                     */
/* ------------------------------------------------------------------------ */
size_t sn_pack_info_workspace_bytes(int32_t n_rays);
int sn_pack_info(const int64_t* ray_indices, int64_t n_samples, int32_t n_rays,
                 int32_t* packed_info, int32_t* n_bad, void* workspace,
                 size_t workspace_bytes, void* stream);

/* pack_info that allocates nothing and takes no workspace: no counterpart   */
/* in the reference; for callers that cannot call hipMalloc (code under HIP  */
/* graph capture, tight serving loops, harnesses that forbid allocation).    */
/* Same result as sn_pack_info.  Its scratch is a fixed device array, one    */
/* per GPU for the life of the process, of sn_pack_info_workspace_bytes(     */
/* sn_pack_info_noalloc_max_rays()) bytes; n_rays above that limit (2^22) is */
/* SN_EINVAL.  Calls on one GPU share the array: overlapping calls race.     */
int64_t sn_pack_info_noalloc_max_rays(void);
int sn_pack_info_noalloc(const int64_t* ray_indices, int64_t n_samples, int32_t n_rays,
                         int32_t* packed_info, int32_t* n_bad, void* stream);

/* Exclusive scan of int32 counts into packed_info rows (base, count); also  */
/* writes the grand total to *total (device int32, nullable).  This is the   */
/* cumsum step of csrc/ray_marching.cu:257-258.                             */
size_t sn_scan_workspace_bytes(int64_t n);
int sn_counts_to_packed_info(const int32_t* counts, int32_t n, int32_t* packed_info,
                             int32_t* total, void* workspace, size_t workspace_bytes,
                             void* stream);

/* p[i] = value for i in [*lo_dev * lo_mult, min(*hi_dev * hi_mult, cap)):  */
/* clears the part of a capacity-sized buffer that a sync-free iteration's   */
/* device counts say is in use (a fill kernel: safe under graph capture).    */
int sn_fill_span_dev(float* p, int64_t cap, const int32_t* lo_dev, int32_t lo_mult,
                     const int32_t* hi_dev, int32_t hi_mult, float value, void* stream);

/* ------------------------------------------------------------------------ */
/* Patch-based render weights: replaces                                      */
/*   _C.weight_from_alpha_patch_based_forward_naive  (csrc/render_weight.cu:502-543)  */
/*   _C.weight_and_transmittance_..._forward_naive    (csrc/render_weight.cu:629-667)  */
/*   _C.weight_from_alpha_patch_based_backward_naive (csrc/render_weight.cu:583-626)  */
/*   _C.weight_and_transmittance_..._backward_naive   (csrc/render_weight.cu:669-712)  */
/* alphas, weights, transmittance: float [n_samples, patch_size] (sample-     */
/* major); patch p owns samples packed_info[p,0] .. +packed_info[p,1].       */
/* w[s,k] = alpha[s,k] * T[s,k], T exclusive product of (1 - alpha) over the */
/* patch's earlier samples, for each ray k of the patch independently.       */
/* Samples not covered by any patch are left untouched (reference: empty).   */
/* n_samples bounds every access: a patch whose [base, base+count) does not  */
/* fit in [0, n_samples) is skipped.  transmittance (fwd) is nullable.       */
/* Backward uses the division-free suffix recurrence                         */
/*   g_alpha[j] = T[j] * (gw[j] - S[j]),  S[j] = gw[j+1]a[j+1] + (1-a[j+1])S[j+1] */
/* which equals the reference's (gw T - accum)/max(1-a,1e-10) wherever that  */
/* is finite; grad_T (nullable) adds d(sum gT*T)/d alpha.                    */
/* ------------------------------------------------------------------------ */
int sn_patch_weight_fwd(const int32_t* packed_info, int32_t n_patches, int32_t patch_size,
                        int64_t n_samples, const float* alphas, float* weights,
                        float* transmittance, void* stream);
int sn_patch_weight_bwd(const int32_t* packed_info, int32_t n_patches, int32_t patch_size,
                        int64_t n_samples, const float* alphas, const float* grad_weights,
                        const float* grad_transmittance, float* grad_alphas, void* stream);

/* Patch-based transmittance only: replaces                                  */
/*   _C.transmittance_from_alpha_patch_based_forward_naive (csrc/render_transmittance.cu:147-176, */
/*   with alphas[ray_id] in place of the reference's alphas[j], line 173)    */
/* The backward is sn_patch_weight_bwd with grad_weights == NULL.            */
int sn_patch_transmittance_fwd(const int32_t* packed_info, int32_t n_patches,
                               int32_t patch_size, int64_t n_samples, const float* alphas,
                               float* transmittance, void* stream);

/* ------------------------------------------------------------------------ */
/* Visibility: replaces render_visibility (NA/nerfacc/vol_rendering.py:680-748) */
/* with the per-ray sequential exclusive product of                         */
/* csrc/render_transmittance.cu:85-112.  keep[i] = T[i] >= early_stop_eps    */
/* && (alpha_thre <= 0 || alpha[i] >= alpha_thre).  keep is uint8 (bool).   */
/* ------------------------------------------------------------------------ */
int sn_visibility(const int32_t* packed_info, int32_t n_rays, int64_t n_samples,
                  const float* alphas, float early_stop_eps, float alpha_thre, uint8_t* keep,
                  void* stream);

/* Stream compaction of marching samples by a keep mask (ray_marching.py:216-220). */
/* Writes the kept (ray_indices, t_starts, t_ends) in order and the count to */
/* *n_out (device int32).  Output capacity must be >= n.                    */
size_t sn_compact_workspace_bytes(int64_t n);
int sn_compact_samples(const uint8_t* keep, int32_t n, const int64_t* ray_indices,
                       const float* t_starts, const float* t_ends, int64_t* out_ray_indices,
                       float* out_t_starts, float* out_t_ends, int32_t* n_out,
                       void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Ray marching with occupancy-grid skipping: replaces _C.ray_marching       */
/* (csrc/ray_marching.cu:81-289), AABB contraction, cone_angle 0 or > 0.     */
/* Pass 1 (count) writes num_steps[r]; pass 2 (fill) uses packed_info from   */
/* sn_counts_to_packed_info.  `max_per_ray` > 0 truncates each ray after     */
/* that many samples (0 = unbounded, the reference's behaviour).  Floating  */
/* point follows the reference kernel with no FMA contraction.              */
/* grid: uint8 [rx, ry, rz] (x-major, index x*ry*rz + y*rz + z).            */
/* step_dev (nullable): device float overriding step_size (no host sync).   */
/* ------------------------------------------------------------------------ */
int sn_ray_march_count(const float* rays_o, const float* rays_d, const float* t_min,
                       const float* t_max, int32_t n_rays, const float* roi_aabb,
                       const uint8_t* grid, int32_t rx, int32_t ry, int32_t rz,
                       float step_size, const float* step_dev, float cone_angle,
                       int32_t max_per_ray, int32_t* num_steps, void* stream);
int sn_ray_march_fill(const float* rays_o, const float* rays_d, const float* t_min,
                      const float* t_max, int32_t n_rays, const float* roi_aabb,
                      const uint8_t* grid, int32_t rx, int32_t ry, int32_t rz,
                      float step_size, const float* step_dev, float cone_angle,
                      int32_t max_per_ray, const int32_t* packed_info, int64_t* ray_indices,
                      float* t_starts, float* t_ends, void* stream);
/* One-pass form when max_per_ray > 0 bounds every ray: ray i's samples go   */
/* to slots [i*max_per_ray, i*max_per_ray + num_steps[i]) of the padded      */
/* arrays; after sn_counts_to_packed_info, sn_march_gather packs them into   */
/* exactly the (ray_indices, t_starts, t_ends) of count + fill.              */
int sn_ray_march_padded(const float* rays_o, const float* rays_d, const float* t_min,
                        const float* t_max, int32_t n_rays, const float* roi_aabb,
                        const uint8_t* grid, int32_t rx, int32_t ry, int32_t rz,
                        float step_size, const float* step_dev, float cone_angle,
                        int32_t max_per_ray, int32_t* num_steps, float* t_starts_pad,
                        float* t_ends_pad, void* stream);
/* ------------------------------------------------------------------------ */
/* Patch sampling: replaces Dataset.gen_random_patches + near_far_from_sphere */
/* (models/dataset_loader.py:223-297) and the stratified start               */
/* (renderer.py:69-71) for given random numbers: cx, cy (int64 [B], drawn as */
/* the reference's torch.randint ranges), sel (int64 [B], index into         */
/* train_idx), u ([B] in [0,1), nullable: no jitter) times *step_dev.        */
/* Dataset tensors: normals [n,H,W,3], v_inv_all [n,H,W,3,3], masks [n,H,W], */
/* k_inv [n,ks,ks] (k_stride ks = 3 or 4), pose [n,4,4].  mask_mode 1:     */
/* mask = (m > 0.5), 0: ones.  Outputs                                       */
/* as gen_random_patches (rays_o, rays_d, plane_n, v_inv, normals, mask      */
/* [B, Ph*Pw]), mask_sum = sum(mask) + 1e-5 in a fixed order, the centre    */
/* rays o_c / d_c and t_min / t_max.  Indices are clamped into range.       */
int sn_patch_sample(const int64_t* cx, const int64_t* cy, const int64_t* sel, const float* u,
                    const float* step_dev, const int64_t* train_idx, const float* normals,
                    const float* v_inv_all, const float* masks, const float* k_inv,
                    int32_t k_stride, const float* pose, int32_t n_img, int32_t n_train,
                    int32_t H, int32_t W, int32_t B, int32_t patch_h, int32_t patch_w,
                    int32_t mask_mode,
                    float* rays_o, float* rays_d, float* plane_n, float* v_inv,
                    float* normals_out, float* mask, float* mask_sum, float* o_c, float* d_c,
                    float* t_min, float* t_max, void* stream);

/* ------------------------------------------------------------------------ */
/* MLP weight packing with weight norm: replaces the torch ops of            */
/* torch.nn.utils.weight_norm (fields.py:69-70, W = g v / ||v|| per output   */
/* row) + pack_mlp's concatenation, and SingleVarianceNetwork's              */
/* inv_s = clip(exp(10 variance), 1e-6, 1e6) (fields.py:133-139).           */
/* `layers` is a HOST array of 6 int64 per layer: offsets of v, g (-1: no    */
/* weight norm, W = v) and b in `flat`, offset of the layer in `packed`,     */
/* out, in.  variance / inv_s may be NULL.                                   */
int sn_mlp_pack_fwd(const float* flat, int32_t n_layers, const int64_t* layers, float* packed,
                    const float* variance, float* inv_s, void* stream);
/* Backward of the packing: grad (same offsets as flat) += dL/dv, dL/dg,     */
/* dL/db from g_packed, and *g_variance += dL/dinv_s * d inv_s/d variance.  */
int sn_mlp_pack_bwd(const float* flat, int32_t n_layers, const int64_t* layers,
                    const float* g_packed, float* grad, const float* variance,
                    const float* g_inv_s, float* g_variance, void* stream);
/* As sn_ray_march_padded, reading occupancy from 4x4x4 brick masks (built  */
/* by sn_occ_pack_bricks from `grid`; NULL: the byte grid): one 8-byte read */
/* serves the 64 cells of a brick.  Same samples.  No counterpart in the    */
/* reference (its grid is the byte array).                                  */
int sn_ray_march_padded_bricks(const float* rays_o, const float* rays_d, const float* t_min,
                               const float* t_max, int32_t n_rays, const float* roi_aabb,
                               const uint8_t* grid, const uint64_t* bricks, int32_t rx,
                               int32_t ry, int32_t rz, float step_size, const float* step_dev,
                               float cone_angle, int32_t max_per_ray, int32_t* num_steps,
                               float* t_starts_pad, float* t_ends_pad, void* stream);
/* bricks[((x>>2)*(ry/4) + (y>>2))*(rz/4) + (z>>2)] bit (x&3)|(y&3)<<2|(z&3)<<4 */
/* = grid[x][y][z] != 0; resolutions must be multiples of 4.                 */
int sn_occ_pack_bricks(const uint8_t* binary, int32_t rx, int32_t ry, int32_t rz,
                       uint64_t* bricks, void* stream);
/* n_capped (device int32, nullable): incremented once per ray whose count  */
/* reached max_per_ray (the cap may have truncated it; the reference marches */
/* unbounded), accumulated over calls -- the caller reads and clears it.     */
int sn_march_gather(const int32_t* packed_info, int32_t n_rays, int32_t max_per_ray,
                    const float* t_starts_pad, const float* t_ends_pad, int64_t* ray_indices,
                    float* t_starts, float* t_ends, int32_t* n_capped, void* stream);

/* ------------------------------------------------------------------------ */
/* Patch accumulation: replaces accumulate_along_rays_patch_based            */
/* (NA/nerfacc/vol_rendering.py:269-335): out[ray_indices[s], k, d] +=       */
/* src[s, k, d] for src float [n, P, D].  `out` must be zeroed by the caller */
/* (float atomics, like the reference's scatter_add_).                      */
/* ------------------------------------------------------------------------ */
int sn_scatter_add_rows(const int64_t* ray_indices, int64_t n, int32_t row_elems,
                        const float* src, float* out, int32_t n_out_rows, void* stream);

/* ------------------------------------------------------------------------ */
/* Occupancy grid (NA/nerfacc/grid.py:196-239, csrc/contraction.cu:36-61).   */
/* sn_occ_cell_points: x[i] = ((coord(idx[i]) + jitter[i]) / res) mapped to  */
/*   the roi, coord(c) = (c / (ry*rz), (c / rz) % ry, c % rz).               */
/* sn_occ_ema_max: occs[c] = max(decay*occs[c], max_{i: idx[i]==c} occ[i]) - */
/*   a deterministic scatter-max (the reference's index_put has undefined   */
/*   order for duplicate indices, grid.py:232).                             */
/* sn_occ_binarize: binary[c] = occs[c] > min(mean(occs), thre); the mean is */
/*   an fp64 sum in fixed chunks of 1024 cells (sequential within a chunk,  */
/*   chunks summed sequentially) so the bits are reproducible.  Also writes  */
/*   the threshold used to *thre_out (device float, nullable).              */
/* sn_occ_nonzero: ordered list of occupied cell indices + count.           */
/* ------------------------------------------------------------------------ */
int sn_occ_cell_points(const int64_t* cell_idx, int32_t n, const float* jitter,
                       int32_t rx, int32_t ry, int32_t rz, const float* roi_aabb, float* x,
                       void* stream);
size_t sn_occ_workspace_bytes(int32_t n_cells);
int sn_occ_ema_max(float* occs, int32_t n_cells, const int64_t* cell_idx, const float* occ,
                   int32_t n, float decay, void* workspace, size_t workspace_bytes,
                   void* stream);
int sn_occ_binarize(const float* occs, int32_t n_cells, float thre, uint8_t* binary,
                    float* thre_out, void* workspace, size_t workspace_bytes, void* stream);
int sn_occ_nonzero(const uint8_t* binary, int32_t n_cells, int64_t* out_idx,
                   int32_t* count, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Multiresolution hash-grid encoding: replaces tinycudann.Encoding with     */
/* otype=HashGrid, 3-D input, linear interpolation (models/fields.py:25-31, */
/* 78).  Level l: scale = base*2^(l*log2(per_level_scale)) - 1, res =       */
/* ceil(scale)+1, pos = fma(scale, x, 0.5), cell = floor(pos) as uint32     */
/* (inputs outside [0,1] wrap, as in tcnn), entry = dense index when res^3  */
/* fits the level's table, else xor-prime hash (1, 2654435761, 805459861),  */
/* modulo the level size.  params: float, level-major, features interleaved. */
/* Output enc[i, l*F + f]; levels >= n_active_levels are written as 0 and   */
/* receive no gradient (the bandwidth mask of fields.py:80-83).             */
/* ------------------------------------------------------------------------ */
/* Fills offsets[0..n_levels] (entries, not floats); returns total entries. */
int64_t sn_hashgrid_layout(int32_t n_levels, int32_t log2_hashmap_size,
                           int32_t base_resolution, float per_level_scale, uint32_t* offsets);
int sn_hashgrid_fwd(const float* x, int64_t n, const float* params, int32_t n_levels,
                    int32_t n_features, int32_t log2_hashmap_size, int32_t base_resolution,
                    float per_level_scale, int32_t n_active_levels, float* enc,
                    void* stream);
/* dL/dparams += sum over points (atomic fp32 adds; caller zeroes).  grad_x  */
/* (nullable) = dL/dx (overwritten).  Before the atomics, contributions to  */
/* the same table row are merged: mode 2 (default, F <= 2) per workgroup in */
/* an LDS hash table over all levels, mode 1 across a wavefront on coarse   */
/* levels, mode 0 none.  sn_hashgrid_set_aggregation(mode) is a process-    */
/* wide switch (no counterpart in the reference; for measurement and tests).*/
void sn_hashgrid_set_aggregation(int32_t mode);
int sn_hashgrid_bwd(const float* x, int64_t n, const float* params, int32_t n_levels,
                    int32_t n_features, int32_t log2_hashmap_size, int32_t base_resolution,
                    float per_level_scale, int32_t n_active_levels, const float* grad_enc,
                    float* grad_params, float* grad_x, void* stream);
/* Backward of the map (grad_enc, params, x) -> grad_x = sum dL/denc * denc/dx */
/* given its cotangent g_gx [n,3] (tcnn's backward_backward_input):         */
/*   grad_grad_enc [n, L*F] (nullable, overwritten)                          */
/*   grad_params  (nullable, atomically accumulated)                         */
/*   grad_x [n,3] (nullable, overwritten): the trilinear cross terms         */
int sn_hashgrid_bwd_bwd(const float* x, int64_t n, const float* params, int32_t n_levels,
                        int32_t n_features, int32_t log2_hashmap_size, int32_t base_resolution,
                        float per_level_scale, int32_t n_active_levels, const float* grad_enc,
                        const float* g_gx, float* grad_grad_enc, float* grad_params,
                        float* grad_x, void* stream);

/* ------------------------------------------------------------------------ */
/* SDF MLP (models/fields.py:39-99 with input_concat): h0 = [x, enc] of     */
/* width d_in = 3 + enc_dims (<= 64); n_layers hidden layers (1 or 2) of     */
/* width d_hidden (32 or 64) with Softplus(beta), threshold 20; output 1.    */
/* Weights are the weight-normed (folded) matrices, packed fp32 row-major:   */
/*   W0[H][d_in] b0[H] { W1[H][H] b1[H] } wo[H] bo                           */
/* Runs on the fp32 matrix cores.  n_cap is the buffer capacity (rows);      */
/* n_dev (nullable, device int32) the number of valid rows, so callers can   */
/* size work on the device without a host round trip.                        */
/* ------------------------------------------------------------------------ */
int sn_mlp_n_params(int32_t d_in, int32_t d_hidden, int32_t n_layers);
/* sdf[n]; dsdf_dh0 (nullable) [n, d_in] = d sdf / d h0 (for the analytic    */
/* x-gradient, fields.py:107-119: grad = dsdf_dh0[:, :3] + the hash-grid     */
/* backward of dsdf_dh0[:, 3:]).                                             */
int sn_mlp_fwd(const float* x, const float* enc, int32_t enc_dims, int64_t n_cap,
               const int32_t* n_dev, const float* mlp, int32_t d_hidden, int32_t n_layers,
               float beta, float* sdf, float* dsdf_dh0, void* stream);
size_t sn_mlp_bwd_workspace_bytes(int64_t n_cap, int32_t enc_dims, int32_t d_hidden,
                                  int32_t n_layers, int32_t analytic);
/* Backward of sum g_sdf * sdf: g_mlp (packed layout; overwritten, or added  */
/* to when accumulate != 0) summed over rows in a fixed order; g_enc [n,E]   */
/* and g_x [n,3] (both nullable) overwritten.                                */
int sn_mlp_bwd(const float* x, const float* enc, int32_t enc_dims, int64_t n_cap,
               const int32_t* n_dev, const float* mlp, int32_t d_hidden, int32_t n_layers,
               float beta, const float* g_sdf, float* g_enc, float* g_x, float* g_mlp,
               int32_t accumulate, void* workspace, size_t workspace_bytes, void* stream);
/* Backward of J = sum g_sdf * sdf + v . (d sdf/dx) (double backward of the   */
/* analytic gradient).  enc_dot [n,E] = directional derivative of enc along */
/* v (sn_hashgrid_bwd_bwd's grad_grad_enc with g_gx = v).  Outputs g_mlp as  */
/* sn_mlp_bwd, g_h0 = dJ/dh0 and g_h0_dot = dJ/d(h0 tangent), both [n, d_in]. */
int sn_mlp_bwd_ad(const float* x, const float* enc, const float* v, const float* enc_dot,
                  int32_t enc_dims, int64_t n_cap, const int32_t* n_dev, const float* mlp,
                  int32_t d_hidden, int32_t n_layers, float beta, const float* g_sdf,
                  float* g_h0, float* g_h0_dot, float* g_mlp, int32_t accumulate,
                  void* workspace, size_t workspace_bytes, void* stream);
/* As sn_mlp_bwd_ad with only the encoding's columns of g_h0 / g_h0_dot,     */
/* written as contiguous [n, enc_dims] rows g_enc / g_enc_dot (the layout the */
/* hash-grid backward reads; positions are constants of the iteration).      */
int sn_mlp_bwd_ad_enc(const float* x, const float* enc, const float* v, const float* enc_dot,
                      int32_t enc_dims, int64_t n_cap, const int32_t* n_dev, const float* mlp,
                      int32_t d_hidden, int32_t n_layers, float beta, const float* g_sdf,
                      float* g_enc, float* g_enc_dot, float* g_mlp, int32_t accumulate,
                      void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Fused patch rendering + losses, forward and backward: replaces the torch  */
/* graph of models/renderer.py:142-276 (NeuS alpha, patch weights, 'dfd' or  */
/* 'ad' per-sample normals, accumulation) and exp_runner.py:169-203 (normal  */
/* l2/l1, eikonal, mask BCE) and its autograd backward.                      */
/* Per patch b (n_patches, patch_h x patch_w rays, centre ray = middle):     */
/*   rays_o, rays_d, n_true [B,P,3], plane_n [B,3], v_inv [B,P,3,3] ('dfd'),  */
/*   mask [B,P] (binarised), packed_info [B,2] of the centre samples,        */
/*   t_starts/t_ends [S] (centre ray), diff_idx [S] = row of the sample's end */
/*   in sdf_ends or -1 when the end is the next sample's start (same patch), */
/*   sdf_starts [S,P], sdf_ends [Nd,P], grad_starts [S,P,3] ('ad', else null),*/
/*   inv_s, mask_sum (= sum(mask) + 1e-5) and n_samples (= S) on the device.  */
/*   sdf_ends = NULL: the ends follow the starts in one buffer (row S*P on,   */
/*   S read on the device), and so do their gradients in g_sdf_starts.       */
/* Outputs: dL/dsdf_starts [S,P], dL/dsdf_ends [Nd,P], dL/dgrad [S,P,3]      */
/* ('ad'), losses[4] = (normal, mask, eikonal, dL/dinv_s) (weighted terms),  */
/* weight_sum [B,P] and comp_normal [B,P,3] (nullable).  partials needs      */
/* 4 * sn_render_step_waves() floats.                                        */
/* ------------------------------------------------------------------------ */
int sn_render_step_waves(int32_t n_patches, int32_t patch_size);
/* Kernel choice (no counterpart in the reference; for tests): 1 (default)   */
/* = sample-parallel kernel for 3x3 and 2x4 patches, 0 = ray-parallel.     */
void sn_render_step_set_impl(int32_t impl);
int sn_render_step(int32_t n_patches, int32_t patch_h, int32_t patch_w, const float* rays_o,
                   const float* rays_d, const float* plane_n, const float* v_inv,
                   const float* n_true, const float* mask, const int32_t* packed_info,
                   const float* t_starts, const float* t_ends, const int32_t* diff_idx,
                   const float* sdf_starts, const float* sdf_ends, const float* grad_starts,
                   const float* inv_s, const float* mask_sum, const int32_t* n_samples,
                   float normal_w, float eik_w, float mask_w, int32_t loss_l1,
                   float* g_sdf_starts, float* g_sdf_ends, float* g_grad_starts,
                   float* losses, float* partials, float* weight_sum, float* comp_normal,
                   void* stream);

/* The same MLP on the fp16 (forward) / bf16 (backward) matrix cores with  */
/* fp32 accumulation: weights rounded once to the MFMA type, activations    */
/* and gradients carried as hi+lo pairs (fields.py:85-99 as sn_mlp_fwd /    */
/* sn_mlp_bwd; sdf only, no dsdf/dh0).  n_dev as in the device-count       */
/* variants below (nullable).  g_enc [n,E] / g_x [n,3] nullable.            */
int sn_mlp16_fwd(const float* x, const float* enc, int32_t enc_dims, int64_t n_cap,
                 const int32_t* n_dev, const float* mlp, int32_t d_hidden, int32_t n_layers,
                 float beta, float* sdf, void* stream);
/* sn_mlp16_fwd plus d sdf / d h0 [n, 3 + enc_dims] (columns [x, enc]): the   */
/* MLP's part of SDFNetwork.gradient (models/fields.py:107-119, the rendered  */
/* normal of the analytic-gradient method, renderer.py:225-226), replacing    */
/* autograd.grad of the torch MLP; the same fp16 weights backward, with the   */
/* cotangents as hi + lo fp16 parts.  The table's part: sn_hashgrid_dx_dev.   */
int sn_mlp16_fwd_grad(const float* x, const float* enc, int32_t enc_dims, int64_t n_cap,
                      const int32_t* n_dev, const float* mlp, int32_t d_hidden, int32_t n_layers,
                      float beta, float* sdf, float* dh0, void* stream);
/* sn_mlp_bwd_ad_enc on the matrix cores (the double backward of the         */
/* analytic gradient, fields.py:107-119 with create_graph=True): the forward  */
/* and its tangent along (v, enc_dot) recomputed on fp16, the two reverse     */
/* passes on bf16 with fp32 accumulation.  Workspace as sn_mlp16_bwd.         */
int sn_mlp16_bwd_ad_enc(const float* x, const float* enc, const float* v, const float* enc_dot,
                        int32_t enc_dims, int64_t n_cap, const int32_t* n_dev, const float* mlp,
                        int32_t d_hidden, int32_t n_layers, float beta, const float* g_sdf,
                        float* g_enc, float* g_enc_dot, float* g_mlp, int32_t accumulate,
                        void* workspace, size_t workspace_bytes, void* stream);
size_t sn_mlp16_bwd_workspace_bytes(int64_t n_cap, int32_t enc_dims, int32_t d_hidden,
                                    int32_t n_layers);
int sn_mlp16_bwd(const float* x, const float* enc, int32_t enc_dims, int64_t n_cap,
                 const int32_t* n_dev, const float* mlp, int32_t d_hidden, int32_t n_layers,
                 float beta, const float* g_sdf, float* g_enc, float* g_x, float* g_mlp,
                 int32_t accumulate, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Surface extraction: replaces PyMCubes' marching cubes in extract_geometry */
/* (renderer.py:27-34) with marching tetrahedra (6 per cube around the main  */
/* diagonal).  values [rx,ry,rz] (x-major); inside = value < iso.  Pass 1:   */
/* counts[cube] = triangles; sn_counts_to_packed_info scans them; pass 2     */
/* writes 9 floats (3 world vertices) per triangle at each cube's base,      */
/* oriented inside -> outside; world = bound_min + grid * (max-min)/(r-1).   */
/* ------------------------------------------------------------------------ */
int sn_mt_count(const float* values, int32_t rx, int32_t ry, int32_t rz, float iso,
                int32_t* counts, void* stream);
int sn_mt_emit(const float* values, int32_t rx, int32_t ry, int32_t rz, float iso,
               const float* bound_min, const float* bound_max, const int32_t* packed_info,
               float* triangles, void* stream);

/* Marching cubes: replaces PyMCubes' marching_cubes(u, threshold) as called */
/* by extract_geometry (models/renderer.py:27-34).  Corner case = sum of     */
/* (value < iso) << corner in the usual corner numbering; one vertex per     */
/* grid edge whose ends straddle iso, at the linear interpolation of the     */
/* values along the edge, so the vertex set is the one any marching-cubes    */
/* table gives; the triangulation of each case (faces with diagonal inside   */
/* corners keep those corners apart, so the surface has no cracks) is        */
/* derived by supernormal_amd/csrc/gen_mc_tables.py.  Triangles are wound so  */
/* that their normal points towards value > iso.  sn_mc_count writes the     */
/* triangle count of each of the (rx-1)(ry-1)(rz-1) cubes (x-major, as the   */
/* values); scan it with sn_counts_to_packed_info; sn_mc_emit writes each    */
/* cube's triangles at its base: 9 floats per triangle (world space, grid    */
/* point (i,j,k) at bound_min + (i,j,k) * (bound_max-bound_min)/(r-1)) and,  */
/* if vertex_ids is not NULL, 3 int64 per triangle: the vertex's edge id      */
/* 3 * (lower grid point's linear index) + axis (equal ids = equal vertex).  */
int sn_mc_count(const float* values, int32_t rx, int32_t ry, int32_t rz, float iso,
                int32_t* counts, void* stream);
int sn_mc_emit(const float* values, int32_t rx, int32_t ry, int32_t rz, float iso,
               const float* bound_min, const float* bound_max, const int32_t* packed_info,
               float* triangles, int64_t* vertex_ids, void* stream);

/* First hit of each ray with a mesh from sn_mc_emit: replaces trimesh's     */
/* mesh.ray.intersects_location(multiple_hits=False) in find_visible_points */
/* (exp_runner.py:618-630).  The ray walks the grid's cubes in order from    */
/* max(t_near, box entry) to min(t_far, box exit) (t_near / t_far nullable:  */
/* 0 and +inf) and tests the triangles of each cube (packed_info from the    */
/* scan of sn_mc_count).  t_hit = ray parameter of the nearest hit (+inf if  */
/* none; rays_d need not be unit), tri_hit = its triangle or -1.  bound_min  */
/* and bound_max are host pointers (3 floats), as in sn_mc_emit.             */
int sn_mesh_first_hit(const float* rays_o, const float* rays_d, const float* t_near,
                      const float* t_far, int32_t n_rays, int32_t rx, int32_t ry, int32_t rz,
                      const float* bound_min, const float* bound_max,
                      const int32_t* packed_info, const float* triangles, float* t_hit,
                      int32_t* tri_hit, void* stream);

/* ------------------------------------------------------------------------ */
/* Adam (torch.optim.Adam defaults, exp_runner.py:97, 205-207) over one flat */
/* fp32 parameter buffer: m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;        */
/* p -= lr/(1-b1^t) * m / (sqrt(v)/sqrt(1-b2^t) + eps).  `lr` is read from   */
/* device memory (lr_dev) when non-null so that a captured graph can replay */
/* with a schedule; else the scalar lr is used.  Likewise the 1-based step  */
/* t is read from step_dev (device int32) when non-null, else `step`.       */
/* ------------------------------------------------------------------------ */
int sn_adam_step(float* params, const float* grads, float* exp_avg, float* exp_avg_sq,
                 int64_t n, float lr, const float* lr_dev, double beta1, double beta2,
                 double eps, int32_t step, const int32_t* step_dev, void* stream);
/* The same step, also refreshing an fp16 copy (copy16, nullable) of params   */
/* [copy_offset, copy_offset + copy_n): the gather copy of the hash table     */
/* (tcnn keeps its parameters' fp16 copy the same way).  run_dev (device     */
/* int32, nullable): when *run_dev <= 0 nothing is updated -- the iteration   */
/* had no sample (exp_runner.py:182 never reaches optimizer.step then).      */
int sn_adam_step_copy16(float* params, const float* grads, float* exp_avg, float* exp_avg_sq,
                        int64_t n, float lr, const float* lr_dev, double beta1, double beta2,
                        double eps, int32_t step, const int32_t* step_dev, void* copy16,
                        int64_t copy_offset, int64_t copy_n, const int32_t* run_dev,
                        void* stream);
/* Device step count of the optimiser (torch.optim.Adam's state['step'],     */
/* exp_runner.py:207): ++*step_dev when *run_dev > 0, else ++*skipped_dev    */
/* (nullable) -- the count of iterations skipped for want of samples.        */
int sn_step_advance(int32_t* step_dev, const int32_t* run_dev, int32_t* skipped_dev,
                    void* stream);

/* Adam steps of several flat parameter buffers in one launch (the serving   */
/* loop's resident jobs; no counterpart in the reference, whose one          */
/* optimiser is torch.optim.Adam, exp_runner.py:97, 207).  segments: a DEVICE */
/* array of n_segments SnAdamSegment in block order: segment s owns blocks   */
/* [block0, block0 + n_blocks) of a launch of total_blocks blocks; n_blocks  */
/* = sn_adam_multi_blocks(n).  Each segment reads its learning rate and step */
/* count (both required: a segment without them is left alone) and run flag */
/* (nullable: always run) from device memory, exactly as                    */
/* sn_adam_step_copy16 with lr_dev / step_dev / run_dev; copy16 (nullable)   */
/* is its fp16 copy of [copy_off, copy_off + copy_n).  Buffers 16-byte        */
/* aligned.                                                                   */
typedef struct SnAdamSegment {
  float* params;
  const float* grads;
  float* exp_avg;
  float* exp_avg_sq;
  int64_t n;
  const float* lr_dev;
  const int32_t* step_dev;
  const int32_t* run_dev;
  void* copy16;
  int64_t copy_off;
  int64_t copy_n;
  int32_t block0;
  int32_t n_blocks;
} SnAdamSegment;
int32_t sn_adam_multi_blocks(int64_t n);
int sn_adam_multi(const SnAdamSegment* segments, int32_t n_segments, int32_t total_blocks,
                  double beta1, double beta2, double eps, void* stream);

/* Step guard (no counterpart in the reference, which has no NaN check):     */
/* run = 1 if *n_samples > 0, the n_losses losses are finite and every one    */
/* of the n gradients is finite, else 0; pass run to sn_step_advance and     */
/* sn_adam_step_copy16 (run_dev) so that such an iteration leaves the        */
/* parameters, moments and step count unchanged.  n_nonfinite (nullable,     */
/* device int32) counts the iterations that had samples but a non-finite     */
/* loss or gradient.  grads: 16-byte aligned.                                */
int sn_step_guard(const int32_t* n_samples, const float* losses, int32_t n_losses,
                  const float* grads, int64_t n, int32_t* run, int32_t* n_nonfinite,
                  void* stream);
/* dst[i] = (fp16) src[i] (initialises such a copy). */
int sn_to_half(const float* src, void* dst, int64_t n, void* stream);

/* ------------------------------------------------------------------------ */
/* Device-count variants for an iteration with no host round trip.  Each    */
/* takes a capacity (`n_cap`, sizes the launch and the buffers) and a device */
/* int32 count `n_dev` (the valid prefix; nullable = n_cap), so sizes that  */
/* depend on the data never travel to the host.  Same results as the plain  */
/* entry points on the valid prefix.                                        */
/* ------------------------------------------------------------------------ */
/* nerfacc.pack_info (NA/nerfacc/pack.py:46-77) over the first *n_dev samples. */
int sn_pack_info_dev(const int64_t* ray_indices, int64_t n_cap, const int32_t* n_dev,
                     int32_t n_rays, int32_t* packed_info, int32_t* n_bad, void* workspace,
                     size_t workspace_bytes, void* stream);
/* Compaction by a keep mask (ray_marching.py:216-220) over *n_dev samples. */
/* With packed_in (the rays' packed_info before compaction, nullable) it    */
/* also writes packed_out = pack_info of the kept samples (pack.py:46-77)   */
/* without counting them again.                                             */
int sn_compact_samples_dev(const uint8_t* keep, int32_t n_cap, const int32_t* n_dev,
                           const int64_t* ray_indices, const float* t_starts,
                           const float* t_ends, int64_t* out_ray_indices, float* out_t_starts,
                           float* out_t_ends, int32_t* n_out, const int32_t* packed_in,
                           int32_t n_rays, int32_t* packed_out, void* workspace,
                           size_t workspace_bytes, void* stream);

/* Visibility pre-pass points (renderer.py:80-100): for the marched centre   */
/* samples, the start points [n] followed by the end points of the samples  */
/* whose end is not the next sample's start on the same ray [n_diff].       */
/* diff_idx[s] = slot of sample s's end point in that tail, or -1.          */
/* *n_diff and *n_eval = n + n_diff are written on the device.              */
/* rays_o_c/rays_d_c: centre-ray origins/directions with row stride         */
/* `centre_stride` floats (9*3 for 3x3 patches read in place).              */
size_t sn_step_workspace_bytes(int64_t n_cap);
int sn_centre_points(const int64_t* ray_indices, const float* t_starts, const float* t_ends,
                     int64_t n_cap, const int32_t* n_dev, const float* rays_o_c,
                     const float* rays_d_c, int32_t centre_stride, int32_t* diff_idx,
                     int32_t* n_diff, float* points, int32_t* n_eval, void* workspace,
                     size_t workspace_bytes, void* stream);
/* NeuS alpha of each centre sample from the pre-pass SDF values           */
/* (renderer.py:102-118) and the transmittance test of                      */
/* ray_marching.py:193-215: keep[s] = T_before(s) >= early_stop_eps.        */
int sn_alpha_visibility(const int32_t* packed_info, int32_t n_rays, const float* sdf,
                        const int32_t* diff_idx, int64_t n_cap, const int32_t* n_dev,
                        const float* inv_s, float early_stop_eps, uint8_t* keep,
                        void* stream);
/* Patch expansion (renderer.py:142-159): each centre sample's start (and    */
/* non-contiguous end) moved onto the patch's marching plane for all P rays  */
/* of the patch; points laid out [n*P starts | n_diff*P ends].              */
int sn_patch_points(const int64_t* ray_indices, const float* t_starts, const float* t_ends,
                    int64_t n_cap, const int32_t* n_dev, const float* rays_o,
                    const float* rays_d, const float* plane_n, int32_t patch_h,
                    int32_t patch_w, int32_t* diff_idx, int32_t* n_diff, float* points,
                    int32_t* n_eval, void* workspace, size_t workspace_bytes, void* stream);
/* Hash-grid encode reading an fp16 copy of the table (blending in fp32).   */
int sn_hashgrid_fwd16_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                          const void* params16, int32_t n_levels, int32_t n_features,
                          int32_t log2_hashmap_size, int32_t base_resolution,
                          float per_level_scale, int32_t n_active_levels,
                        const int32_t* n_active_dev, float* enc,
                          void* stream);
/* Hash-grid encode / backward / double backward over the first *n_dev      */
/* points of an n_cap-point buffer (tcnn Encoding, as sn_hashgrid_*).       */
/* n_active_dev (nullable): device copy of the active level count, at most  */
/* n_active_levels, read by the kernels, so that a captured launch follows  */
/* the bandwidth schedule without re-capture.                                */
int sn_hashgrid_fwd_dev(const float* x, int64_t n_cap, const int32_t* n_dev, const float* params,
                        int32_t n_levels, int32_t n_features, int32_t log2_hashmap_size,
                        int32_t base_resolution, float per_level_scale, int32_t n_active_levels,
                        const int32_t* n_active_dev,
                        float* enc, void* stream);
int sn_hashgrid_bwd_dev(const float* x, int64_t n_cap, const int32_t* n_dev, const float* params,
                        int32_t n_levels, int32_t n_features, int32_t log2_hashmap_size,
                        int32_t base_resolution, float per_level_scale, int32_t n_active_levels,
                        const int32_t* n_active_dev,
                        const float* grad_enc, float* grad_params, float* grad_x, void* stream);
/* The analytic x-gradient of the SDF given the MLP's d sdf / d h0 rows      */
/* (sn_mlp_fwd's dsdf_dh0, h0 = [x, enc], row stride ld >= 3 + L*F):         */
/* grad [n,3] = dsdf_dh0[:, :3] + sum over levels/corners of the table's     */
/* contribution (fields.py:107-119: autograd.grad of the sdf w.r.t. x).      */
/* Reads the fp32 parameters, or the fp16 copy params16 when not NULL.       */
/* n_dev / n_active_dev: device counts as in sn_hashgrid_fwd_dev.            */
int sn_hashgrid_dx_dev(const float* x, int64_t n_cap, const int32_t* n_dev, const float* params,
                       const void* params16, int32_t n_levels, int32_t n_features,
                       int32_t log2_hashmap_size, int32_t base_resolution,
                       float per_level_scale, int32_t n_active_levels,
                       const int32_t* n_active_dev, const float* dsdf_dh0, int32_t ld,
                       float* grad, void* stream);
/* Backward as sn_hashgrid_bwd_dev with a workspace (F <= 2, mode 2), no   */
/* fp32 atomics: corner contributions (summed across the lanes of a         */
/* wavefront whose points share a cell) become records in per-bucket lists  */
/* (8192 table rows per bucket); a second kernel sums each list as 64-bit   */
/* fixed point scaled by the level's largest value and adds the sums to the */
/* rows it owns.  Records beyond a list's capacity take fp32 atomics; with  */
/* none, the result is bit-reproducible.  The lists share one pool, whose   */
/* size follows from workspace_bytes (sn_hashgrid_bwd_workspace_bytes gives */
/* the bytes for records_per_bucket times the bucket count; -1: unsupported */
/* config).  Each call shares the pool out by the counts the previous call  */
/* on the same workspace left in it (5/4 of each list's count plus a floor, */
/* at most 262144 records or the even split if larger; scaled down to the   */
/* pool, or sharing what they leave of it evenly up to the same cap; the    */
/* even split for a workspace no call has used): the                        */
/* workspace carries this hint between calls, any content is valid input,  */
/* and calls that share a workspace must not overlap.  Same sums as         */
/* sn_hashgrid_bwd_dev to fp32 rounding.                                     */
int64_t sn_hashgrid_bwd_workspace_bytes(int32_t n_levels, int32_t n_features,
                                        int32_t log2_hashmap_size, int32_t base_resolution,
                                        float per_level_scale, int32_t records_per_bucket);
int sn_hashgrid_bwd_ws_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                           const float* params, int32_t n_levels, int32_t n_features,
                           int32_t log2_hashmap_size, int32_t base_resolution,
                           float per_level_scale, int32_t n_active_levels,
                        const int32_t* n_active_dev, const float* grad_enc,
                           float* grad_params, float* grad_x, void* workspace,
                           size_t workspace_bytes, void* stream);

int sn_hashgrid_bwd_bwd_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                            const float* params, int32_t n_levels, int32_t n_features,
                            int32_t log2_hashmap_size, int32_t base_resolution,
                            float per_level_scale, int32_t n_active_levels,
                        const int32_t* n_active_dev,
                            const float* grad_enc, const float* g_gx, float* grad_grad_enc,
                            float* grad_params, float* grad_x, void* stream);
/* The encoding's directional derivative along g_gx (sn_hashgrid_bwd_bwd_dev */
/* with only grad_grad_enc: tcnn's backward_backward_input's dL/d(dL/dy)    */
/* term, models/fields.py:107-119) from the fp32 table or, when params16 is */
/* given, its fp16 copy (sn_hashgrid_fwd16_dev's); F = 2 with the levels a   */
/* multiple of 4 gathers 4 levels at a time, other configs need params.     */
int sn_hashgrid_tangent_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                            const float* params, const void* params16, int32_t n_levels,
                            int32_t n_features, int32_t log2_hashmap_size,
                            int32_t base_resolution, float per_level_scale,
                            int32_t n_active_levels, const int32_t* n_active_dev,
                            const float* g_gx, float* grad_grad_enc, void* stream);
/* The table gradient of the analytic-gradient iteration in one pass: the   */
/* first-order term of grad_enc (as sn_hashgrid_bwd_ws_dev) plus the double */
/* backward's table term that sn_hashgrid_bwd_bwd_dev adds for              */
/* (grad_enc = grad_enc2, g_gx): per point, level and corner               */
/* scale_l (g_gx . d w_c / d frac) grad_enc2[l, f].  Replaces tcnn's        */
/* backward + backward_backward_input pair behind                           */
/* autograd.grad(create_graph=True) (models/fields.py:107-119) for the     */
/* table.  With a workspace the two terms share each corner's bucket record */
/* (fixed-point sums, no fp32 atomic); without one (or F > 2) the second    */
/* term is added with fp32 atomics.  Accumulates into grad_params.          */
int sn_hashgrid_bwd_ad_ws_dev(const float* x, int64_t n_cap, const int32_t* n_dev,
                              const float* params, int32_t n_levels, int32_t n_features,
                              int32_t log2_hashmap_size, int32_t base_resolution,
                              float per_level_scale, int32_t n_active_levels,
                              const int32_t* n_active_dev, const float* grad_enc,
                              const float* g_gx, const float* grad_enc2, float* grad_params,
                              void* workspace, size_t workspace_bytes, void* stream);
/* Cells of an occupancy update after warm-up (NA/nerfacc/grid.py:181-194): */
/* n_cells/4 uniform cells then up to n_cells/4 occupied cells (all of them */
/* in order when fewer), chosen from device uniforms.  *n_out on device.    */
size_t sn_occ_sample_workspace_bytes(int32_t n_cells);
int sn_occ_sample_cells(const uint8_t* binary, int32_t n_cells, const float* u_uniform,
                        const float* u_select, int64_t* out_idx, int32_t* n_out,
                        void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SN_API_H */
