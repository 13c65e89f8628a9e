/*
 * sanitize_check.c -- runs every function of sn_oracle.c on small seeded
 * inputs held in heap buffers of exactly the size each function may touch,
 * so that an AddressSanitizer / UndefinedBehaviorSanitizer build of the
 * oracle (make -C oracle sanitize) reports any access outside them or any
 * undefined arithmetic.  TEST INFRASTRUCTURE ONLY, like sn_oracle.c.
 *
 * Cases: empty rays and patches (zero samples), a zero-size batch, the
 * per-ray sample cap, out-of-range ray indices in pack_info, axis-aligned
 * ray directions (infinite 1/d), NaN occupancies, non-cubic grids, dense and
 * hashed hash-grid levels.  Each case also checks a property the result must
 * have (counts summing to the sample count, weights summing to 1 - T_end,
 * bounded transmittance), so a sanitized build that runs but computes
 * nonsense fails too.  Exit status 0 when every check holds.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int or_pack_info(const int64_t* ri, int64_t n, int32_t n_rays, int32_t* packed);
void or_patch_weight_fwd(const int32_t* packed, int32_t n_patches, int32_t P,
                         const float* alphas, float* w, float* trans);
void or_patch_weight_bwd_ref(const int32_t* packed, int32_t n_patches, int32_t P,
                             const float* alphas, const float* w, const float* gw, float* ga);
void or_patch_weight_bwd_f64(const int32_t* packed, int32_t n_patches, int32_t P,
                             const double* alphas, const double* gw, const double* gT,
                             double* ga);
void or_visibility(const int32_t* packed, int32_t n_rays, const float* alphas, float eps,
                   float alpha_thre, uint8_t* keep);
void or_ray_march_count(const float* rays_o, const float* rays_d, const float* tmin,
                        const float* tmax, int32_t n_rays, const float* roi,
                        const uint8_t* grid, int32_t rx, int32_t ry, int32_t rz, float step,
                        float cone, int32_t max_per_ray, int32_t* num_steps);
void or_ray_march_fill(const float* rays_o, const float* rays_d, const float* tmin,
                       const float* tmax, int32_t n_rays, const float* roi, const uint8_t* grid,
                       int32_t rx, int32_t ry, int32_t rz, float step, float cone,
                       int32_t max_per_ray, const int32_t* packed, int64_t* ri, float* t0,
                       float* t1);
void or_occ_cell_points(const int64_t* idx, int32_t n, const float* jit, int32_t rx, int32_t ry,
                        int32_t rz, const float* roi, float* x);
void or_occ_ema_max(float* occs, int32_t n_cells, const int64_t* idx, const float* occ, int32_t n,
                    float decay);
float or_occ_binarize(const float* occs, int32_t n_cells, float thre, uint8_t* binary);
void or_hashgrid_corners(const float* x, int64_t n, int32_t n_levels, const float* scale,
                         const uint32_t* res, const uint32_t* hsize, const uint8_t* dense,
                         uint32_t* idx_out, float* frac_out, float* pos_out);

static int failures = 0;
#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);                       \
      fprintf(stderr, "\n");                              \
      ++failures;                                         \
    }                                                     \
  } while (0)

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static float urand(void) { /* splitmix64 -> [0, 1) */
  uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (float)(z >> 40) / (float)(1u << 24);
}

/* exact-size allocation; a zero-size request still returns a unique pointer */
static void* xalloc(size_t n) {
  void* p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "out of memory\n");
    exit(2);
  }
  return p;
}

/* packed_info of n_rays rays with the given counts */
static int32_t* packed_of(const int32_t* counts, int32_t n_rays, int64_t* total) {
  int32_t* packed = xalloc(sizeof(int32_t) * 2 * (size_t)n_rays);
  int64_t c = 0;
  for (int32_t r = 0; r < n_rays; ++r) {
    packed[2 * r] = (int32_t)c;
    packed[2 * r + 1] = counts[r];
    c += counts[r];
  }
  *total = c;
  return packed;
}

static void check_pack_info(void) {
  const int32_t n_rays = 37;
  for (int pass = 0; pass < 2; ++pass) {
    const int64_t n = pass ? 0 : 500;
    int64_t* ri = xalloc(sizeof(int64_t) * (size_t)n);
    int expect_bad = 0;
    for (int64_t i = 0; i < n; ++i) {
      ri[i] = (int64_t)(urand() * (n_rays + 4)) - 2; /* a few below 0 and above n_rays - 1 */
      expect_bad += ri[i] < 0 || ri[i] >= n_rays;
    }
    int32_t* packed = xalloc(sizeof(int32_t) * 2 * n_rays);
    const int bad = or_pack_info(ri, n, n_rays, packed);
    CHECK(bad == expect_bad, "pack_info bad %d != %d", bad, expect_bad);
    const int64_t last = (int64_t)packed[2 * (n_rays - 1)] + packed[2 * (n_rays - 1) + 1];
    CHECK(last == n - bad, "pack_info total %lld != %lld", (long long)last, (long long)(n - bad));
    free(packed);
    free(ri);
  }
}

static void check_patch_render(void) {
  const int32_t n_patches = 11, P = 9;
  int32_t counts[11];
  for (int32_t p = 0; p < n_patches; ++p) counts[p] = (p % 4 == 0) ? 0 : (int32_t)(urand() * 40);
  counts[n_patches - 1] = 0; /* an empty last patch */
  int64_t n;
  int32_t* packed = packed_of(counts, n_patches, &n);
  const size_t m = (size_t)n * P;
  float* a = xalloc(sizeof(float) * m);
  float* w = xalloc(sizeof(float) * m);
  float* T = xalloc(sizeof(float) * m);
  float* gw = xalloc(sizeof(float) * m);
  float* ga = xalloc(sizeof(float) * m);
  double* a64 = xalloc(sizeof(double) * m);
  double* gw64 = xalloc(sizeof(double) * m);
  double* gT64 = xalloc(sizeof(double) * m);
  double* ga64 = xalloc(sizeof(double) * m);
  for (size_t i = 0; i < m; ++i) {
    a[i] = urand() * 0.3f;
    if (i % 97 == 5) a[i] = 1.0f; /* opaque sample: 1 - a = 0 in the division */
    gw[i] = urand() - 0.5f;
    a64[i] = a[i];
    gw64[i] = gw[i];
    gT64[i] = 0.0;
  }
  or_patch_weight_fwd(packed, n_patches, P, a, w, T);
  for (int32_t p = 0; p < n_patches; ++p)
    for (int32_t k = 0; k < P; ++k) {
      double s = 0.0, t_end = 1.0;
      for (int32_t j = 0; j < counts[p]; ++j) {
        const size_t idx = ((size_t)packed[2 * p] + j) * P + k;
        CHECK(T[idx] >= 0.f && T[idx] <= 1.f, "transmittance %g out of [0, 1]", T[idx]);
        s += w[idx];
        t_end *= 1.0 - a[idx];
      }
      CHECK(fabs(s - (1.0 - t_end)) < 1e-5, "weights sum %g != 1 - T_end %g", s, 1.0 - t_end);
    }
  or_patch_weight_fwd(packed, n_patches, P, a, w, NULL); /* transmittance not asked for */
  or_patch_weight_bwd_ref(packed, n_patches, P, a, w, gw, ga);
  or_patch_weight_bwd_f64(packed, n_patches, P, a64, gw64, gT64, ga64);
  or_patch_weight_bwd_f64(packed, n_patches, P, a64, gw64, NULL, ga64);
  for (size_t i = 0; i < m; ++i) CHECK(isfinite(ga64[i]), "f64 backward not finite at %zu", i);
  /* visibility over the rays of the same samples (P = 1) */
  uint8_t* keep = xalloc((size_t)n);
  or_visibility(packed, n_patches, a, 1e-4f, 0.f, keep);
  or_visibility(packed, n_patches, a, 1e-4f, 0.1f, keep);
  /* the zero-size batch */
  or_patch_weight_fwd(packed, 0, P, a, w, T);
  or_visibility(packed, 0, a, 1e-4f, 0.f, keep);
  free(keep);
  free(ga64);
  free(gT64);
  free(gw64);
  free(a64);
  free(ga);
  free(gw);
  free(T);
  free(w);
  free(a);
  free(packed);
}

static void check_march(void) {
  const int32_t n_rays = 64, rx = 12, ry = 8, rz = 16;
  const float roi[6] = {-1.f, -1.f, -1.f, 1.f, 1.f, 1.f};
  uint8_t* grid = xalloc((size_t)rx * ry * rz);
  for (int32_t c = 0; c < rx * ry * rz; ++c) grid[c] = urand() < 0.4f;
  float* o = xalloc(sizeof(float) * 3 * n_rays);
  float* d = xalloc(sizeof(float) * 3 * n_rays);
  float* tmin = xalloc(sizeof(float) * n_rays);
  float* tmax = xalloc(sizeof(float) * n_rays);
  for (int32_t i = 0; i < n_rays; ++i) {
    float v[3], nrm = 0.f;
    for (int k = 0; k < 3; ++k) {
      o[3 * i + k] = urand() * 0.4f - 0.2f;
      v[k] = urand() * 2.f - 1.f;
      nrm += v[k] * v[k];
    }
    nrm = sqrtf(nrm);
    for (int k = 0; k < 3; ++k) d[3 * i + k] = v[k] / nrm;
    if (i % 8 == 1) { /* axis-aligned: two zero components */
      d[3 * i] = 1.f;
      d[3 * i + 1] = 0.f;
      d[3 * i + 2] = 0.f;
    }
    tmin[i] = 0.f;
    tmax[i] = (i % 8 == 3) ? 0.f : 2.f; /* some empty rays */
  }
  int32_t* num = xalloc(sizeof(int32_t) * n_rays);
  for (int pass = 0; pass < 2; ++pass) {
    const int32_t cap = pass ? 7 : 0; /* no cap / a per-ray cap */
    or_ray_march_count(o, d, tmin, tmax, n_rays, roi, grid, rx, ry, rz, 0.02f, 0.f, cap, num);
    int64_t n;
    int32_t* packed = packed_of(num, n_rays, &n);
    int64_t* ri = xalloc(sizeof(int64_t) * (size_t)n);
    float* t0 = xalloc(sizeof(float) * (size_t)n);
    float* t1 = xalloc(sizeof(float) * (size_t)n);
    or_ray_march_fill(o, d, tmin, tmax, n_rays, roi, grid, rx, ry, rz, 0.02f, 0.f, cap, packed,
                      ri, t0, t1);
    CHECK(n > 10 * n_rays / (cap ? 4 : 1), "only %lld samples marched", (long long)n);
    for (int32_t i = 0; i < n_rays; ++i) {
      if (cap) CHECK(num[i] <= cap, "ray %d: %d samples above the cap %d", i, num[i], cap);
      if (i % 8 == 3) CHECK(num[i] == 0, "empty ray %d has %d samples", i, num[i]);
      for (int32_t j = 0; j < num[i]; ++j) {
        const int64_t s = (int64_t)packed[2 * i] + j;
        CHECK(ri[s] == i, "sample %lld has ray %lld", (long long)s, (long long)ri[s]);
        CHECK(t1[s] > t0[s], "sample %lld: t1 %g <= t0 %g", (long long)s, t1[s], t0[s]);
        if (j) CHECK(t0[s] >= t1[s - 1], "samples of ray %d overlap", i);
      }
    }
    free(t1);
    free(t0);
    free(ri);
    free(packed);
  }
  free(num);
  free(tmax);
  free(tmin);
  free(d);
  free(o);
  free(grid);
}

static void check_occupancy(void) {
  const int32_t rx = 6, ry = 5, rz = 7, n_cells = rx * ry * rz, n = 150;
  const float roi[6] = {-1.f, -0.5f, -2.f, 1.f, 0.5f, 2.f};
  int64_t* idx = xalloc(sizeof(int64_t) * n);
  float* jit = xalloc(sizeof(float) * 3 * n);
  float* x = xalloc(sizeof(float) * 3 * n);
  float* occ = xalloc(sizeof(float) * n);
  float* occs = xalloc(sizeof(float) * n_cells);
  uint8_t* binary = xalloc((size_t)n_cells);
  for (int32_t i = 0; i < n; ++i) {
    idx[i] = (int64_t)(urand() * (n_cells + 6)) - 3; /* some outside the grid */
    for (int k = 0; k < 3; ++k) jit[3 * i + k] = urand();
    occ[i] = (i == 17) ? NAN : urand();
  }
  for (int32_t c = 0; c < n_cells; ++c) occs[c] = urand() * 0.1f;
  /* cell points only for cells inside the grid, as the caller guarantees */
  int64_t* inside = xalloc(sizeof(int64_t) * n);
  int32_t m = 0;
  for (int32_t i = 0; i < n; ++i)
    if (idx[i] >= 0 && idx[i] < n_cells) inside[m++] = idx[i];
  or_occ_cell_points(inside, m, jit, rx, ry, rz, roi, x);
  for (int32_t i = 0; i < m; ++i)
    for (int k = 0; k < 3; ++k)
      CHECK(x[3 * i + k] >= roi[k] && x[3 * i + k] <= roi[3 + k], "cell point outside the box");
  or_occ_ema_max(occs, n_cells, idx, occ, n, 0.95f);
  int n_nan = 0;
  for (int32_t c = 0; c < n_cells; ++c) n_nan += occs[c] != occs[c];
  CHECK(n_nan <= 1, "%d NaN cells from one NaN occupancy", n_nan);
  for (int32_t c = 0; c < n_cells; ++c)
    if (occs[c] != occs[c]) occs[c] = 0.f;
  const float t = or_occ_binarize(occs, n_cells, 0.01f, binary);
  CHECK(t <= 0.01f, "threshold %g above the given one", t);
  for (int32_t c = 0; c < n_cells; ++c) CHECK(binary[c] == (occs[c] > t), "binarize cell %d", c);
  free(inside);
  free(binary);
  free(occs);
  free(occ);
  free(x);
  free(jit);
  free(idx);
}

static void check_hash_corners(void) {
  const int32_t L = 4;
  const int64_t n = 50;
  const float scale[4] = {3.f, 7.5f, 31.f, 255.f};
  const uint32_t res[4] = {5, 10, 33, 257};
  const uint32_t hsize[4] = {128, 1000, 4096, 4096};
  const uint8_t dense[4] = {1, 1, 0, 0};
  float* x = xalloc(sizeof(float) * 3 * n);
  for (int64_t i = 0; i < 3 * n; ++i) x[i] = urand();
  x[0] = 0.f;
  x[1] = 1.f; /* the box's faces */
  uint32_t* idx = xalloc(sizeof(uint32_t) * n * L * 8);
  float* frac = xalloc(sizeof(float) * n * L * 3);
  float* pos = xalloc(sizeof(float) * n * L * 3);
  or_hashgrid_corners(x, n, L, scale, res, hsize, dense, idx, frac, pos);
  or_hashgrid_corners(x, n, L, scale, res, hsize, dense, idx, frac, NULL);
  for (int64_t i = 0; i < n * L * 8; ++i)
    CHECK(idx[i] < hsize[(i / 8) % L], "corner index %u outside level %lld", idx[i],
          (long long)((i / 8) % L));
  for (int64_t i = 0; i < n * L * 3; ++i)
    CHECK(frac[i] >= 0.f && frac[i] < 1.f, "fraction %g outside [0, 1)", frac[i]);
  free(pos);
  free(frac);
  free(idx);
  free(x);
}

int main(void) {
  check_pack_info();
  check_patch_render();
  check_march();
  check_occupancy();
  check_hash_corners();
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("oracle sanitize check: all functions ran clean\n");
  return 0;
}
