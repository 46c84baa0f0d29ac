/*
 * lbm_cpu_abi.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU implementation of the lbm_* entry points of include/lbcuda_amd.h on
 * top of the FP64 oracle, so the Fortran host program (host/) can be linked
 * and run end to end in a container without a GPU.  Only tests/ link it; the
 * product host links lbcuda_amd/_lib/liblbcuda_amd.so.
 */
#include <cstdio>
#include <cstring>
#include <vector>

#include "../include/lbcuda_amd.h"
#include "lbm_oracle.h"

struct lbm_ctx {
  orc_ctx* o;
  lbm_config cfg;
  size_t N;
  int np;
};

extern "C" {

void lbm_config_default(lbm_config* cfg, int32_t n) {
  std::memset(cfg, 0, sizeof *cfg);
  cfg->nx = cfg->ny = cfg->nz = n;
  cfg->tau_r = cfg->tau_b = 1.0f;
  cfg->sigma = 0.1f;
  cfg->beta = 0.999f;
  cfg->store_rho0 = 0.5f;
  cfg->min_rho = -10.0f;
  cfg->max_rho = 10.0f;
  cfg->radius = 5.5f;
  cfg->rmass = 1.0f / 1500.0f;
  cfg->wetting = LBM_WET_JANUS;
  cfg->hertz = 1;
  cfg->hz_k = 10.0f; cfg->hz_rmin = 10.5f; cfg->hz_rcap = 9.0f;
  cfg->lub_k = 10.0f; cfg->lub_rmin = 10.0f; cfg->lub_rcap = 9.0f;
  cfg->dens_r = cfg->dens_b = 1.0f;
}

int lbm_create(const lbm_config* cfg, lbm_ctx** out) {
  lbm_ctx* c = new lbm_ctx;
  c->cfg = *cfg;
  c->N = (size_t)cfg->nx * cfg->ny * cfg->nz;
  c->np = 0;
  int rc = orc_create(cfg, 64, &c->o);
  if (rc) { delete c; return rc; }
  *out = c;
  return 0;
}

int lbm_destroy(lbm_ctx* c) {
  if (c) { orc_destroy(c->o); delete c; }
  return 0;
}

int lbm_init_fields(lbm_ctx* c, int32_t t, const float u0[3], const float* rr, const float* rb, const float* vel) {
  return orc_init_fields(c->o, t, u0, rr, rb, vel);
}

int lbm_set_pops(lbm_ctx* c, const float* pr, const float* pb) { return orc_set_pops(c->o, pr, pb); }

int lbm_set_particles(lbm_ctx* c, int32_t n, const double* pos) {
  c->np = n;
  return orc_set_particles(c->o, n, pos);
}

int lbm_step(lbm_ctx* c, int32_t n) { return orc_step(c->o, n); }

/* moments2Fl (write_output.f90:975-1041): colour densities, velocity and
 * phase field on fluid nodes with positive total density, zeros elsewhere. */
int lbm_get_moments(lbm_ctx* c, float* rr, float* rb, float* vel, float* ph) {
  std::vector<double> pr(c->N * 19), pb(c->N * 19);
  orc_get_pops(c->o, 0, pr.data());
  orc_get_pops(c->o, 1, pb.data());
  std::vector<int8_t> fl(c->N);
  orc_get_flags(c->o, fl.data());
  static const int ex[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0};
  static const int ey[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1};
  static const int ez[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
  for (size_t n = 0; n < c->N; ++n) {
    double a = 0, b = 0, j[3] = {0, 0, 0};
    for (int l = 0; l < 19; ++l) {
      const double r = pr[n + c->N * l], q = pb[n + c->N * l];
      a += r;
      b += q;
      j[0] += (r + q) * ex[l];
      j[1] += (r + q) * ey[l];
      j[2] += (r + q) * ez[l];
    }
    const bool ok = fl[n] == 1 && a + b > 0.0;  // 1 = fluid node (lbm_get_flags)
    if (rr) rr[n] = ok ? (float)a : 0.0f;
    if (rb) rb[n] = ok ? (float)b : 0.0f;
    if (ph) ph[n] = ok ? (float)((a - b) / (a + b)) : 0.0f;
    if (vel)
      for (int d = 0; d < 3; ++d) vel[3 * n + d] = ok ? (float)(j[d] / (a + b)) : 0.0f;
  }
  return 0;
}

int lbm_get_pops(lbm_ctx* c, int32_t color, float* out) {
  if (!c || !out || color < 0 || color > 1) return LBM_ERR_ARG;
  std::vector<double> p(c->N * 19);
  int rc = orc_get_pops(c->o, color, p.data());
  for (size_t i = 0; i < p.size(); ++i) out[i] = (float)p[i];
  return rc;
}

int lbm_get_rho(lbm_ctx* c, float* rr, float* rb) {
  if (!c) return LBM_ERR_ARG;
  std::vector<double> a(c->N), b(c->N);
  int rc = orc_get_rho(c->o, a.data(), b.data());
  for (size_t n = 0; n < c->N; ++n) {
    if (rr) rr[n] = (float)a[n];
    if (rb) rb[n] = (float)b[n];
  }
  return rc;
}

int lbm_get_flags(lbm_ctx* c, int8_t* out) {
  if (!c || !out) return LBM_ERR_ARG;
  return orc_get_flags(c->o, out);
}

int lbm_get_particles(lbm_ctx* c, double* pos, double* force) { return orc_get_particles(c->o, pos, force); }

/* The CPU test implementation keeps no checkpoint image: size reports
 * LBM_ERR_ARG and callers fall back to the restart files. */
int lbm_checkpoint_size(lbm_ctx*, int64_t*) { return LBM_ERR_ARG; }
int lbm_checkpoint_save(lbm_ctx*, void*, int64_t) { return LBM_ERR_ARG; }
int lbm_checkpoint_load(lbm_ctx*, const void*, int64_t) { return LBM_ERR_ARG; }

int lbm_get_sums(lbm_ctx* c, double out[5]) {
  std::vector<double> pr(c->N * 19), pb(c->N * 19);
  orc_get_pops(c->o, 0, pr.data());
  orc_get_pops(c->o, 1, pb.data());
  std::vector<int8_t> fl(c->N);
  orc_get_flags(c->o, fl.data());
  for (int q = 0; q < 5; ++q) out[q] = 0.0;
  for (size_t n = 0; n < c->N; ++n) {
    if (fl[n] != 1) continue;
    for (int l = 0; l < 19; ++l) {
      out[0] += pr[n + c->N * l];
      out[1] += pb[n + c->N * l];
    }
  }
  return 0;
}

int lbm_status(lbm_ctx* c, int32_t* code, int32_t* step, char* msg, int32_t len) {
  int cd;
  long st;
  orc_status(c->o, &cd, &st, msg, len);
  if (code) *code = cd;
  if (step) *step = (int32_t)st;
  return 0;
}

int lbm_fill_random_rho(uint64_t seed, double mean, double amp, int32_t nx, int32_t ny, int32_t nz,
                        float* rho_r, float* rho_b) {
  const uint64_t n = (uint64_t)nx * (uint64_t)ny * (uint64_t)nz;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t z = (seed ^ i) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    rho_r[i] = (float)(mean + amp * u);
    rho_b[i] = 1.0f - rho_r[i];
  }
  return 0;
}

/* The CPU implementation has a single collide-stream path and no tiles:
 * the kernel-selection calls accept the request and report tile 0. */
int lbm_set_kernel(lbm_ctx* c, int32_t variant, int32_t, int32_t) {
  return (c && (variant == LBM_KERNEL_FUSED || variant == LBM_KERNEL_TWOPASS)) ? 0 : LBM_ERR_ARG;
}

int lbm_autotune(lbm_ctx* c, int32_t reps, int32_t* tile, int32_t* zchunk, double* best_ms) {
  if (!c || reps <= 0) return LBM_ERR_ARG;
  if (tile) *tile = 0;
  if (zchunk) *zchunk = 1;
  if (best_ms) *best_ms = 0.0;
  return 0;
}

}  /* extern "C" */
