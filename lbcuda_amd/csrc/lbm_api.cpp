/*
 * lbm_api.cpp -- host side of the C ABI declared in include/lbcuda_amd.h.
 *
 * Owns the device memory of one box and sequences the kernels of one time
 * step in the order of the reference step loop (main.CUF:127-278):
 *   [particles] flags from positions, erase (init_isfluid_CG, del_fluid_CG)
 *   densities + phase field              (init_rhoB_CG, init_rhoR_CG)
 *   [particles] fill, wetting            (make_fluid_CG, compute_*_rho)
 *   collide + stream (+ bounce-back copy) (time_step_CG ... stream_CG_z)
 *   [particles] link forces + moving wall, pair forces, rigid-body update
 *                                        (partBB_CG*, inter_force_particles,
 *                                         time_step_force_particles)
 * All launches go to one stream; the host synchronises only when it reads
 * something back or polls the error word at the end of lbm_step.
 */
#include "lbm_common.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

int lbm_particles_upload_belt2();

namespace {

int hip_ok(hipError_t e) { return e == hipSuccess ? 0 : LBM_ERR_HIP; }

inline void dfree(void* p) {
  if (p) (void)hipFree(p);
}

void set_host_err(lbm_ctx* c, int code, const char* msg) {
  if (c->host_err == 0) {
    c->host_err = code;
    c->host_err_step = c->step;
    std::snprintf(c->host_msg, sizeof c->host_msg, "%s", msg);
  }
}

template <class T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return 0;
  return hipMalloc((void**)p, sizeof(T) * count) == hipSuccess ? 0 : LBM_ERR_ALLOC;
}

/* spherical_template (main.CUF:1990-2130): non-fluid offsets of one particle */
void build_template(float rdim, int* rmax_out, std::vector<int8_t>& tmpl) {
  const float sqrcut = rdim * rdim;
  const int rmin = (int)std::floor(rdim);
  const int rmax = (int)(rdim + 1.0f);
  const int w = 2 * rmax + 1;
  std::vector<int8_t> s((size_t)w * w * w, lbm::FL_FLUID);
  auto S = [&](int i, int j, int k) -> int8_t& { return s[(size_t)((i + rmax) + w * ((j + rmax) + w * (k + rmax)))]; };
  for (int k = -rmax; k <= rmax; ++k)
    for (int j = -rmax; j <= rmax; ++j)
      for (int i = -rmax; i <= rmax; ++i)
        if ((float)(i * i + j * j + k * k) <= sqrcut) S(i, j, k) = lbm::FL_DEAD;
  for (int k = -rmin; k <= rmin; ++k)
    for (int j = -rmin; j <= rmin; ++j)
      for (int i = -rmin; i <= rmin; ++i)
        if (S(i, j, k) == lbm::FL_DEAD)
          for (int l = 1; l < 19; ++l)
            if (S(i + lbm::EX(l), j + lbm::EY(l), k + lbm::EZ(l)) == lbm::FL_FLUID) {
              S(i, j, k) = lbm::FL_LIST;
              break;
            }
  S(0, 0, 0) = lbm::FL_CM;
  tmpl.clear();
  for (int k = -rmax; k <= rmax; ++k)
    for (int j = -rmax; j <= rmax; ++j)
      for (int i = -rmax; i <= rmax; ++i)
        if (S(i, j, k) != lbm::FL_FLUID) {
          tmpl.push_back((int8_t)i); tmpl.push_back((int8_t)j); tmpl.push_back((int8_t)k);
          tmpl.push_back(S(i, j, k));
        }
  *rmax_out = rmax;
}

/* offsets with i^2+j^2+k^2 <= rmax^2 (the distance tests of del_fluid,
 * make_fluid: <= sphereMax; partBB_CG, partBB_CG_p2n2: < sphereMax) */
void build_ball(int rmax, std::vector<int8_t>& ball) {
  ball.clear();
  const int smax = rmax * rmax;
  for (int k = -rmax; k <= rmax; ++k)
    for (int j = -rmax; j <= rmax; ++j)
      for (int i = -rmax; i <= rmax; ++i) {
        int d = i * i + j * j + k * k;
        if (d <= smax) {
          ball.push_back((int8_t)i); ball.push_back((int8_t)j); ball.push_back((int8_t)k);
          ball.push_back((int8_t)(d < smax ? 1 : 0));
        }
      }
}

/* D3Q27 ring of the template's surface nodes (offsets outside the template):
 * the fluid nodes whose densities the particle kernels read; z, y, x order
 * so consecutive threads read consecutive x */
void build_ring(const std::vector<int8_t>& tmpl, std::vector<int8_t>& ring) {
  std::vector<int> seen;
  const size_t nt = tmpl.size() / 4;
  auto key = [](int i, int j, int k) { return (i + 32) + 64 * ((j + 32) + 64 * (k + 32)); };
  for (size_t t = 0; t < nt; ++t) seen.push_back(key(tmpl[4 * t], tmpl[4 * t + 1], tmpl[4 * t + 2]));
  std::vector<std::array<int, 3>> out;
  for (size_t t = 0; t < nt; ++t) {
    if (tmpl[4 * t + 3] != lbm::FL_LIST) continue;
    for (int k = -1; k <= 1; ++k)
      for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) {
          if (!(i || j || k)) continue;
          const int a = tmpl[4 * t] + i, b = tmpl[4 * t + 1] + j, cc = tmpl[4 * t + 2] + k;
          const int q = key(a, b, cc);
          if (std::find(seen.begin(), seen.end(), q) != seen.end()) continue;
          seen.push_back(q);
          out.push_back({a, b, cc});
        }
  }
  std::sort(out.begin(), out.end(), [](const std::array<int, 3>& u, const std::array<int, 3>& v) {
    if (u[2] != v[2]) return u[2] < v[2];
    if (u[1] != v[1]) return u[1] < v[1];
    return u[0] < v[0];
  });
  ring.clear();
  for (auto& e : out) {
    ring.push_back((int8_t)e[0]); ring.push_back((int8_t)e[1]); ring.push_back((int8_t)e[2]); ring.push_back(0);
  }
}

/* default tile for the fused kernel: the whole-row tile when one fits */
void default_kernel(lbm_ctx* c) {
  c->variant = c->cfg.diffdens ? LBM_KERNEL_TWOPASS : LBM_KERNEL_FUSED;   /* the density-ratio model is
                                                                             two-kernel only */
  c->tile = 0;
  const int prefer[] = {5, 7, 9};   /* 256x2, 128x4, 64x4 with x wrap */
  for (int t : prefer)
    if (lbm_fused_tile_valid(c, t, 1)) { c->tile = t; break; }
  c->zchunk = c->nz < 32 ? c->nz : 32;
}

int restamp_initial(lbm_ctx* c) {
  /* step-0 node flags (setupRho_Pops, main.CUF:2596-2605): the current level
     from the positions, the other level stays all fluid. */
  int rc = 0;
  for (int b = 0; b < 2; ++b) {
    rc |= hip_ok(hipMemsetAsync(c->d_key[b], 0xff, sizeof(uint32_t) * c->N, c->stream));
    std::vector<int> inv((size_t)3 * c->np, INT32_MIN);
    rc |= hip_ok(hipMemcpyAsync(c->d_stamp[b], inv.data(), sizeof(int) * inv.size(), hipMemcpyHostToDevice, c->stream));
    rc |= hip_ok(hipStreamSynchronize(c->stream));
  }
  if (rc) return LBM_ERR_HIP;
  return lbm_particles_flags(c, c->cur);
}

}  // namespace

extern "C" {

void lbm_config_default(lbm_config* cfg, int32_t n) {
  std::memset(cfg, 0, sizeof *cfg);
  cfg->nx = cfg->ny = cfg->nz = n;
  cfg->tau_r = cfg->tau_b = 1.0f;
  cfg->sigma = 0.1f;
  cfg->beta = 0.999f;
  cfg->storage = LBM_STORE_F32;
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

int lbm_fill_random_rho(uint64_t seed, double mean, double amp, int32_t nx, int32_t ny, int32_t nz,
                        float* rho_r, float* rho_b) {
  if (!rho_r || !rho_b || nx <= 0 || ny <= 0 || nz <= 0) return LBM_ERR_ARG;
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
  return LBM_OK;
}

const char* lbm_version(void) {
  return "lbcuda_amd 0.1 gfx950 (storage: f32, f16-shifted, bf16-shifted)";
}

int lbm_create(const lbm_config* cfg, lbm_ctx** out) {
  if (!cfg || !out) return LBM_ERR_ARG;
  *out = nullptr;
  if (cfg->nx < 3 || cfg->ny < 3 || cfg->nz < 3) return LBM_ERR_ARG;
  if (cfg->tau_r <= 0.5f || cfg->tau_b <= 0.5f) return LBM_ERR_ARG;
  if (cfg->storage < LBM_STORE_F32 || cfg->storage > LBM_STORE_BF16) return LBM_ERR_ARG;
  /* shifted 16-bit storage: the fused kernel re-stores non-fluid values
     through float, exact while w_0 rho0 + |f - w rho0| < 0.5 */
  if (cfg->storage != LBM_STORE_F32 && !(cfg->store_rho0 >= 0.0f && cfg->store_rho0 <= 1.2f)) return LBM_ERR_ARG;
  /* DIFFDENS: "the density of at least one of the two fluids should be one"
     (SetConstants, main.CUF:1058-1062) */
  if (cfg->diffdens && !(cfg->dens_r > 0.0f && cfg->dens_b > 0.0f && (cfg->dens_r == 1.0f || cfg->dens_b == 1.0f)))
    return LBM_ERR_ARG;
  /* template offsets are int8: radius + 3 must stay below 127 */
  if (cfg->with_particles && (cfg->radius < 1.0f || cfg->radius > 60.0f)) return LBM_ERR_ARG;
  /* particle kernels wrap offsets of up to rmax + 2 nodes once per axis */
  if (cfg->with_particles) {
    const int rmx = (int)(cfg->radius + 1.0f) + 3;
    if (cfg->nx < rmx || cfg->ny < rmx || cfg->nz < rmx) return LBM_ERR_ARG;
  }
  /* 32-bit node offsets inside the kernels */
  if ((double)cfg->nx * cfg->ny * cfg->nz > (double)(1u << 30)) return LBM_ERR_ARG;
  /* the fused kernel addresses a time level through buffer resources over
     groups of four direction planes with 32-bit offsets: four planes of N
     population pairs (2 * esize bytes each) must fit in 4 GiB, i.e. up to
     512^3 nodes with float storage and 645^3 with 16-bit storage */
  if ((double)cfg->nx * cfg->ny * cfg->nz * (cfg->storage == LBM_STORE_F32 ? 8.0 : 4.0) > 1073741824.0)
    return LBM_ERR_ARG;
  lbm_ctx* c = new (std::nothrow) lbm_ctx;
  if (!c) return LBM_ERR_ALLOC;
  std::memset((void*)c, 0, sizeof *c);
  c->cfg = *cfg;
  c->nx = cfg->nx; c->ny = cfg->ny; c->nz = cfg->nz;
  c->N = (size_t)c->nx * c->ny * c->nz;
  c->storage = cfg->storage;
  c->esize = cfg->storage == LBM_STORE_F32 ? 4 : 2;
  default_kernel(c);
  int rc = 0;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return LBM_ERR_HIP; }
  for (int b = 0; b < 2; ++b) {
    rc |= hipMalloc(&c->pop[b], c->esize * 38 * c->N) == hipSuccess ? 0 : LBM_ERR_ALLOC;
    rc |= dalloc(&c->flags[b], c->N);
  }
  rc |= dalloc(&c->rho_r, c->N);
  rc |= dalloc(&c->rho_b, c->N);
  rc |= dalloc(&c->phi, c->N);
  rc |= dalloc(&c->d_status, 1);
  if (!rc) rc |= hip_ok(hipMemset(c->d_status, 0, sizeof(lbm::DevStatus)));
  if (!rc && cfg->with_particles) {
    std::vector<int8_t> tmpl, ball, ring, tlist;
    build_template(cfg->radius, &c->rmax, tmpl);
    build_ball(c->rmax, ball);
    build_ring(tmpl, ring);
    for (size_t t = 0; t < tmpl.size() / 4; ++t)
      if (tmpl[4 * t + 3] == lbm::FL_LIST) tlist.insert(tlist.end(), &tmpl[4 * t], &tmpl[4 * t] + 4);
    c->ntmpl = (int)(tmpl.size() / 4);
    c->nball = (int)(ball.size() / 4);
    c->nring = (int)(ring.size() / 4);
    c->ntlist = (int)(tlist.size() / 4);
    /* the boxed shell pass indexes a (2 rmax + 1)^3 LDS box by these offsets */
    c->shell_box_ok = true;
    for (const std::vector<int8_t>* tb : {&tmpl, &ring})
      for (size_t t = 0; t < tb->size() / 4; ++t)
        for (int d = 0; d < 3; ++d)
          if (std::abs((int)(*tb)[4 * t + d]) > c->rmax) c->shell_box_ok = false;
    rc |= dalloc(&c->d_tmpl, tmpl.size());
    rc |= dalloc(&c->d_ring, ring.size());
    rc |= dalloc(&c->d_tlist, tlist.size());
    rc |= dalloc(&c->d_ball, ball.size());
    for (int b = 0; b < 2; ++b) rc |= dalloc(&c->d_key[b], c->N);
    rc |= dalloc(&c->d_claim, c->N);
    if (!rc) {
      rc |= hip_ok(hipMemcpy(c->d_tmpl, tmpl.data(), tmpl.size(), hipMemcpyHostToDevice));
      rc |= hip_ok(hipMemcpy(c->d_ball, ball.data(), ball.size(), hipMemcpyHostToDevice));
      rc |= hip_ok(hipMemcpy(c->d_ring, ring.data(), ring.size(), hipMemcpyHostToDevice));
      rc |= hip_ok(hipMemcpy(c->d_tlist, tlist.data(), tlist.size(), hipMemcpyHostToDevice));
      for (int b = 0; b < 2; ++b) rc |= hip_ok(hipMemset(c->d_key[b], 0xff, sizeof(uint32_t) * c->N));
      rc |= hip_ok(hipMemset(c->d_claim, 0xff, sizeof(int) * c->N));
      rc |= lbm_particles_upload_belt2();
    }
  }
  if (rc) {
    lbm_destroy(c);
    return rc & LBM_ERR_ALLOC ? LBM_ERR_ALLOC : LBM_ERR_HIP;
  }
  *out = c;
  return LBM_OK;
}

int lbm_destroy(lbm_ctx* c) {
  if (!c) return LBM_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int b = 0; b < 2; ++b) {
    dfree(c->pop[b]);
    dfree(c->flags[b]);
    dfree(c->d_pos[b]);
    dfree(c->d_force[b]);
    dfree(c->d_stamp[b]);
    dfree(c->d_key[b]);
  }
  dfree(c->rho_r); dfree(c->rho_b); dfree(c->phi);
  dfree(c->d_status);
  dfree(c->d_myf); dfree(c->d_myf2); dfree(c->d_myt); dfree(c->d_myt2); dfree(c->d_restamp);
  dfree(c->d_claim); dfree(c->d_tmpl); dfree(c->d_ball); dfree(c->d_ring); dfree(c->d_tlist);
  dfree(c->d_nbr); dfree(c->d_nnbr); dfree(c->d_contrib); dfree(c->d_smask); dfree(c->d_closure);
  for (int e = 0; e < 2 * c->ev_cap; ++e) (void)hipEventDestroy(c->ev_pairs[e]);
  delete[] c->ev_pairs;
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return LBM_OK;
}

int lbm_init_fields(lbm_ctx* c, int32_t type, const float u0[3], const float* rho_r, const float* rho_b,
                    const float* vel) {
  if (!c) return LBM_ERR_ARG;
  if (type < LBM_INIT_DROPLET || type > LBM_INIT_VELFIELD) return LBM_ERR_ARG;
  const bool host_rho = type == LBM_INIT_RHO || type == LBM_INIT_RANDOM || type == LBM_INIT_VELFIELD ||
                        type == LBM_INIT_POPS;
  if (host_rho && (!rho_r || !rho_b)) return LBM_ERR_ARG;
  if (type == LBM_INIT_VELFIELD && !vel) return LBM_ERR_ARG;
  const float zero[3] = {0.0f, 0.0f, 0.0f};
  if (!u0) u0 = zero;
  int rc = 0;
  c->cur = 0;
  c->step = 0;
  rc |= hip_ok(hipMemsetAsync(c->d_status, 0, sizeof(lbm::DevStatus), c->stream));
  rc |= hip_ok(hipMemsetAsync(c->rho_r, 0, sizeof(float) * c->N, c->stream));
  rc |= hip_ok(hipMemsetAsync(c->rho_b, 0, sizeof(float) * c->N, c->stream));
  if (host_rho) {
    rc |= hip_ok(hipMemcpyAsync(c->rho_r, rho_r, sizeof(float) * c->N, hipMemcpyHostToDevice, c->stream));
    rc |= hip_ok(hipMemcpyAsync(c->rho_b, rho_b, sizeof(float) * c->N, hipMemcpyHostToDevice, c->stream));
  }
  float* d_vel = nullptr;
  if (type == LBM_INIT_VELFIELD) {
    rc |= dalloc(&d_vel, 3 * c->N);
    if (!rc) rc |= hip_ok(hipMemcpyAsync(d_vel, vel, sizeof(float) * 3 * c->N, hipMemcpyHostToDevice, c->stream));
  }
  if (!rc) rc |= lbm_launch_init_eq(c, type, u0, d_vel);
  rc |= hip_ok(hipStreamSynchronize(c->stream));
  dfree(d_vel);
  if (rc) return rc;
  c->host_err = 0;
  c->initialised = true;
  if (c->np > 0) rc = restamp_initial(c);
  if (!rc) rc = hip_ok(hipStreamSynchronize(c->stream));
  return rc;
}

int lbm_set_pops(lbm_ctx* c, const float* pr, const float* pb) {
  if (!c || !pr || !pb) return LBM_ERR_ARG;
  float *dr = nullptr, *db = nullptr;
  int rc = dalloc(&dr, 19 * c->N) | dalloc(&db, 19 * c->N);
  if (!rc) {
    rc |= hip_ok(hipMemcpyAsync(dr, pr, sizeof(float) * 19 * c->N, hipMemcpyHostToDevice, c->stream));
    rc |= hip_ok(hipMemcpyAsync(db, pb, sizeof(float) * 19 * c->N, hipMemcpyHostToDevice, c->stream));
    if (!rc) rc |= lbm_launch_pack_pops(c, dr, db);
    rc |= hip_ok(hipStreamSynchronize(c->stream));
  }
  dfree(dr);
  dfree(db);
  return rc;
}

int lbm_set_particles(lbm_ctx* c, int32_t n, const double* pos) {
  if (!c || !pos || n <= 0) return LBM_ERR_ARG;
  if (!c->cfg.with_particles) return LBM_ERR_ARG;
  if (n > (1 << 28)) return LBM_ERR_ARG;
  /* SetupParts sanity checks, main.CUF:1204-1222 */
  const double R = c->cfg.radius;
  if (c->nx < 2 * R || c->ny < 2 * R || c->nz < 2 * R) return LBM_ERR_PART_SETUP;
  for (int a = 0; a < n; ++a) {
    const double* p = pos + 13 * (size_t)a;
    if (p[0] < 1 || p[0] > c->nx || p[1] < 1 || p[1] > c->ny || p[2] < 1 || p[2] > c->nz) return LBM_ERR_PART_SETUP;
    for (int b = a + 1; b < n; ++b) {
      const double* q = pos + 13 * (size_t)b;
      double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
      if (std::sqrt(dx * dx + dy * dy + dz * dz) < 2 * R + 1) return LBM_ERR_PART_SETUP;
    }
  }
  if (n != c->np) {
    for (int b = 0; b < 2; ++b) {
      dfree(c->d_pos[b]); dfree(c->d_force[b]); dfree(c->d_stamp[b]);
      c->d_pos[b] = nullptr; c->d_force[b] = nullptr; c->d_stamp[b] = nullptr;
    }
    dfree(c->d_restamp);
    c->d_restamp = nullptr;
    dfree(c->d_myf); dfree(c->d_myf2); dfree(c->d_myt); dfree(c->d_myt2);
    c->d_myf = c->d_myf2 = c->d_myt = c->d_myt2 = nullptr;
    int rc = 0;

    dfree(c->d_nbr); dfree(c->d_nnbr); dfree(c->d_contrib); dfree(c->d_smask);
    c->d_nbr = nullptr; c->d_nnbr = nullptr; c->d_contrib = nullptr; c->d_smask = nullptr;
    rc |= dalloc(&c->d_nbr, (size_t)lbm_particles_max_neighbours() * n);
    rc |= dalloc(&c->d_contrib, (size_t)6 * lbm_particles_max_neighbours() * n);
    rc |= dalloc(&c->d_nnbr, (size_t)n);
    rc |= dalloc(&c->d_smask, (size_t)c->ntlist * n);
    for (int b = 0; b < 2; ++b) {
      rc |= dalloc(&c->d_pos[b], (size_t)13 * n);
      rc |= dalloc(&c->d_force[b], (size_t)12 * n);
      rc |= dalloc(&c->d_stamp[b], (size_t)3 * n);
    }
    rc |= dalloc(&c->d_restamp, (size_t)n);
    rc |= dalloc(&c->d_myf, (size_t)3 * n);
    rc |= dalloc(&c->d_myf2, (size_t)3 * n);
    rc |= dalloc(&c->d_myt, (size_t)3 * n);
    rc |= dalloc(&c->d_myt2, (size_t)3 * n);
    if (rc) { c->np = 0; return LBM_ERR_ALLOC; }
    c->np = n;
    c->generation += 1;
  }
  /* read_input_atom (main.CUF:1300-1340): both levels start equal; forces
     start at the external force/torque, bounce-back parts at zero */
  const lbm_config& g = c->cfg;
  std::vector<double> f12((size_t)12 * n, 0.0), m3((size_t)3 * n, 0.0), e3((size_t)3 * n), t3((size_t)3 * n);
  for (int a = 0; a < n; ++a)
    for (int d = 0; d < 3; ++d) {
      f12[12 * (size_t)a + d] = g.ext_force[d];
      f12[12 * (size_t)a + 6 + d] = g.lrotate ? g.ext_torque[d] : 0.0;
      e3[3 * (size_t)a + d] = g.ext_force[d];
      t3[3 * (size_t)a + d] = g.lrotate ? g.ext_torque[d] : 0.0;
    }
  int rc = 0;
  for (int b = 0; b < 2; ++b) {
    rc |= hip_ok(hipMemcpy(c->d_pos[b], pos, sizeof(double) * 13 * n, hipMemcpyHostToDevice));
    rc |= hip_ok(hipMemcpy(c->d_force[b], f12.data(), sizeof(double) * f12.size(), hipMemcpyHostToDevice));
  }
  rc |= hip_ok(hipMemcpy(c->d_myf, m3.data(), sizeof(double) * m3.size(), hipMemcpyHostToDevice));
  rc |= hip_ok(hipMemcpy(c->d_myt, m3.data(), sizeof(double) * m3.size(), hipMemcpyHostToDevice));
  rc |= hip_ok(hipMemcpy(c->d_myf2, e3.data(), sizeof(double) * e3.size(), hipMemcpyHostToDevice));
  rc |= hip_ok(hipMemcpy(c->d_myt2, t3.data(), sizeof(double) * t3.size(), hipMemcpyHostToDevice));
  if (rc) return rc;
  if (c->initialised) {
    rc = restamp_initial(c);
    if (!rc) rc = hip_ok(hipStreamSynchronize(c->stream));
  }
  return rc;
}

int lbm_step(lbm_ctx* c, int32_t nsteps) {
  if (!c || nsteps < 0) return LBM_ERR_ARG;
  if (!c->initialised) return LBM_ERR_STATE;
  const bool parts = c->np > 0;
  int rc = 0;
  for (int s = 0; s < nsteps && !rc; ++s) {
    c->step += 1;
    if (parts) {
      rc |= lbm_particles_flags(c, c->cur);
      rc |= lbm_particles_erase(c);
    }
    const bool fused = c->variant == LBM_KERNEL_FUSED;
    if (!fused) rc |= lbm_launch_density(c);
    if (parts) {
      rc |= lbm_particles_fill(c);
      rc |= fused ? lbm_particles_shell_wetting(c) : lbm_particles_wetting(c);
    }
    rc |= fused ? lbm_launch_fused(c) : lbm_launch_collide(c);
    if (parts) {
      rc |= lbm_particles_bounce(c);
      rc |= lbm_particles_integrate(c);
    }
    c->cur = 1 - c->cur;
  }
  if (rc) {
    set_host_err(c, LBM_ERR_HIP, "kernel launch failed");
    return LBM_ERR_HIP;
  }
  lbm::DevStatus st;
  if (hipMemcpyAsync(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    set_host_err(c, LBM_ERR_HIP, "device error while stepping");
    return LBM_ERR_HIP;
  }
  return st.code;
}

int lbm_get_pops(lbm_ctx* c, int32_t color, float* out) {
  if (!c || !out || color < 0 || color > 1) return LBM_ERR_ARG;
  float* d = nullptr;
  int rc = dalloc(&d, 19 * c->N);
  if (!rc) {
    rc |= lbm_launch_unpack_pops(c, color, d);
    rc |= hip_ok(hipMemcpyAsync(out, d, sizeof(float) * 19 * c->N, hipMemcpyDeviceToHost, c->stream));
    rc |= hip_ok(hipStreamSynchronize(c->stream));
  }
  dfree(d);
  return rc;
}

int lbm_get_rho(lbm_ctx* c, float* rr, float* rb) {
  if (!c) return LBM_ERR_ARG;
  int rc = 0;
  if (c->variant == LBM_KERNEL_FUSED && c->step > 0) {
    /* the single-pass step keeps densities only where particle kernels need
       them: form the rest from the level the last step started from */
    float *d_rr = nullptr, *d_rb = nullptr;
    rc |= dalloc(&d_rr, c->N);
    rc |= dalloc(&d_rb, c->N);
    if (!rc) {
      rc |= lbm_launch_rho_from_pops(c, d_rr, d_rb);
      if (rr) rc |= hip_ok(hipMemcpyAsync(rr, d_rr, sizeof(float) * c->N, hipMemcpyDeviceToHost, c->stream));
      if (rb) rc |= hip_ok(hipMemcpyAsync(rb, d_rb, sizeof(float) * c->N, hipMemcpyDeviceToHost, c->stream));
      rc |= hip_ok(hipStreamSynchronize(c->stream));
    }
    dfree(d_rr);
    dfree(d_rb);
    return rc;
  }
  if (rr) rc |= hip_ok(hipMemcpyAsync(rr, c->rho_r, sizeof(float) * c->N, hipMemcpyDeviceToHost, c->stream));
  if (rb) rc |= hip_ok(hipMemcpyAsync(rb, c->rho_b, sizeof(float) * c->N, hipMemcpyDeviceToHost, c->stream));
  rc |= hip_ok(hipStreamSynchronize(c->stream));
  return rc;
}

int lbm_get_moments(lbm_ctx* c, float* rr, float* rb, float* vel, float* ph) {
  if (!c) return LBM_ERR_ARG;
  float *d_rr = nullptr, *d_rb = nullptr, *d_v = nullptr, *d_ph = nullptr;
  int rc = 0;
  if (rr) rc |= dalloc(&d_rr, c->N);
  if (rb) rc |= dalloc(&d_rb, c->N);
  if (vel) rc |= dalloc(&d_v, 3 * c->N);
  if (ph) rc |= dalloc(&d_ph, c->N);
  if (!rc) {
    rc |= lbm_launch_moments(c, d_rr, d_rb, d_v, d_ph);
    if (rr) rc |= hip_ok(hipMemcpyAsync(rr, d_rr, sizeof(float) * c->N, hipMemcpyDeviceToHost, c->stream));
    if (rb) rc |= hip_ok(hipMemcpyAsync(rb, d_rb, sizeof(float) * c->N, hipMemcpyDeviceToHost, c->stream));
    if (vel) rc |= hip_ok(hipMemcpyAsync(vel, d_v, sizeof(float) * 3 * c->N, hipMemcpyDeviceToHost, c->stream));
    if (ph) rc |= hip_ok(hipMemcpyAsync(ph, d_ph, sizeof(float) * c->N, hipMemcpyDeviceToHost, c->stream));
    rc |= hip_ok(hipStreamSynchronize(c->stream));
  }
  dfree(d_rr); dfree(d_rb); dfree(d_v); dfree(d_ph);
  return rc;
}

int lbm_get_flags(lbm_ctx* c, int8_t* out) {
  if (!c || !out) return LBM_ERR_ARG;
  int rc = hip_ok(hipMemcpyAsync(out, c->flags[c->cur], c->N, hipMemcpyDeviceToHost, c->stream));
  rc |= hip_ok(hipStreamSynchronize(c->stream));
  return rc;
}

int lbm_get_particles(lbm_ctx* c, double* pos, double* force) {
  if (!c || c->np <= 0) return LBM_ERR_ARG;
  int rc = 0;
  if (pos)
    rc |= hip_ok(hipMemcpyAsync(pos, c->d_pos[c->cur], sizeof(double) * 13 * c->np, hipMemcpyDeviceToHost, c->stream));
  if (force)
    rc |= hip_ok(hipMemcpyAsync(force, c->d_force[1 - c->cur], sizeof(double) * 12 * c->np, hipMemcpyDeviceToHost,
                                c->stream));
  rc |= hip_ok(hipStreamSynchronize(c->stream));
  return rc;
}

int lbm_get_sums(lbm_ctx* c, double out[5]) {
  if (!c || !out) return LBM_ERR_ARG;
  double* d = nullptr;
  int rc = dalloc(&d, 5);
  if (!rc) {
    rc |= lbm_launch_sums(c, d);
    rc |= hip_ok(hipMemcpyAsync(out, d, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
    rc |= hip_ok(hipStreamSynchronize(c->stream));
  }
  dfree(d);
  return rc;
}

/* ---- checkpoint image --------------------------------------------------
 * Header, then every device buffer that carries state from one step to the
 * next, in a fixed order: the populations of both levels, the flags,
 * densities and phase field, and, with particles, both time levels of the
 * particle state, the force accumulators, the stamped centres, the owner keys
 * and the fill claims.  Everything else a step touches is rebuilt inside it. */
namespace {

struct CkptHeader {
  char magic[8];          /* "LBCKPT1\0"                                   */
  lbm_config cfg;         /* must equal the loading context's config       */
  int32_t np, cur;
  int64_t step;
  int64_t bytes;          /* whole image                                   */
};

struct CkptPart {
  void* dev;
  size_t bytes;
};

void ckpt_parts(lbm_ctx* c, std::vector<CkptPart>& v) {
  v.clear();
  const size_t N = c->N;
  /* both levels: slots that no node pushes into (those next to a sphere's
     interior) keep the value of two steps before, and the surface densities
     read them */
  for (int b = 0; b < 2; ++b) v.push_back({c->pop[b], 2 * 19 * N * c->esize});
  for (int b = 0; b < 2; ++b) v.push_back({c->flags[b], N});
  v.push_back({c->rho_r, sizeof(float) * N});
  v.push_back({c->rho_b, sizeof(float) * N});
  v.push_back({c->phi, sizeof(float) * N});
  if (c->np > 0) {
    const size_t n = (size_t)c->np;
    for (int b = 0; b < 2; ++b) {
      v.push_back({c->d_pos[b], sizeof(double) * 13 * n});
      v.push_back({c->d_force[b], sizeof(double) * 12 * n});
      v.push_back({c->d_stamp[b], sizeof(int) * 3 * n});
      v.push_back({c->d_key[b], sizeof(uint32_t) * N});
    }
    v.push_back({c->d_myf, sizeof(double) * 3 * n});
    v.push_back({c->d_myf2, sizeof(double) * 3 * n});
    v.push_back({c->d_myt, sizeof(double) * 3 * n});
    v.push_back({c->d_myt2, sizeof(double) * 3 * n});
    v.push_back({c->d_claim, sizeof(int) * N});
  }
}

int64_t ckpt_bytes(lbm_ctx* c) {
  int64_t s = sizeof(CkptHeader);
  std::vector<CkptPart> v;
  ckpt_parts(c, v);
  for (const CkptPart& p : v) s += (int64_t)p.bytes;
  return s;
}

}  // namespace

int lbm_checkpoint_size(lbm_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return LBM_ERR_ARG;
  if (!c->initialised) return LBM_ERR_STATE;
  *bytes = ckpt_bytes(c);
  return 0;
}

int lbm_checkpoint_save(lbm_ctx* c, void* buf, int64_t bytes) {
  if (!c || !buf) return LBM_ERR_ARG;
  if (!c->initialised) return LBM_ERR_STATE;
  if (bytes < ckpt_bytes(c)) return LBM_ERR_ARG;
  CkptHeader h;
  std::memset(&h, 0, sizeof h);
  std::memcpy(h.magic, "LBCKPT1", 8);
  h.cfg = c->cfg;
  h.np = c->np;
  h.cur = c->cur;
  h.step = c->step;
  h.bytes = ckpt_bytes(c);
  std::memcpy(buf, &h, sizeof h);
  char* p = static_cast<char*>(buf) + sizeof h;
  int rc = hip_ok(hipStreamSynchronize(c->stream));
  std::vector<CkptPart> parts;
  ckpt_parts(c, parts);
  for (const CkptPart& s : parts) {
    if (rc) break;
    rc |= hip_ok(hipMemcpy(p, s.dev, s.bytes, hipMemcpyDeviceToHost));
    p += s.bytes;
  }
  return rc;
}

int lbm_checkpoint_load(lbm_ctx* c, const void* buf, int64_t bytes) {
  if (!c || !buf || bytes < (int64_t)sizeof(CkptHeader)) return LBM_ERR_ARG;
  CkptHeader h;
  std::memcpy(&h, buf, sizeof h);
  if (std::memcmp(h.magic, "LBCKPT1", 8) != 0) return LBM_ERR_ARG;
  if (std::memcmp(&h.cfg, &c->cfg, sizeof(lbm_config)) != 0) return LBM_ERR_ARG;
  if (h.cur != 0 && h.cur != 1) return LBM_ERR_ARG;
  /* particle buffers exist once lbm_set_particles has run with the same count */
  if (h.np != c->np) return LBM_ERR_STATE;
  const int cur = c->cur;
  c->cur = h.cur;
  const int64_t need = ckpt_bytes(c);
  c->cur = cur;
  if (h.bytes != need || bytes < need) return LBM_ERR_ARG;
  int rc = hip_ok(hipStreamSynchronize(c->stream));
  if (rc) return rc;
  c->cur = h.cur;
  const char* p = static_cast<const char*>(buf) + sizeof h;
  std::vector<CkptPart> parts;
  ckpt_parts(c, parts);
  for (const CkptPart& s : parts) {
    if (rc) break;
    rc |= hip_ok(hipMemcpy(s.dev, p, s.bytes, hipMemcpyHostToDevice));
    p += s.bytes;
  }
  rc |= hip_ok(hipMemsetAsync(c->d_status, 0, sizeof(lbm::DevStatus), c->stream));
  rc |= hip_ok(hipStreamSynchronize(c->stream));
  if (rc) return rc;
  c->step = h.step;
  c->host_err = 0;
  c->initialised = true;
  return 0;
}

int lbm_status(lbm_ctx* c, int32_t* code, int32_t* step, char* msg, int32_t len) {
  if (!c) return LBM_ERR_ARG;
  lbm::DevStatus st{0, 0};
  int rc = hip_ok(hipMemcpyAsync(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost, c->stream));
  rc |= hip_ok(hipStreamSynchronize(c->stream));
  int cd = c->host_err ? c->host_err : st.code;
  int sp = c->host_err ? (int)c->host_err_step : st.step;
  if (code) *code = cd;
  if (step) *step = sp;
  if (msg && len > 0) {
    const char* m = "";
    if (c->host_err) m = c->host_msg;
    else if (st.code == LBM_ERR_RHO_RANGE) m = "init_rho_CG] Range error rho (kernels_fluid_CG.CUF:438)";
    else if (st.code == LBM_ERR_BIG_FORCE) m = "nve_lf] Big force (kernels_fluid_PART.CUF:1532)";
    else if (st.code == LBM_ERR_EXTREME_VEL) m = "nve_lf] Extreme vel (kernels_fluid_PART.CUF:1545)";
    else if (st.code == LBM_ERR_OUT_OF_BOX) m = "nve] particle out of box (kernels_fluid_PART.CUF:1581)";
    else if (st.code == LBM_ERR_LIST_OVERFLOW) m = "particle neighbour list full (cf. kernels_fluid_PART.CUF:2436)";
    else if (st.code == LBM_ERR_STORE_RANGE) m = "rest population beyond the range of 16-bit storage";
    std::snprintf(msg, (size_t)len, "%s", m);
  }
  return rc;
}

int lbm_kernel_timing(lbm_ctx* c, int32_t on) {
  if (!c) return LBM_ERR_ARG;
  if (on && !c->ev_pairs) {
    const int cap = 4096;
    c->ev_pairs = new (std::nothrow) hipEvent_t[2 * cap];
    if (!c->ev_pairs) return LBM_ERR_ALLOC;
    for (int e = 0; e < 2 * cap; ++e)
      if (hipEventCreate(&c->ev_pairs[e]) != hipSuccess) {
        for (int q = 0; q < e; ++q) (void)hipEventDestroy(c->ev_pairs[q]);
        delete[] c->ev_pairs;
        c->ev_pairs = nullptr;
        return LBM_ERR_HIP;
      }
    c->ev_cap = cap;
  }
  c->timing = on != 0;
  c->ev_used = 0;
  c->t_total_ms = 0.0;
  c->t_launches = 0;
  return LBM_OK;
}

}  /* extern "C" */

/* Events are recorded on the library stream around each collide-stream
 * launch; the elapsed times are summed only when asked for, so timing adds
 * no host synchronisation to the step loop. */
void lbm_timing_begin(lbm_ctx* c) {
  if (!c->timing) return;
  if (c->ev_used == c->ev_cap) {   /* pool full: fold the recorded pairs in */
    int64_t n;
    double ms;
    lbm_get_kernel_timing(c, &ms, &n);
  }
  (void)hipEventRecord(c->ev_pairs[2 * c->ev_used], c->stream);
}

void lbm_timing_end(lbm_ctx* c) {
  if (!c->timing) return;
  (void)hipEventRecord(c->ev_pairs[2 * c->ev_used + 1], c->stream);
  c->ev_used += 1;
}

extern "C" {

int lbm_get_kernel_timing(lbm_ctx* c, double* total_ms, int64_t* launches) {
  if (!c) return LBM_ERR_ARG;
  if (c->ev_used > 0) {
    if (hipEventSynchronize(c->ev_pairs[2 * c->ev_used - 1]) != hipSuccess) return LBM_ERR_HIP;
    for (int e = 0; e < c->ev_used; ++e) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, c->ev_pairs[2 * e], c->ev_pairs[2 * e + 1]) != hipSuccess) return LBM_ERR_HIP;
      c->t_total_ms += ms;
    }
    c->t_launches += c->ev_used;
    c->ev_used = 0;
  }
  if (total_ms) *total_ms = c->t_total_ms;
  if (launches) *launches = c->t_launches;
  return LBM_OK;
}

int lbm_kernel_info(lbm_ctx* c, char* name, int32_t len, double* bytes_per_lu) {
  if (!c) return LBM_ERR_ARG;
  const char* st = c->storage == LBM_STORE_F32 ? "f32" : (c->storage == LBM_STORE_F16 ? "f16" : "bf16");
  if (name && len > 0) {
    if (c->variant == LBM_KERNEL_FUSED) {
      int tx, ty, wx;
      lbm_fused_tile_shape(c->tile, &tx, &ty, &wx);
      std::snprintf(name, (size_t)len, "k_fused<%s,%s,%dx%d%s,zc=%d>", st, c->np > 0 ? "part" : "fluid", tx, ty,
                    wx ? "w" : "", c->zchunk);
    } else {
      std::snprintf(name, (size_t)len, "k_collide_push<%s,%s>", st, c->np > 0 ? "part" : "fluid");
    }
  }
  if (bytes_per_lu) *bytes_per_lu = 2.0 * 19.0 * 2.0 * (double)c->esize + (c->np > 0 ? 1.0 : 0.0);
  return LBM_OK;
}

}  /* extern "C" */

extern "C" {

int lbm_set_kernel(lbm_ctx* c, int32_t variant, int32_t tile, int32_t zchunk) {
  if (!c) return LBM_ERR_ARG;
  if (variant == LBM_KERNEL_TWOPASS) {
    c->variant = variant;
    return LBM_OK;
  }
  if (variant != LBM_KERNEL_FUSED || c->cfg.diffdens) return LBM_ERR_ARG;
  if (zchunk <= 0) zchunk = c->nz < 32 ? c->nz : 32;
  if (zchunk > c->nz) zchunk = c->nz;
  if (!lbm_fused_tile_valid(c, tile, zchunk)) return LBM_ERR_ARG;
  c->variant = variant;
  c->tile = tile;
  c->zchunk = zchunk;
  return LBM_OK;
}

int lbm_get_kernel(lbm_ctx* c, int32_t* variant, int32_t* tile, int32_t* zchunk) {
  if (!c) return LBM_ERR_ARG;
  if (variant) *variant = c->variant;
  if (tile) *tile = c->tile;
  if (zchunk) *zchunk = c->zchunk;
  return LBM_OK;
}

int32_t lbm_num_tiles(void) { return lbm_fused_num_tiles(); }

int lbm_tile_shape(int32_t tile, int32_t* tx, int32_t* ty, int32_t* wrap) {
  if (tile < 0 || tile >= lbm_fused_num_tiles()) return LBM_ERR_ARG;
  int a, b, w;
  lbm_fused_tile_shape(tile, &a, &b, &w);
  if (tx) *tx = a;
  if (ty) *ty = b;
  if (wrap) *wrap = w;
  return LBM_OK;
}

int lbm_autotune(lbm_ctx* c, int32_t reps, int32_t* tile_out, int32_t* zc_out, double* best_ms) {
  if (!c || reps <= 0 || c->cfg.diffdens) return LBM_ERR_ARG;
  if (!c->initialised) return LBM_ERR_STATE;
  /* the fused kernel only reads the current level, so repeated launches into
     a scratch level leave the state as it was; the status word is restored */
  void* scratch = nullptr;
  if (hipMalloc(&scratch, c->esize * 38 * c->N) != hipSuccess) return LBM_ERR_ALLOC;
  lbm::DevStatus saved;
  int rc = hip_ok(hipMemcpyAsync(&saved, c->d_status, sizeof saved, hipMemcpyDeviceToHost, c->stream));
  rc |= hip_ok(hipStreamSynchronize(c->stream));
  hipEvent_t e0, e1;
  rc |= hip_ok(hipEventCreate(&e0));
  rc |= hip_ok(hipEventCreate(&e1));
  const int zcs[] = {8, 16, 32, 64, 128, 256};
  double best = 1e30;
  int bt = -1, bz = 0;
  const bool parts = c->np > 0;
  /* every candidate is timed in two passes over the whole list and keeps
     its faster pass, so a drift of the clock or memory rate during the sweep
     does not favour the candidates timed early or late */
  std::vector<std::pair<int, int>> cand;
  for (int t = 0; t < lbm_fused_num_tiles(); ++t) {
    int last_zc = -1;
    for (int zc0 : zcs) {
      const int zc = zc0 > c->nz ? c->nz : zc0;
      if (zc == last_zc) continue;
      last_zc = zc;
      if (lbm_fused_tile_valid(c, t, zc)) cand.push_back({t, zc});
    }
  }
  std::vector<double> tmin(cand.size(), 1e30);
  for (int pass = 0; !rc && pass < 2; ++pass) {
    for (size_t k = 0; !rc && k < cand.size(); ++k) {
      const int t = cand[k].first, zc = cand[k].second;
      rc |= lbm_launch_fused_with(c, t, zc, scratch);   /* warm-up */
      rc |= hip_ok(hipEventRecord(e0, c->stream));
      for (int r = 0; r < reps; ++r) rc |= lbm_launch_fused_with(c, t, zc, scratch);
      rc |= hip_ok(hipEventRecord(e1, c->stream));
      rc |= hip_ok(hipEventSynchronize(e1));
      float ms = 0.0f;
      rc |= hip_ok(hipEventElapsedTime(&ms, e0, e1));
      if (!rc) tmin[k] = std::min(tmin[k], (double)ms / reps);
    }
  }
  for (size_t k = 0; !rc && k < cand.size(); ++k)
    if (tmin[k] < best) { best = tmin[k]; bt = cand[k].first; bz = cand[k].second; }
  (void)parts;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  rc |= hip_ok(hipMemcpyAsync(c->d_status, &saved, sizeof saved, hipMemcpyHostToDevice, c->stream));
  rc |= hip_ok(hipStreamSynchronize(c->stream));
  (void)hipFree(scratch);
  if (rc) return LBM_ERR_HIP;
  if (bt < 0) return LBM_ERR_ARG;
  c->variant = LBM_KERNEL_FUSED;
  c->tile = bt;
  c->zchunk = bz;
  if (tile_out) *tile_out = bt;
  if (zc_out) *zc_out = bz;
  if (best_ms) *best_ms = best;
  return LBM_OK;
}

}  /* extern "C" */

extern "C" {

int lbm_copy_bandwidth(int64_t bytes, int32_t reps, double* gbs) {
  if (bytes < 1024 || reps <= 0 || !gbs) return LBM_ERR_ARG;
  const size_t n4 = (size_t)bytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n4 * 16) != hipSuccess) return LBM_ERR_ALLOC;
  if (hipMalloc(&b, n4 * 16) != hipSuccess) { (void)hipFree(a); return LBM_ERR_ALLOC; }
  hipStream_t st;
  int rc = hip_ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  rc |= hip_ok(hipEventCreate(&e0));
  rc |= hip_ok(hipEventCreate(&e1));
  rc |= hip_ok(hipMemsetAsync(a, 0, n4 * 16, st));
  rc |= lbm_launch_copy_probe(a, b, n4, st);   /* warm-up */
  rc |= hip_ok(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) rc |= lbm_launch_copy_probe(a, b, n4, st);
  rc |= hip_ok(hipEventRecord(e1, st));
  rc |= hip_ok(hipEventSynchronize(e1));
  float ms = 0.0f;
  rc |= hip_ok(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  if (rc) return LBM_ERR_HIP;
  *gbs = 2.0 * (double)(n4 * 16) * reps / ((double)ms * 1e-3) / 1e9;
  return LBM_OK;
}

int lbm_stream_bandwidth(int64_t nodes, int32_t reps, double* gbs) {
  if (nodes < 256 || nodes > ((int64_t)1 << 27) || reps <= 0 || !gbs) return LBM_ERR_ARG;
  const size_t bytes = (size_t)19 * (size_t)nodes * sizeof(float) * 2;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return LBM_ERR_ALLOC;
  if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return LBM_ERR_ALLOC; }
  hipStream_t st;
  int rc = hip_ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  rc |= hip_ok(hipEventCreate(&e0));
  rc |= hip_ok(hipEventCreate(&e1));
  rc |= hip_ok(hipMemsetAsync(a, 0, bytes, st));
  rc |= lbm_launch_stream_probe(a, b, (uint32_t)nodes, st);   /* warm-up */
  rc |= hip_ok(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) rc |= lbm_launch_stream_probe(a, b, (uint32_t)nodes, st);
  rc |= hip_ok(hipEventRecord(e1, st));
  rc |= hip_ok(hipEventSynchronize(e1));
  float ms = 0.0f;
  rc |= hip_ok(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  if (rc) return LBM_ERR_HIP;
  *gbs = 2.0 * (double)bytes * reps / ((double)ms * 1e-3) / 1e9;
  return LBM_OK;
}

}  /* extern "C" */

/* ---------------------------------------------------------------------- */
/* Ensemble of independent boxes, advanced together: one persistent launch
 * for the fluid of all boxes and one launch per particle kernel for all of
 * their particles, on one stream (no reference counterpart: the reference
 * runs one box per process). */
struct lbm_ensemble {
  std::vector<lbm_ctx*> box;
  hipStream_t stream;
  lbm_fluid_batch* fb;
  lbm_part_batch* pb;
  long gen;
  int tile, zc;
  int storage;
};

namespace {

long ensemble_gen(const lbm_ensemble* e) {
  long g = 0;
  for (lbm_ctx* c : e->box) g += c->generation * 7 + c->np;
  return g;
}

/* (re)build the per-box argument blocks when a box's buffers changed */
int ensemble_prepare(lbm_ensemble* e) {
  const long g = ensemble_gen(e);
  if (e->fb && e->pb && g == e->gen) return 0;
  lbm_fluid_batch_destroy(e->fb);
  lbm_part_batch_destroy(e->pb);
  e->fb = nullptr;
  e->pb = nullptr;
  int rc = 0;
  e->fb = lbm_fluid_batch_create(e->box.data(), (int)e->box.size(), &rc);
  if (!rc) e->pb = lbm_part_batch_create(e->box.data(), (int)e->box.size(), &rc);
  if (!rc) rc = lbm_fluid_batch_config(e->fb, e->tile, e->zc);
  if (rc) return rc;
  e->gen = g;
  return 0;
}

int ensemble_check_levels(const lbm_ensemble* e) {
  for (lbm_ctx* c : e->box) {
    if (!c->initialised) return LBM_ERR_STATE;
    if (c->cur != e->box[0]->cur || c->step != e->box[0]->step) return LBM_ERR_BATCH;
  }
  return 0;
}

}  // namespace

extern "C" {

int lbm_ensemble_create(int32_t nbox, const lbm_config* cfgs, lbm_ensemble** out) {
  if (!out || !cfgs || nbox <= 0) return LBM_ERR_ARG;
  *out = nullptr;
  for (int i = 0; i < nbox; ++i)
    if (cfgs[i].storage != cfgs[0].storage || cfgs[i].diffdens) return LBM_ERR_BATCH;
  lbm_ensemble* e = new (std::nothrow) lbm_ensemble();
  if (!e) return LBM_ERR_ALLOC;
  e->fb = nullptr;
  e->pb = nullptr;
  e->gen = -1;
  e->tile = 1;    /* 64 x 8 until lbm_ensemble_autotune / set_tile */
  e->zc = 16;
  e->storage = cfgs[0].storage;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return LBM_ERR_HIP;
  }
  for (int i = 0; i < nbox; ++i) {
    lbm_ctx* c = nullptr;
    const int rc = lbm_create(&cfgs[i], &c);
    if (rc) {
      lbm_ensemble_destroy(e);
      return rc;
    }
    e->box.push_back(c);
  }
  *out = e;
  return LBM_OK;
}

int lbm_ensemble_destroy(lbm_ensemble* e) {
  if (!e) return LBM_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  lbm_fluid_batch_destroy(e->fb);
  lbm_part_batch_destroy(e->pb);
  for (lbm_ctx* c : e->box) lbm_destroy(c);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return LBM_OK;
}

int lbm_ensemble_box(lbm_ensemble* e, int32_t i, lbm_ctx** box) {
  if (!e || !box || i < 0 || i >= (int)e->box.size()) return LBM_ERR_ARG;
  *box = e->box[i];
  return LBM_OK;
}

int lbm_ensemble_set_tile(lbm_ensemble* e, int32_t tile, int32_t zchunk) {
  if (!e || !lbm_fluid_batch_tile_ok(tile) || zchunk < 1) return LBM_ERR_ARG;
  e->tile = tile;
  e->zc = zchunk;
  if (e->fb) return lbm_fluid_batch_config(e->fb, tile, zchunk);
  return LBM_OK;
}

int lbm_ensemble_step(lbm_ensemble* e, int32_t nsteps) {
  if (!e || nsteps < 0) return LBM_ERR_ARG;
  int rc = ensemble_check_levels(e);
  if (rc) return rc;
  for (lbm_ctx* c : e->box) rc |= hip_ok(hipStreamSynchronize(c->stream));
  if (rc) return rc;
  rc = ensemble_prepare(e);
  if (rc) return rc;
  int cur = e->box[0]->cur;
  long step = e->box[0]->step;
  for (int s = 0; s < nsteps && !rc; ++s) {
    step += 1;
    rc |= lbm_part_batch_pre(e->pb, cur, step, e->stream);
    rc |= lbm_fluid_batch_launch(e->fb, cur, step, e->stream, nullptr);
    rc |= lbm_part_batch_post(e->pb, cur, step, e->stream);
    cur = 1 - cur;
  }
  for (lbm_ctx* c : e->box) {
    c->cur = cur;
    c->step = step;
  }
  if (rc) return LBM_ERR_HIP;
  /* first error of any box */
  std::vector<lbm::DevStatus> st(e->box.size());
  for (size_t i = 0; i < e->box.size(); ++i)
    rc |= hip_ok(hipMemcpyAsync(&st[i], e->box[i]->d_status, sizeof(lbm::DevStatus), hipMemcpyDeviceToHost, e->stream));
  rc |= hip_ok(hipStreamSynchronize(e->stream));
  if (rc) return LBM_ERR_HIP;
  for (auto& x : st)
    if (x.code) return x.code;
  return LBM_OK;
}

int lbm_ensemble_autotune(lbm_ensemble* e, int32_t reps, int32_t* tile_out, int32_t* zc_out, double* best_ms) {
  if (!e || reps <= 0) return LBM_ERR_ARG;
  int rc = ensemble_check_levels(e);
  if (rc) return rc;
  for (lbm_ctx* c : e->box) rc |= hip_ok(hipStreamSynchronize(c->stream));
  if (rc) return rc;
  rc = ensemble_prepare(e);
  if (rc) return rc;
  /* scratch output levels: the timed launches leave every box as it was */
  std::vector<void*> scratch(e->box.size(), nullptr);
  for (size_t i = 0; i < e->box.size() && !rc; ++i)
    if (hipMalloc(&scratch[i], e->box[i]->esize * 38 * e->box[i]->N) != hipSuccess) rc = LBM_ERR_ALLOC;
  const int par = e->box[0]->cur;
  lbm::FluidArgs* d_sa = rc ? nullptr : lbm_fluid_batch_scratch_args(e->fb, par, scratch.data());
  if (!rc && !d_sa) rc = LBM_ERR_ALLOC;
  std::vector<lbm::DevStatus> saved(e->box.size());
  for (size_t i = 0; i < e->box.size() && !rc; ++i)
    rc |= hip_ok(hipMemcpy(&saved[i], e->box[i]->d_status, sizeof(lbm::DevStatus), hipMemcpyDeviceToHost));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) rc |= hip_ok(hipEventCreate(&e0)) | hip_ok(hipEventCreate(&e1));
  const int zcs[] = {8, 16, 32, 64};
  double best = 1e30;
  int bt = e->tile, bz = e->zc;
  for (int t = 0; !rc && t < lbm_fused_num_tiles(); ++t) {
    if (!lbm_fluid_batch_tile_ok(t)) continue;
    for (int zc : zcs) {
      rc |= lbm_fluid_batch_config(e->fb, t, zc);
      rc |= lbm_fluid_batch_launch(e->fb, par, e->box[0]->step, e->stream, d_sa);   /* warm-up */
      rc |= hip_ok(hipEventRecord(e0, e->stream));
      for (int r = 0; r < reps; ++r) rc |= lbm_fluid_batch_launch(e->fb, par, e->box[0]->step, e->stream, d_sa);
      rc |= hip_ok(hipEventRecord(e1, e->stream));
      rc |= hip_ok(hipEventSynchronize(e1));
      float ms = 0.0f;
      rc |= hip_ok(hipEventElapsedTime(&ms, e0, e1));
      if (rc) break;
      if (ms / reps < best) { best = ms / reps; bt = t; bz = zc; }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (size_t i = 0; i < e->box.size(); ++i)
    if (!rc) rc |= hip_ok(hipMemcpy(e->box[i]->d_status, &saved[i], sizeof(lbm::DevStatus), hipMemcpyHostToDevice));
  (void)hipFree(d_sa);
  for (void* p : scratch) (void)hipFree(p);
  if (rc) return rc;
  e->tile = bt;
  e->zc = bz;
  rc = lbm_fluid_batch_config(e->fb, bt, bz);
  if (tile_out) *tile_out = bt;
  if (zc_out) *zc_out = bz;
  if (best_ms) *best_ms = best;
  return rc;
}

}  /* extern "C" */

extern "C" {

int lbm_set_closure(lbm_ctx* c, const float* block) {
  if (!c) return LBM_ERR_ARG;
  if (!block) {
    dfree(c->d_closure);
    c->d_closure = nullptr;
    c->generation += 1;
    return LBM_OK;
  }
  if (!c->cfg.closure || !c->cfg.with_particles) return LBM_ERR_ARG;
  const float* sc = block + LBM_MLP_NPARAM + 4;
  for (int k = 0; k < 7; ++k)
    if (!(sc[k] != 0.0f)) return LBM_ERR_ARG;   /* scales must be non-zero */
  const float gmax = block[LBM_CLOSURE_NFLOAT - 1];
  if (!(gmax > 0.0f && gmax < 10.0f)) return LBM_ERR_ARG;
  if (!c->d_closure) {
    int rc = dalloc(&c->d_closure, LBM_CLOSURE_NFLOAT);
    if (rc) return rc;
  }
  int rc = hip_ok(hipMemcpy(c->d_closure, block, sizeof(float) * LBM_CLOSURE_NFLOAT, hipMemcpyHostToDevice));
  c->closure_gap_max = gmax;
  c->generation += 1;
  return rc;
}

}  /* extern "C" */
