/*
 * lbm_oracle.cpp -- CPU ORACLE.  TEST INFRASTRUCTURE ONLY.
 *
 * A line-by-line C++/OpenMP restatement of the reference LBcuda time step
 * (two-component colour-gradient D3Q19 fluid + rigid colloids), templated on
 * the real type: the double instantiation is the parity oracle, the float
 * instantiation is the timed "reference CPU path".  It keeps the reference's
 * own data layout and step order so a reader can check it against the CUDA
 * Fortran side by side:
 *   - arrays (0:nx+1, 0:ny+1, 0:nz+1, 0:18, 1:2), halo width 1, Fortran order
 *     (kernels_fluid.CUF:9-48, main.CUF:2224-2376);
 *   - periodic halo refresh by face/edge copies (fixPeriodic, main.CUF:577-610;
 *     bc_per_*2, bc_edge_*2 kernels_fluid_CG.CUF:1215-1426);
 *   - in-place collision, then push streaming into the other buffer
 *     (time_step_CG :531-848, flipflop*Pop0 :866-892, stream*_CG :895-1141);
 *   - the particle kernels of kernels_fluid_PART.CUF in the order of
 *     main.CUF:125-241.
 * Nothing in the product library links or calls this file.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Deterministic choices where the reference is not deterministic:
 *   - particle candidate lists are filled in ascending particle order
 *     (makeListAtomsGPU appends with atomics, order unspecified,
 *     kernels_fluid_PART.CUF:2433-2443);
 *   - each list covers the tiles touched by [nint(x)-6, nint(x)+6] on every
 *     axis, a superset of every node any particle kernel tests.  The
 *     reference's box [int(x-rdim), int(x+rdim)] (PART:2346-2409) can miss the
 *     node nint(x)+6 that make_fluid needs after a move in -x (a reference
 *     defect, tile-shape dependent); the superset removes that dependence.
 *   - FP64 force sums are accumulated per thread and combined in thread order
 *     (the reference uses FP64 atomics, PART:1978-1980).
 * The lubrication branch is applied to the FP64 sums myf2 with a viscosity
 * from tau (the reference adds it non-atomically to forceAtoms, which
 * addForce_particle_bounce_back then overwrites, PART:1463-1501, 1280-1282).
 *
 * The density-ratio model (config diffdens: the reference's DIFFDENS build,
 * equilCG kernels_fluid_CG.CUF:40-49, psi/xi terms :755-792, alphaCG
 * main.CUF:1077-1083, setup2 :163-182) is pinned by
 * tests/test_oracle_diffdens.py: equilibrium moments against the closed forms
 * of a colour-dependent sound speed, pressure balance of the pure phases,
 * mass and momentum conservation of the correction terms.  The alphaCG
 * branch for densR != 1 is the reference's expression as written and is
 * parity unpinned (no fact fixes it; it does not balance the pressures).
 */
#include "../include/lbcuda_amd.h"
#include "lbm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* D3Q19 lattice, dimensions_m.CUF:212-219 */
const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0};
const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1};
const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
/* D3Q27 links for the gradient and belts, dimensions_m.CUF:275-282 */
const int EX27[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1};
const int EY27[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1};
const int EZ27[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1, 1, -1};

/* node types, dimensions_m.CUF:305-309 */
const int8_t FL_WALL = 0, FL_FLUID = 1, FL_LIST = 2, FL_DEAD = 4, FL_CM = 5;

inline long nint_d(double x) { return std::lround(x); } /* Fortran NINT: half away from 0 */

template <class R>
struct Oracle {
  lbm_config cfg;
  int nx, ny, nz;
  long sx, sy, sz, n3;
  int flip = 1;
  long step_no = 0;
  bool initialised = false;
  int err = 0;
  long err_step = 0;
  char err_msg[160] = {0};

  /* constants (SetConstants, main.CUF:1043-1162) */
  R cssq, p[19], bl[19], ad27[27], pd27[27], rec_fact[19], buz;
  R viscR, viscB, densR, densB, sigma, beta;
  /* DIFFDENS (dimensions_m.CUF:230-262, main.CUF:1047-1085) */
  bool dd = false;
  R alphaCG[2], phi_t[19], varphi_t[19], psi_t[19], xi_t[19];

  std::vector<R> pops[2];   /* colour 0 = R, 1 = B; (x,y,z,l,flip) */
  std::vector<R> rho[2];
  std::vector<int8_t> myfluid;   /* (x,y,z,flip) */
  std::vector<int32_t> owner;    /* debugfluid_d */

  /* particles */
  int natoms = 0;
  int rmax = 0, sphereMax = 0;
  std::vector<int8_t> issub;     /* (-rmax:rmax)^3 */
  std::vector<R> pos;            /* pos(13, N, 2) */
  std::vector<R> fatm;           /* forceAtoms(12, N, 2) */
  std::vector<double> myf, myf2, myt, myt2;   /* (3, N) */
  int tile = 4, ntx = 0, nty = 0, ntz = 0;
  std::vector<std::vector<int>> lists;
  std::vector<int> exdouble, eydouble, ezdouble;

  /* ------------------------------------------------------------------ */
  long idx(long i, long j, long k) const { return i + sx * (j + sy * k); }
  R& P(int c, int fl, int l, long i, long j, long k) {
    return pops[c][((long)(fl - 1) * 19 + l) * n3 + idx(i, j, k)];
  }
  int8_t& F(int fl, long i, long j, long k) { return myfluid[(long)(fl - 1) * n3 + idx(i, j, k)]; }
  R& RH(int c, long i, long j, long k) { return rho[c][idx(i, j, k)]; }
  R& POS(int c, int n, int fl) { return pos[((long)(fl - 1) * natoms + n) * 13 + c]; }
  R& FA(int c, int n, int fl) { return fatm[((long)(fl - 1) * natoms + n) * 12 + c]; }

  void raise(int code, const char* msg) {
    if (err == 0) {
      err = code;
      err_step = step_no;
      std::snprintf(err_msg, sizeof err_msg, "%s", msg);
    }
  }

  int create(const lbm_config* c) {
    cfg = *c;
    nx = c->nx; ny = c->ny; nz = c->nz;
    if (nx < 2 || ny < 2 || nz < 2) return LBM_ERR_ARG;
    sx = nx + 2; sy = ny + 2; sz = nz + 2;
    n3 = sx * sy * sz;
    /* dimensions_m.CUF:96-108, 201-268, 283-300; main.CUF:1066-1103 */
    cssq = R(1) / R(3);
    R p0 = R(1) / R(3), p1 = R(1) / R(18), p2 = R(1) / R(36);
    for (int l = 0; l < 19; ++l) p[l] = l == 0 ? p0 : (l < 7 ? p1 : p2);
    R b_xi = R(0.5);
    R b0 = -(R(2) + R(2) * b_xi) / (R(3) * b_xi + R(12));
    R b1 = b_xi / (R(6) * b_xi + R(24));
    R b2 = R(1) / (R(6) * b_xi + R(24));
    for (int l = 0; l < 19; ++l) bl[l] = l == 0 ? b0 : (l < 7 ? b1 : b2);
    R q0 = (R(2) / R(3)) * (R(2) / R(3)) * (R(2) / R(3));
    R q1 = (R(2) / R(3)) * (R(2) / R(3)) * (R(1) / R(6));
    R q2 = (R(1) / R(6)) * (R(1) / R(6)) * (R(2) / R(3));
    R q3 = (R(1) / R(6)) * (R(1) / R(6)) * (R(1) / R(6));
    for (int l = 0; l < 27; ++l) {
      pd27[l] = l == 0 ? q0 : (l < 7 ? q1 : (l < 19 ? q2 : q3));
      ad27[l] = l == 0 ? R(0) : pd27[l] / cssq;
    }
    for (int l = 0; l < 19; ++l)
      rec_fact[l] = std::sqrt(R(EX[l] * EX[l] + EY[l] * EY[l] + EZ[l] * EZ[l]));
    buz = R(2) / cssq;
    densR = R(1); densB = R(1);
    dd = c->diffdens != 0;
    if (dd) {
      if (!(c->dens_r > 0 && c->dens_b > 0 && (c->dens_r == 1.0f || c->dens_b == 1.0f))) return LBM_ERR_ARG;
      densR = R(c->dens_r);
      densB = R(c->dens_b);
    }
    /* alphaCG: equal pressures rho_R (1 - alpha_R) = rho_B (1 - alpha_B) when
       densR = 1; the other branch is the reference's expression as written */
    if (densR == R(1)) {
      alphaCG[0] = p0;
      alphaCG[1] = (-R(1) + alphaCG[0]) * (densR / densB) + R(1);
    } else {
      alphaCG[1] = p0;
      alphaCG[0] = (R(1) / (-R(1) + alphaCG[1])) * (densB / densR) + R(1);
    }
    for (int l = 0; l < 19; ++l) {
      const int g = l == 0 ? 0 : (l < 7 ? 1 : 2);
      const R ph[3] = {R(0), R(1) / R(12), R(1) / R(24)};
      const R vp[3] = {R(1), -R(1) / R(12), -R(1) / R(24)};
      const R ps[3] = {-R(5) / R(2), -R(1) / R(6), R(1) / R(24)};
      const R xs[3] = {R(0), R(1) / R(4), R(1) / R(8)};
      phi_t[l] = ph[g]; varphi_t[l] = vp[g]; psi_t[l] = ps[g]; xi_t[l] = xs[g];
    }
    viscR = cssq * (R(c->tau_r) - R(0.5));
    viscB = cssq * (R(c->tau_b) - R(0.5));
    sigma = R(c->sigma);
    beta = R(c->beta);
    for (int col = 0; col < 2; ++col) {
      pops[col].assign((size_t)n3 * 19 * 2, R(0));
      rho[col].assign((size_t)n3, R(0));
    }
    myfluid.assign((size_t)n3 * 2, FL_WALL);
    owner.assign((size_t)n3, -1);
    /* second-belt offsets, main.CUF:1145-1161 */
    for (int k = -2; k <= 2; ++k)
      for (int j = -2; j <= 2; ++j)
        for (int i = -2; i <= 2; ++i)
          if (i == -2 || i == 2 || j == -2 || j == 2 || k == -2 || k == 2) {
            exdouble.push_back(i); eydouble.push_back(j); ezdouble.push_back(k);
          }
    spherical_template();
    return LBM_OK;
  }

  /* spherical_template, main.CUF:1990-2130 */
  void spherical_template() {
    R rdim = R(cfg.radius);
    R sqrcut = rdim * rdim;
    int rmin = (int)std::floor(rdim);
    rmax = (int)(rdim + R(1));
    int w = 2 * rmax + 1;
    issub.assign((size_t)w * w * w, FL_FLUID);
    auto S = [&](int i, int j, int k) -> int8_t& {
      return issub[(size_t)((i + rmax) + w * ((j + rmax) + w * (k + rmax)))];
    };
    for (int k = -rmax; k <= rmax; ++k)
      for (int j = -rmax; j <= rmax; ++j)
        for (int i = -rmax; i <= rmax; ++i) {
          R rdist = R(i * i + j * j + k * k);
          if (rdist <= sqrcut) S(i, j, k) = FL_DEAD;
        }
    for (int k = -rmin; k <= rmin; ++k)
      for (int j = -rmin; j <= rmin; ++j)
        for (int i = -rmin; i <= rmin; ++i)
          if (S(i, j, k) == FL_DEAD)
            for (int l = 1; l < 19; ++l)
              if (S(i + EX[l], j + EY[l], k + EZ[l]) == FL_FLUID) {
                S(i, j, k) = FL_LIST;
                break;
              }
    S(0, 0, 0) = FL_CM;
    sphereMax = rmax * rmax;
  }
  int8_t issub_at(int i, int j, int k) const {
    int w = 2 * rmax + 1;
    return issub[(size_t)((i + rmax) + w * ((j + rmax) + w * (k + rmax)))];
  }

  /* equil, kernels_fluid_CG.CUF:30-38 */
  R equil(R rh, R u, R v, R w, int l) const {
    R uv = (R(1) / cssq) * (u * R(EX[l]) + v * R(EY[l]) + w * R(EZ[l]));
    return rh * p[l] * (R(1) + uv + R(0.5) * (uv * uv) - (R(0.5) / cssq) * (u * u + v * v + w * w));
  }

  /* ---------------- periodic halo refresh ----------------------------- */
  /* fixPeriodic serial branch, main.CUF:601-610, kernels_fluid_CG.CUF:1215-1426 */
  void fixPeriodic(bool faiPops, bool faiRho) {
    int fl = flip;
    auto cp = [&](long di, long dj, long dk, long si, long sj, long sk) {
      if (faiPops)
        for (int l = 1; l < 19; ++l)
          for (int c = 0; c < 2; ++c) P(c, fl, l, di, dj, dk) = P(c, fl, l, si, sj, sk);
      if (faiRho)
        for (int c = 0; c < 2; ++c) RH(c, di, dj, dk) = RH(c, si, sj, sk);
    };
    /* bc_per_x2 */
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j) { cp(nx + 1, j, k, 1, j, k); cp(0, j, k, nx, j, k); }
    /* bc_per_y2 */
    for (long k = 1; k <= nz; ++k)
      for (long i = 1; i <= nx; ++i) { cp(i, ny + 1, k, i, 1, k); cp(i, 0, k, i, ny, k); }
    /* bc_per_z2 */
    for (long j = 1; j <= ny; ++j)
      for (long i = 1; i <= nx; ++i) { cp(i, j, nz + 1, i, j, 1); cp(i, j, 0, i, j, nz); }
    /* bc_edge_x2 */
    for (long i = 1; i <= nx; ++i) {
      cp(i, 0, 0, i, ny, nz); cp(i, 0, nz + 1, i, ny, 1);
      cp(i, ny + 1, 0, i, 1, nz); cp(i, ny + 1, nz + 1, i, 1, 1);
    }
    /* bc_edge_y2 */
    for (long j = 1; j <= ny; ++j) {
      cp(0, j, 0, nx, j, nz); cp(0, j, nz + 1, nx, j, 1);
      cp(nx + 1, j, 0, 1, j, nz); cp(nx + 1, j, nz + 1, 1, j, 1);
    }
    /* bc_edge_z2 (k = 0..nz+1 also fills the corners) */
    for (long k = 0; k <= nz + 1; ++k) {
      cp(0, 0, k, nx, ny, k); cp(0, ny + 1, k, nx, 1, k);
      cp(nx + 1, 0, k, 1, ny, k); cp(nx + 1, ny + 1, k, 1, 1, k);
    }
  }

  /* fixPeriodic_isfluid serial branch, main.CUF:504-516; kernels :1438-1520 */
  void fixPeriodic_isfluid(int fl) {
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j) { F(fl, nx + 1, j, k) = F(fl, 1, j, k); F(fl, 0, j, k) = F(fl, nx, j, k); }
    for (long k = 1; k <= nz; ++k)
      for (long i = 1; i <= nx; ++i) { F(fl, i, ny + 1, k) = F(fl, i, 1, k); F(fl, i, 0, k) = F(fl, i, ny, k); }
    for (long j = 1; j <= ny; ++j)
      for (long i = 1; i <= nx; ++i) { F(fl, i, j, nz + 1) = F(fl, i, j, 1); F(fl, i, j, 0) = F(fl, i, j, nz); }
    for (long i = 1; i <= nx; ++i) {
      F(fl, i, 0, 0) = F(fl, i, ny, nz); F(fl, i, 0, nz + 1) = F(fl, i, ny, 1);
      F(fl, i, ny + 1, 0) = F(fl, i, 1, nz); F(fl, i, ny + 1, nz + 1) = F(fl, i, 1, 1);
    }
    for (long j = 1; j <= ny; ++j) {
      F(fl, 0, j, 0) = F(fl, nx, j, nz); F(fl, 0, j, nz + 1) = F(fl, nx, j, 1);
      F(fl, nx + 1, j, 0) = F(fl, 1, j, nz); F(fl, nx + 1, j, nz + 1) = F(fl, 1, j, 1);
    }
    for (long k = 0; k <= nz + 1; ++k) {
      F(fl, 0, 0, k) = F(fl, nx, ny, k); F(fl, 0, ny + 1, k) = F(fl, nx, 1, k);
      F(fl, nx + 1, 0, k) = F(fl, 1, ny, k); F(fl, nx + 1, ny + 1, k) = F(fl, 1, 1, k);
    }
  }

  /* ---------------- initial state -------------------------------------- */
  static R fcut(R r, R r1, R r2) {   /* kernels_fluid_CG.CUF:127-138 */
    if (r <= r1) return R(1);
    if (r > r2) return R(0);
    return R(0.5) * std::cos((r - r1) * R(3.141592653589793238462643383279502884) / (r2 - r1)) + R(0.5);
  }

  /* equilCG, kernels_fluid_CG.CUF:40-49 (colour c = 0 red, 1 blue) */
  R equil_cg(R rh, R u, R v, R w, int c, int l) const {
    R uv = (R(1) / cssq) * (u * R(EX[l]) + v * R(EY[l]) + w * R(EZ[l]));
    return rh * (phi_t[l] + varphi_t[l] * alphaCG[c] +
                 p[l] * (uv + R(0.5) * (uv * uv) - (R(0.5) / cssq) * (u * u + v * v + w * w)));
  }

  void set_node_eq(long i, long j, long k, R rR, R rB, R u, R v, R w, bool cg = false) {
    for (int l = 0; l < 19; ++l) {
      P(0, 1, l, i, j, k) = cg ? equil_cg(rR, u, v, w, 0, l) : equil(rR, u, v, w, l);
      P(1, 1, l, i, j, k) = cg ? equil_cg(rB, u, v, w, 1, l) : equil(rB, u, v, w, l);
      P(0, 2, l, i, j, k) = R(0);
      P(1, 2, l, i, j, k) = R(0);
    }
    F(1, i, j, k) = FL_FLUID;
    F(2, i, j, k) = FL_FLUID;
    owner[idx(i, j, k)] = -1;
  }

  int init_fields(int type, const float u0[3], const float* rr, const float* rb, const float* vel) {
    R vx = R(u0 ? u0[0] : 0), vy = R(u0 ? u0[1] : 0), vz = R(u0 ? u0[2] : 0);
    long glx = nx, gly = ny, glz = nz;
    flip = 1;
    step_no = 0;
    err = 0;
    if ((type == LBM_INIT_RHO || type == LBM_INIT_RANDOM || type == LBM_INIT_VELFIELD ||
         type == LBM_INIT_POPS) && (!rr || !rb))
      return LBM_ERR_ARG;
    if (type == LBM_INIT_VELFIELD && !vel) return LBM_ERR_ARG;
    if (type < LBM_INIT_DROPLET || type > LBM_INIT_VELFIELD) return LBM_ERR_ARG;
    for (int c = 0; c < 2; ++c) std::fill(rho[c].begin(), rho[c].end(), R(0));
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i) {
          R rR = 0, rB = 0, u = vx, v = vy, w = vz;
          long h = (i - 1) + nx * ((j - 1) + (long)ny * (k - 1));
          if (type == LBM_INIT_DROPLET) {       /* setup2 */
            R radius = R(glz / 8);
            R dx = R(i) - (R(glx / 2) + R(0.5));
            R dy = R(j) - (R(gly / 2) + R(0.5));
            R dz = R(k) - (R(glz / 2) + R(0.5));
            R rdist = std::sqrt(dx * dx + dy * dy + dz * dz);
            R t = fcut(rdist, radius, radius + R(0.1));
            rR = t; rB = R(1) - t;
            if (dd) { rR = t * densR; rB = (R(1) - t) * densB; }   /* setup2, :163-169 */
          } else if (type == LBM_INIT_ZPLANES) { /* setup2_zplanes */
            R t = (glz * 1 / 8 < k && k < glz * 5 / 8) ? R(0) : R(1);
            rR = t; rB = R(1) - t;
          } else if (type == LBM_INIT_CYLINDER) { /* setupCyl */
            R radius = R(glz / 8);
            R dx = R(i) - (R(glx / 2) + R(0.5));
            R dy = R(j) - (R(gly / 2) + R(0.5));
            R rdist = std::sqrt(dx * dx + dy * dy);
            R t = fcut(rdist, radius, radius + R(0.1));
            rR = t; rB = R(1) - t;
          } else if (type == LBM_INIT_RHO || type == LBM_INIT_RANDOM ||
                     type == LBM_INIT_POPS) {   /* setupPops */
            rR = R(rr[h]); rB = R(rb[h]);
          } else if (type == LBM_INIT_VELFIELD) {
            rR = R(rr[h]); rB = R(rb[h]);
            u = R(vel[3 * h]); v = R(vel[3 * h + 1]); w = R(vel[3 * h + 2]);
          }
          RH(0, i, j, k) = rR;
          RH(1, i, j, k) = rB;
          set_node_eq(i, j, k, rR, rB, u, v, w, dd && type == LBM_INIT_DROPLET);
        }
    initialised = true;
    if (natoms > 0) init_particles_step0();
    fixPeriodic(true, true);   /* main.CUF:111 */
    return LBM_OK;
  }

  int set_pops(const float* pr, const float* pb) {
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i)
          for (int l = 0; l < 19; ++l) {
            long h = (i - 1) + nx * ((j - 1) + (long)ny * ((k - 1) + (long)nz * l));
            P(0, flip, l, i, j, k) = R(pr[h]);
            P(1, flip, l, i, j, k) = R(pb[h]);
          }
    fixPeriodic(true, false);
    return LBM_OK;
  }

  /* ---------------- particles: set-up ---------------------------------- */
  int set_particles(int n, const double* p13) {
    if (!cfg.with_particles || n <= 0) return LBM_ERR_ARG;
    natoms = n;
    pos.assign((size_t)13 * n * 2, R(0));
    fatm.assign((size_t)12 * n * 2, R(0));
    myf.assign((size_t)3 * n, 0.0); myt.assign((size_t)3 * n, 0.0);
    myf2.assign((size_t)3 * n, 0.0); myt2.assign((size_t)3 * n, 0.0);
    /* read_input_atom, main.CUF:1300-1340: both time levels equal */
    for (int a = 0; a < n; ++a)
      for (int fl = 1; fl <= 2; ++fl)
        for (int c = 0; c < 13; ++c) POS(c, a, fl) = R(p13[13 * a + c]);
    for (int a = 0; a < n; ++a)
      for (int fl = 1; fl <= 2; ++fl) {
        for (int c = 0; c < 3; ++c) FA(c, a, fl) = R(cfg.ext_force[c]);
        if (cfg.lrotate)
          for (int c = 0; c < 3; ++c) FA(6 + c, a, fl) = R(cfg.ext_torque[c]);
      }
    /* myf2/myt2 start at the external force (time_step_force_particles resets
       them to ext_* at the end of every step, PART:1126-1145); the reference
       leaves the step-1 values to device initialisation. */
    for (int a = 0; a < n; ++a)
      for (int c = 0; c < 3; ++c) {
        myf2[3 * a + c] = cfg.ext_force[c];
        myt2[3 * a + c] = cfg.lrotate ? cfg.ext_torque[c] : 0.0;
      }
    /* SetupParts sanity checks, main.CUF:1204-1222 */
    R rdim = R(cfg.radius);
    if (nx < 2 * rdim || ny < 2 * rdim || nz < 2 * rdim) return LBM_ERR_PART_SETUP;
    for (int a = 0; a < n; ++a) {
      R x = POS(0, a, 1), y = POS(1, a, 1), z = POS(2, a, 1);
      if (x < 1 || x > nx || y < 1 || y > ny || z < 1 || z > nz) return LBM_ERR_PART_SETUP;
      for (int b = a + 1; b < n; ++b) {
        R dx = x - POS(0, b, 1), dy = y - POS(1, b, 1), dz = z - POS(2, b, 1);
        R d = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (d < 2 * rdim + 1) return LBM_ERR_PART_SETUP;
      }
    }
    ntx = (nx + tile - 1) / tile; nty = (ny + tile - 1) / tile; ntz = (nz + tile - 1) / tile;
    lists.assign((size_t)ntx * nty * ntz, {});
    updateListAtoms(flip);
    if (initialised) {
      init_particles_step0();
      fixPeriodic(true, true);
    }
    return LBM_OK;
  }

  /* setupRho_Pops particle part, main.CUF:2596-2605 */
  void init_particles_step0() {
    init_isfluid(flip);
    fixPeriodic_isfluid(flip);
  }

  /* candidate lists (zeroListAtomsGPU/makeListAtomsGPU, PART:2315-2451) */
  void updateListAtoms(int fl) {
    for (auto& L : lists) L.clear();
    std::vector<int> mark(lists.size(), -1);
    for (int a = 0; a < natoms; ++a) {
      long c[3] = {nint_d(POS(0, a, fl)), nint_d(POS(1, a, fl)), nint_d(POS(2, a, fl))};
      int nn[3] = {nx, ny, nz};
      std::vector<int> tl[3];
      for (int d = 0; d < 3; ++d) {
        std::vector<char> seen((size_t)(nn[d] + tile - 1) / tile, 0);
        for (long o = -rmax; o <= rmax; ++o) {
          long g = ((c[d] + o - 1) % nn[d] + nn[d]) % nn[d];  /* 0-based wrapped */
          int t = (int)(g / tile);
          if (!seen[t]) { seen[t] = 1; tl[d].push_back(t); }
        }
      }
      for (int tk : tl[2])
        for (int tj : tl[1])
          for (int ti : tl[0]) {
            size_t id = (size_t)ti + ntx * ((size_t)tj + nty * tk);
            if (mark[id] != a) { mark[id] = a; lists[id].push_back(a); }
          }
    }
  }
  const std::vector<int>& list_of(long i, long j, long k) const {
    size_t id = (size_t)((i - 1) / tile) + ntx * ((size_t)((j - 1) / tile) + nty * (size_t)((k - 1) / tile));
    return lists[id];
  }

  /* getDeltai/j/k, PART:182-228 */
  int getDelta(long i, long c, int n) const {
    R rdim = R(cfg.radius);
    long d = i - c;
    if (c - rdim < 1 || c + rdim > n) {
      if (c + rdim > n && i < rdim) d = i - c + n;
      else if (c - rdim < 1 && i > n - rdim) d = i - c - n;
    }
    return (int)d;
  }

  /* init_isfluid_CG + setmyfluid, PART:85-117, 230-264 */
  void init_isfluid(int fl) {
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i) {
          bool updated = false;
          owner[idx(i, j, k)] = -1;
          for (int a : list_of(i, j, k)) {
            if (updated) break;
            long xc = nint_d(POS(0, a, fl)), yc = nint_d(POS(1, a, fl)), zc = nint_d(POS(2, a, fl));
            int di = getDelta(i, xc, nx), dj = getDelta(j, yc, ny), dk = getDelta(k, zc, nz);
            if (std::abs(di) <= rmax && std::abs(dj) <= rmax && std::abs(dk) <= rmax) {
              int8_t val = issub_at(di, dj, dk);
              if (val != FL_FLUID) {
                updated = true;
                F(fl, i, j, k) = val;
                owner[idx(i, j, k)] = a;
              }
            }
          }
          if (!updated) F(fl, i, j, k) = FL_FLUID;
        }
  }

  bool particle_moved(int a) {   /* PART:363-380 */
    for (int c = 0; c < 3; ++c)
      if (nint_d(POS(c, a, 1)) != nint_d(POS(c, a, 2))) return true;
    return false;
  }

  /* velocity moments as compute_u/v/w_1fl, PART:40-70 */
  R mom(int c, int fl, long i, long j, long k, int d, R invrho) {
    auto f = [&](int l) { return P(c, fl, l, i, j, k); };
    if (d == 0) return invrho * (f(1) - f(2) + f(7) - f(8) - f(9) + f(10) + f(11) - f(12) - f(13) + f(14));
    if (d == 1) return invrho * (f(3) - f(4) + f(7) - f(8) + f(9) - f(10) + f(15) - f(16) - f(17) + f(18));
    return invrho * (f(5) - f(6) + f(11) - f(12) + f(13) - f(14) + f(15) - f(16) + f(17) - f(18));
  }

  void min_image(R r[3]) const {
    R aaa = R(1) / R(nx), bbb = R(1) / R(ny), ccc = R(1) / R(nz);
    r[0] = r[0] - R(nx) * R(nint_d(aaa * r[0]));
    r[1] = r[1] - R(ny) * R(nint_d(bbb * r[1]));
    r[2] = r[2] - R(nz) * R(nint_d(ccc * r[2]));
  }
  static void cross(const R a[3], const R b[3], R o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  }

  struct Acc {   /* per-thread FP64 force/torque partial sums */
    std::vector<double> f, t;
  };

  void torque_arm(int a, long i, long j, long k, bool minimg, R rt[3]) {
    R rdim = R(cfg.radius);
    rt[0] = R(i) - POS(0, a, flip);
    rt[1] = R(j) - POS(1, a, flip);
    rt[2] = R(k) - POS(2, a, flip);
    if (minimg) min_image(rt);
    R modr = std::sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    R s = rdim / modr;
    rt[0] = s * rt[0]; rt[1] = s * rt[1]; rt[2] = s * rt[2];
  }

  /* del_fluid_CG + del_fluid, PART:120-149, 383-561 */
  void del_fluid_pass() {
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    std::vector<Acc> acc(nthr);
    for (auto& A : acc) { A.f.assign((size_t)3 * natoms, 0.0); A.t.assign((size_t)3 * natoms, 0.0); }
    std::vector<char> moved(natoms);
    for (int a = 0; a < natoms; ++a) moved[a] = particle_moved(a);
    int fo = 3 - flip;
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k) {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      Acc& A = acc[tid];
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i)
          for (int a : list_of(i, j, k)) {
            if (!moved[a]) continue;
            long xc = nint_d(POS(0, a, fo)), yc = nint_d(POS(1, a, fo)), zc = nint_d(POS(2, a, fo));
            int di = getDelta(i, xc, nx), dj = getDelta(j, yc, ny), dk = getDelta(k, zc, nz);
            if (di * di + dj * dj + dk * dk > sphereMax) continue;
            int8_t fn = F(flip, i, j, k);
            if (F(fo, i, j, k) == FL_FLUID && (fn == FL_LIST || fn == FL_DEAD)) {
              R rR = RH(0, i, j, k), rB = RH(1, i, j, k);
              R rs = rR + rB;
              R inv = R(1) / rs;
              R ft[3];
              for (int d = 0; d < 3; ++d)
                ft[d] = (rR + rB) * (mom(0, flip, i, j, k, d, inv) + mom(1, flip, i, j, k, d, inv));
              for (int d = 0; d < 3; ++d) A.f[3 * a + d] += (double)ft[d];
              if (cfg.lrotate) {
                R rt[3], tt[3];
                torque_arm(a, i, j, k, fn == FL_LIST, rt);  /* dead: no min-image (PART:537-545) */
                cross(rt, ft, tt);
                for (int d = 0; d < 3; ++d) A.t[3 * a + d] += (double)tt[d];
              }
            }
          }
    }
    for (int t = 0; t < nthr; ++t)
      for (size_t q = 0; q < (size_t)3 * natoms; ++q) { myf2[q] += acc[t].f[q]; myt2[q] += acc[t].t[q]; }
  }

  /* compute_onebelt_density_twofluids, PART:783-813 */
  void onebelt(long i, long j, long k, R& Rs, R& Bs, int& D) {
    Rs = 0; Bs = 0; D = 0;
    for (int l = 1; l < 27; ++l) {
      long i1 = i + EX27[l], j1 = j + EY27[l], k1 = k + EZ27[l];
      if (F(flip, i1, j1, k1) == FL_FLUID && F(3 - flip, i1, j1, k1) == FL_FLUID) {
        Rs = Rs + RH(0, i1, j1, k1);
        Bs = Bs + RH(1, i1, j1, k1);
        D = D + 1;
      }
    }
    if (D != 0) { Rs = Rs / R(D); Bs = Bs / R(D); }
  }
  long pimage(long i, long n) const { return (i + n - 1) % n + 1; }  /* PART:2115-2121 */
  /* compute_secbelt_density_twofluids, PART:815-852 */
  void secbelt(long i, long j, long k, R& Rs, R& Bs, int& D) {
    Rs = 0; Bs = 0; D = 0;
    for (size_t l = 0; l < exdouble.size(); ++l) {
      long i1 = pimage(i + exdouble[l], nx), j1 = pimage(j + eydouble[l], ny), k1 = pimage(k + ezdouble[l], nz);
      if (F(flip, i1, j1, k1) == FL_FLUID && F(3 - flip, i1, j1, k1) == FL_FLUID) {
        Rs = Rs + RH(0, i1, j1, k1);
        Bs = Bs + RH(1, i1, j1, k1);
        D = D + 1;
      }
    }
    if (D != 0) { Rs = Rs / R(D); Bs = Bs / R(D); }
  }

  /* make_fluid_CG + make_fluid, PART:152-180, 563-747 */
  void make_fluid_pass() {
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    std::vector<Acc> acc(nthr);
    for (auto& A : acc) { A.f.assign((size_t)3 * natoms, 0.0); A.t.assign((size_t)3 * natoms, 0.0); }
    std::vector<char> moved(natoms);
    for (int a = 0; a < natoms; ++a) moved[a] = particle_moved(a);
    int fo = 3 - flip;
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k) {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      Acc& A = acc[tid];
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i)
          for (int a : list_of(i, j, k)) {
            if (!moved[a]) continue;
            long xc = nint_d(POS(0, a, fo)), yc = nint_d(POS(1, a, fo)), zc = nint_d(POS(2, a, fo));
            int di = getDelta(i, xc, nx), dj = getDelta(j, yc, ny), dk = getDelta(k, zc, nz);
            if (di * di + dj * dj + dk * dk > sphereMax) continue;
            int8_t fold = F(fo, i, j, k);
            if ((fold == FL_LIST || fold == FL_DEAD) && F(flip, i, j, k) == FL_FLUID) {
              R Rs, Bs;
              int D;
              onebelt(i, j, k, Rs, Bs, D);
              if (D == 0) {
                secbelt(i, j, k, Rs, Bs, D);
                /* D == 0: fix_onebelt_density_twofluids is empty (PART:854-862) */
              }
              R u = POS(3, a, flip), v = POS(4, a, flip), w = POS(5, a, flip);
              /* initialize_newnode_fluids, PART:896-913 */
              RH(0, i, j, k) = Rs;
              RH(1, i, j, k) = Bs;
              for (int l = 0; l < 19; ++l) {
                P(0, flip, l, i, j, k) = equil(Rs, u, v, w, l);
                P(1, flip, l, i, j, k) = equil(Bs, u, v, w, l);
              }
              R ft[3] = {-(Rs + Bs) * u, -(Rs + Bs) * v, -(Rs + Bs) * w};
              for (int d = 0; d < 3; ++d) A.f[3 * a + d] += (double)ft[d];
              if (cfg.lrotate) {
                R rt[3], tt[3];
                torque_arm(a, i, j, k, fold == FL_LIST, rt);  /* PART:722-727 */
                cross(rt, ft, tt);
                for (int d = 0; d < 3; ++d) A.t[3 * a + d] += (double)tt[d];
              }
            }
          }
    }
    for (int t = 0; t < nthr; ++t)
      for (size_t q = 0; q < (size_t)3 * natoms; ++q) { myf2[q] += acc[t].f[q]; myt2[q] += acc[t].t[q]; }
  }

  /* quaternion helpers, PART:1830-1871 */
  static void qmult(const R a[4], const R b[4], R o[4]) {
    R c[3];
    cross(a + 1, b + 1, c);
    o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    for (int d = 0; d < 3; ++d) o[1 + d] = a[0] * b[1 + d] + b[0] * a[1 + d] + c[d];
  }
  static void qtrimult(const R a[4], const R b[4], const R c[4], R o[4]) {
    R t[4];
    qmult(b, c, t);
    qmult(a, t, o);
  }

  /* compute_virtual_rho, PART:299-352 (Janus) */
  void virtual_rho_pass() {
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i) {
          if (F(flip, i, j, k) != FL_LIST) continue;
          int a = owner[idx(i, j, k)];
          R rt[3] = {R(i) - POS(0, a, flip), R(j) - POS(1, a, flip), R(k) - POS(2, a, flip)};
          min_image(rt);
          R modr = std::sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
          for (int d = 0; d < 3; ++d) rt[d] = rt[d] / modr;
          R q[4] = {POS(9, a, flip), POS(10, a, flip), POS(11, a, flip), POS(12, a, flip)};
          R qc[4] = {q[0], -q[1], -q[2], -q[3]};
          R xv[4] = {0, 1, 0, 0};
          R q2[4];
          qtrimult(q, xv, qc, q2);
          R ac = q2[1] * rt[0] + q2[2] * rt[1] + q2[3] * rt[2];
          if (ac < 0) { RH(0, i, j, k) = 0; RH(1, i, j, k) = 1; }
          else { RH(0, i, j, k) = 1; RH(1, i, j, k) = 0; }
        }
  }

  /* compute_densities_wall + compute_onebelt_density_2fl_weights, PART:267-297, 864-894 */
  void densities_wall_pass() {
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i) {
          int8_t f = F(flip, i, j, k);
          if (!(f == FL_WALL || f == FL_LIST)) continue;
          R Rs = 0, Bs = 0, Ds = 0;
          for (int l = 1; l < 27; ++l) {
            long i1 = i + EX27[l], j1 = j + EY27[l], k1 = k + EZ27[l];
            if (F(flip, i1, j1, k1) == FL_FLUID) {
              Rs = Rs + pd27[l] * RH(0, i1, j1, k1);
              Bs = Bs + pd27[l] * RH(1, i1, j1, k1);
              Ds = Ds + pd27[l];
            }
          }
          if (Ds != R(0)) { Rs = Rs / Ds; Bs = Bs / Ds; }
          if (Ds == R(0) && f == FL_LIST) {
            int D;
            secbelt(i, j, k, Rs, Bs, D);
            if (D != 0) { RH(0, i, j, k) = Rs; RH(1, i, j, k) = Bs; }
          } else {
            RH(0, i, j, k) = Rs;
            RH(1, i, j, k) = Bs;
          }
        }
  }

  /* init_rhoR_CG / init_rhoB_CG, kernels_fluid_CG.CUF:398-509 */
  void init_rho_pass() {
    R minR = R(cfg.min_rho), maxR = R(cfg.max_rho);
    bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i)
          for (int c = 0; c < 2; ++c) {
            if (F(flip, i, j, k) == FL_FLUID) {
              R s = P(c, flip, 0, i, j, k);
              for (int l = 1; l < 19; ++l) s = s + P(c, flip, l, i, j, k);
              if (s < minR || s > maxR) bad = true;
              if (s < R(0)) s = R(0);   /* MINDENS, defines.h:33 */
              RH(c, i, j, k) = s;
            } else {
              RH(c, i, j, k) = R(0);
            }
          }
    if (bad) raise(LBM_ERR_RHO_RANGE, "init_rho_CG] Range error rho (kernels_fluid_CG.CUF:438)");
  }

  /* time_step_CG, kernels_fluid_CG.CUF:531-848 (HIGHGRAD, no DIFFDENS) */
  void collide_pass() {
    const R mylimit = R(1.e-20);
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i) {
          if (F(flip, i, j, k) != FL_FLUID) continue;
          R tR[19], tB[19], fd[19];
          for (int l = 0; l < 19; ++l) { tR[l] = P(0, flip, l, i, j, k); tB[l] = P(1, flip, l, i, j, k); }
          R rR = RH(0, i, j, k), rB = RH(1, i, j, k);
          R rs = rR + rB;
          R phf = (rR / densR - rB / densB) / (rR / densR + rB / densB);
          R visc = R(1) / (R(0.5) * (R(1) + phf) * (R(1) / viscR) + R(0.5) * (R(1) - phf) * (R(1) / viscB));
          R omega = R(1) / (visc / cssq + R(0.5));
          R omm = R(1) - omega;
          R inv = R(1) / rs;
          auto mu = [&](const R* f) { return inv * (f[1] - f[2] + f[7] - f[8] - f[9] + f[10] + f[11] - f[12] - f[13] + f[14]); };
          auto mv = [&](const R* f) { return inv * (f[3] - f[4] + f[7] - f[8] + f[9] - f[10] + f[15] - f[16] - f[17] + f[18]); };
          auto mw = [&](const R* f) { return inv * (f[5] - f[6] + f[11] - f[12] + f[13] - f[14] + f[15] - f[16] + f[17] - f[18]); };
          R u = mu(tR) + mu(tB), v = mv(tR) + mv(tB), w = mw(tR) + mw(tB);
          R px = 0, py = 0, pz = 0;
          R gR[3] = {0, 0, 0}, gB[3] = {0, 0, 0};   /* psi_R, psi_B (DIFFDENS, :755-762) */
          for (int l = 1; l < 27; ++l) {
            long i1 = i + EX27[l], j1 = j + EY27[l], k1 = k + EZ27[l];
            R a = RH(0, i1, j1, k1), b = RH(1, i1, j1, k1);
            R ps = (a / densR - b / densB) / (a / densR + b / densB);
            px = px + ad27[l] * R(EX27[l]) * ps;
            py = py + ad27[l] * R(EY27[l]) * ps;
            pz = pz + ad27[l] * R(EZ27[l]) * ps;
            if (dd) {
              gR[0] = gR[0] + ad27[l] * R(EX27[l]) * a; gR[1] = gR[1] + ad27[l] * R(EY27[l]) * a;
              gR[2] = gR[2] + ad27[l] * R(EZ27[l]) * a;
              gB[0] = gB[0] + ad27[l] * R(EX27[l]) * b; gB[1] = gB[1] + ad27[l] * R(EY27[l]) * b;
              gB[2] = gB[2] + ad27[l] * R(EZ27[l]) * b;
            }
          }
          if (dd) {
            /* :767-792: G = v (x) psi + transpose; only its diagonal meets
               cmat(i,i,l) = e_li^2 */
            const R vel[3] = {u, v, w};
            const R* gs[2] = {gR, gB};
            R* ts[2] = {tR, tB};
            const R rk[2] = {rR, rB};
            for (int c = 0; c < 2; ++c) {
              R vdot = vel[0] * gs[c][0] + vel[1] * gs[c][1] + vel[2] * gs[c][2];
              R G[3];
              for (int d = 0; d < 3; ++d) G[d] = vel[d] * gs[c][d] + vel[d] * gs[c][d];
              for (int l = 0; l < 19; ++l) {
                R tens = G[0] * R(EX[l] * EX[l]) + G[1] * R(EY[l] * EY[l]) + G[2] * R(EZ[l] * EZ[l]);
                ts[c][l] = ts[c][l] * omm + (equil_cg(rk[c], u, v, w, c, l) + visc * (psi_t[l] * vdot + xi_t[l] * tens)) * omega;
              }
            }
          } else {
            for (int l = 0; l < 19; ++l) tR[l] = tR[l] * omm + equil(rR, u, v, w, l) * omega;
            for (int l = 0; l < 19; ++l) tB[l] = tB[l] * omm + equil(rB, u, v, w, l) * omega;
          }
          R pn2 = px * px + py * py + pz * pz;
          if (pn2 > mylimit) {
            R pn = std::sqrt(pn2);
            R ac = R(9) / R(4) * omega * sigma;
            for (int l = 0; l < 19; ++l) {
              R ed = R(EX[l]) * px + R(EY[l]) * py + R(EZ[l]) * pz;
              R t = pn * (p[l] * (ed * ed) / pn2 - bl[l]);
              fd[l] = tR[l] + tB[l] + ac * t;
            }
            for (int l = 0; l < 19; ++l) {
              R feq = dd ? equil_cg(rR, 0, 0, 0, 0, l) + equil_cg(rB, 0, 0, 0, 1, l)
                         : equil(rR, 0, 0, 0, l) + equil(rB, 0, 0, 0, l);
              R ed = R(EX[l]) * px + R(EY[l]) * py + R(EZ[l]) * pz;
              R t1 = rec_fact[l] * pn;
              R cph = (t1 <= mylimit) ? R(0) : ed / t1;
              R t = beta * rR * rB * cph / (rs * rs);
              P(0, flip, l, i, j, k) = fd[l] * rR / rs + t * feq;
              P(1, flip, l, i, j, k) = fd[l] * rB / rs - t * feq;
            }
          } else {
            for (int l = 0; l < 19; ++l) { P(0, flip, l, i, j, k) = tR[l]; P(1, flip, l, i, j, k) = tB[l]; }
          }
        }
  }

  /* partBB_CG + node_to_particle_bounce_back_bc2, PART:916-973, 1874-2113 */
  void n2p_pass() {
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    std::vector<Acc> acc(nthr);
    for (auto& A : acc) { A.f.assign((size_t)3 * natoms, 0.0); A.t.assign((size_t)3 * natoms, 0.0); }
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k) {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      Acc& A = acc[tid];
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i) {
          if (F(flip, i, j, k) != FL_LIST) continue;
          for (int a : list_of(i, j, k)) {
            long xc = nint_d(POS(0, a, flip)), yc = nint_d(POS(1, a, flip)), zc = nint_d(POS(2, a, flip));
            int di = getDelta(i, xc, nx), dj = getDelta(j, yc, ny), dk = getDelta(k, zc, nz);
            if (di * di + dj * dj + dk * dk >= sphereMax) continue;
            link_loop(a, i, j, k, &A, false);
          }
        }
    }
    for (int t = 0; t < nthr; ++t)
      for (size_t q = 0; q < (size_t)3 * natoms; ++q) { myf[q] += acc[t].f[q]; myt[q] += acc[t].t[q]; }
  }

  /* partBB_CG_p2n2 + phase2 (wall=true) or n2p (wall=false) link loop */
  void link_loop(int a, long i, long j, long k, Acc* A, bool wall) {
    R vs[3] = {POS(3, a, flip), POS(4, a, flip), POS(5, a, flip)};
    R v[3] = {vs[0], vs[1], vs[2]};
    R ot[3] = {0, 0, 0}, rv[3] = {0, 0, 0};
    if (cfg.lrotate) {
      R q[4] = {POS(9, a, flip), POS(10, a, flip), POS(11, a, flip), POS(12, a, flip)};
      R qv[4] = {0, POS(6, a, flip), POS(7, a, flip), POS(8, a, flip)};
      R qc[4] = {q[0], -q[1], -q[2], -q[3]};
      R oat[4];
      qtrimult(q, qv, qc, oat);
      ot[0] = oat[1]; ot[1] = oat[2]; ot[2] = oat[3];
      torque_arm(a, i, j, k, true, rv);
    }
    for (int il = 1; il <= 9; ++il) {
      int lo = 2 * il - 1, hi = 2 * il;
      long ii = i + EX[lo], jj = j + EY[lo], kk = k + EZ[lo];
      long io = i + EX[hi], jo = j + EY[hi], ko = k + EZ[hi];
      bool fi = F(flip, ii, jj, kk) == FL_FLUID, fo = F(flip, io, jo, ko) == FL_FLUID;
      for (int side = 0; side < 2; ++side) {
        /* side 0: fluid at ii, wall value into lo; side 1: fluid at io, into hi */
        if (side == 0 && !(fi && !fo)) continue;
        if (side == 1 && !(!fi && fo)) continue;
        long ni = side == 0 ? ii : io, nj = side == 0 ? jj : jo, nk = side == 0 ? kk : ko;
        int lin = side == 0 ? lo : hi;   /* direction from s toward the fluid node */
        int lout = side == 0 ? hi : lo;  /* direction from the fluid node toward s */
        R rt[3] = {0, 0, 0};
        if (cfg.lrotate) {
          for (int d = 0; d < 3; ++d) rt[d] = rv[d];
          rt[0] = rv[0] + R(0.5) * R(EX[lin]);
          rt[1] = rv[1] + R(0.5) * R(EY[lin]);
          rt[2] = rv[2] + R(0.5) * R(EZ[lin]);
          R c[3];
          cross(ot, rt, c);
          v[0] = vs[0] + c[0]; v[1] = vs[1] + c[1]; v[2] = vs[2] + c[2];
        }
        R ev = R(EX[lout]) * v[0] + R(EY[lout]) * v[1] + R(EZ[lout]) * v[2];
        if (!wall) {
          R f2pR = R(2) * P(0, flip, lout, ni, nj, nk) - p[lout] * buz * RH(0, ni, nj, nk) * ev;
          R f2pB = R(2) * P(1, flip, lout, ni, nj, nk) - p[lout] * buz * RH(1, ni, nj, nk) * ev;
          R f2p = f2pR + f2pB;
          R ft[3] = {f2p * R(EX[lout]), f2p * R(EY[lout]), f2p * R(EZ[lout])};
          for (int d = 0; d < 3; ++d) A->f[3 * a + d] += (double)ft[d];
          if (cfg.lrotate) {
            R tt[3];
            cross(rt, ft, tt);
            tt[0] = tt[0] * R(std::abs(EX[lout]));
            tt[1] = tt[1] * R(std::abs(EY[lout]));
            tt[2] = tt[2] * R(std::abs(EZ[lout]));
            for (int d = 0; d < 3; ++d) A->t[3 * a + d] += (double)tt[d];
          }
        } else {
          R f2pR = p[lout] * buz * RH(0, ni, nj, nk) * ev;
          P(0, flip, lin, i, j, k) = P(0, flip, lin, i, j, k) - f2pR;
          R f2pB = p[lout] * buz * RH(1, ni, nj, nk) * ev;
          P(1, flip, lin, i, j, k) = P(1, flip, lin, i, j, k) - f2pB;
        }
      }
    }
  }

  /* partBB_CG_p2n + phase1, PART:975-988, 2153-2199 */
  void p2n1_pass() {
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i) {
          if (F(flip, i, j, k) != FL_LIST) continue;
          for (int il = 1; il <= 9; ++il) {
            int lo = 2 * il - 1, hi = 2 * il;
            long ii = i + EX[lo], jj = j + EY[lo], kk = k + EZ[lo];
            long io = i + EX[hi], jo = j + EY[hi], ko = k + EZ[hi];
            bool fi = F(flip, ii, jj, kk) == FL_FLUID, fo = F(flip, io, jo, ko) == FL_FLUID;
            if (fi && !fo)
              for (int c = 0; c < 2; ++c) P(c, flip, lo, i, j, k) = P(c, flip, hi, ii, jj, kk);
            if (!fi && fo)
              for (int c = 0; c < 2; ++c) P(c, flip, hi, i, j, k) = P(c, flip, lo, io, jo, ko);
          }
        }
  }

  /* partBB_CG_p2n2, PART:990-1040 */
  void p2n2_pass() {
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i) {
          if (F(flip, i, j, k) != FL_LIST) continue;
          for (int a : list_of(i, j, k)) {
            long xc = nint_d(POS(0, a, flip)), yc = nint_d(POS(1, a, flip)), zc = nint_d(POS(2, a, flip));
            int di = getDelta(i, xc, nx), dj = getDelta(j, yc, ny), dk = getDelta(k, zc, nz);
            if (di * di + dj * dj + dk * dk >= sphereMax) continue;
            link_loop(a, i, j, k, nullptr, true);
          }
        }
  }

  /* flipflop*Pop0_CG + stream*_CG + stream_CG_x/_y/_z, kernels_fluid_CG.CUF:866-1141 */
  void stream_pass() {
    int fl = flip, fo = 3 - flip;
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i)
          for (int c = 0; c < 2; ++c) P(c, fo, 0, i, j, k) = P(c, fl, 0, i, j, k);
    /* interior push: every (target, l) slot has exactly one source node, so
       the planes can be pushed in parallel without conflicts. */
#pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i) {
          if (F(fl, i, j, k) == FL_DEAD) continue;
          for (int c = 0; c < 2; ++c)
            for (int l = 1; l < 19; ++l)
              P(c, fo, l, i + EX[l], j + EY[l], k + EZ[l]) = P(c, fl, l, i, j, k);
        }
    /* halo shells push into the interior */
    auto halo_push = [&](long i, long j, long k) {
      if (F(fl, i, j, k) == FL_DEAD) return;
      for (int l = 1; l < 19; ++l) {
        long i1 = i + EX[l], j1 = j + EY[l], k1 = k + EZ[l];
        if (1 <= i1 && i1 <= nx && 1 <= j1 && j1 <= ny && 1 <= k1 && k1 <= nz)
          for (int c = 0; c < 2; ++c) P(c, fo, l, i1, j1, k1) = P(c, fl, l, i, j, k);
      }
    };
    for (long k = 0; k <= nz + 1; ++k)
      for (long j = 0; j <= ny + 1; ++j) { halo_push(0, j, k); halo_push(nx + 1, j, k); }
    for (long k = 0; k <= nz + 1; ++k)
      for (long i = 0; i <= nx + 1; ++i) { halo_push(i, 0, k); halo_push(i, ny + 1, k); }
    for (long j = 0; j <= ny + 1; ++j)
      for (long i = 0; i <= nx + 1; ++i) { halo_push(i, j, 0); halo_push(i, j, nz + 1); }
  }

  /* inter_force_particles + addHertzian, PART:1052-1068, 1385-1503 */
  void inter_force_pass() {
    R rdim = R(cfg.radius);
    R hz_k = R(cfg.hz_k), hz_rmin = R(cfg.hz_rmin), hz_rcap = R(cfg.hz_rcap);
    R rsqcut = hz_rmin * hz_rmin;
    R gmin = R(5) * R(0.5) * hz_k * std::pow(hz_rmin - hz_rcap, R(1.5));
    R rpar = R(2) * rdim;
    R lubfactor = R(cfg.lub_k) * R(6) * R(3.141592653589793238462643383279502884) * (rdim * rdim) * (rdim * rdim) / (rpar * rpar);
    R rparcut = rpar + R(cfg.lub_rmin);
    R rparcap = rpar + R(cfg.lub_rcap);
    R mxrsqcut = std::max(rparcut * rparcut, rsqcut);
    R visc = (R(cfg.tau_r) - R(0.5)) / R(3);   /* dynamic viscosity at rho = 1 */
    for (int ia = 0; ia < natoms; ++ia)
      for (int ja = ia + 1; ja < natoms; ++ja) {
        R xdf = POS(0, ja, flip) - POS(0, ia, flip);
        R ydf = POS(1, ja, flip) - POS(1, ia, flip);
        R zdf = POS(2, ja, flip) - POS(2, ia, flip);
        xdf = xdf - R(nx) * R(nint_d(xdf / R(nx)));
        ydf = ydf - R(ny) * R(nint_d(ydf / R(ny)));
        zdf = zdf - R(nz) * R(nint_d(zdf / R(nz)));
        R rsq = xdf * xdf + ydf * ydf + zdf * zdf;
        if (rsq > mxrsqcut) continue;
        R rrr = std::sqrt(rsq);
        if (cfg.hertz && rrr <= hz_rmin) {
          R g = rrr <= hz_rcap ? gmin : R(5) * R(0.5) * hz_k * std::pow(hz_rmin - rrr, R(1.5));
          R ft[3] = {g * xdf / rrr, g * ydf / rrr, g * zdf / rrr};
          for (int d = 0; d < 3; ++d) {
            myf2[3 * ia + d] += -(double)ft[d];
            myf2[3 * ja + d] += (double)ft[d];
          }
        }
        if (cfg.lubrication && rrr <= rparcut) {
          R ux = xdf / rrr, uy = ydf / rrr, uz = zdf / rrr;
          R rr = rrr <= rparcap ? rparcap : rrr;
          R dotuv = ux * (POS(3, ia, flip) - POS(3, ja, flip)) + uy * (POS(4, ia, flip) - POS(4, ja, flip)) +
                    uz * (POS(5, ia, flip) - POS(5, ja, flip));
          R s = visc * lubfactor * dotuv * (R(1) / (rr - rpar) - R(1));
          R ft[3] = {s * ux, s * uy, s * uz};
          for (int d = 0; d < 3; ++d) {
            myf2[3 * ia + d] += -(double)ft[d];
            myf2[3 * ja + d] += (double)ft[d];
          }
        }
      }
  }

  /* time_step_force_particles -> addForce_particle_bounce_back -> nve_lf ->
     rot_midstep_impl_lf, PART:1098-1166, 1266-1382, 1506-1614, 1621-1827 */
  void particle_update_pass() {
    int fl = flip, fo = 3 - flip;
    R rmass = R(cfg.rmass), rdim = R(cfg.radius);
    R maxv = R(cfg.max_vel);
    R rotin = rmass > 0 ? R(2) / R(5) * (R(1) / rmass) * rdim * rdim : R(0);
    R cx = R(nx + 1) * R(0.5), cy = R(ny + 1) * R(0.5), cz = R(nz + 1) * R(0.5);
    for (int a = 0; a < natoms; ++a) {
      /* addForce_particle_bounce_back */
      for (int d = 0; d < 3; ++d) FA(d, a, fl) = R(myf2[3 * a + d]);
      R fx = FA(0, a, fl), fy = FA(1, a, fl), fz = FA(2, a, fl);
      for (int d = 0; d < 3; ++d) FA(3 + d, a, fl) = R(myf[3 * a + d]);
      fx = fx + (FA(3, a, fl) + FA(3, a, fo)) * R(0.5);
      fy = fy + (FA(4, a, fl) + FA(4, a, fo)) * R(0.5);
      fz = fz + (FA(5, a, fl) + FA(5, a, fo)) * R(0.5);
      FA(0, a, fl) = fx; FA(1, a, fl) = fy; FA(2, a, fl) = fz;
      if (cfg.lrotate) {
        for (int d = 0; d < 3; ++d) FA(6 + d, a, fl) = R(myt2[3 * a + d]);
        for (int d = 0; d < 3; ++d) FA(9 + d, a, fl) = R(myt[3 * a + d]);
        for (int d = 0; d < 3; ++d) FA(6 + d, a, fl) = FA(6 + d, a, fl) + (FA(9 + d, a, fl) + FA(9 + d, a, fo)) * R(0.5);
      }
      /* nve_lf */
      R xo = POS(0, a, fl), yo = POS(1, a, fl), zo = POS(2, a, fl);
      R vx = POS(3, a, fl), vy = POS(4, a, fl), vz = POS(5, a, fl);
      if (std::fabs(fx) * rmass > R(0.5) || std::fabs(fy) * rmass > R(0.5) || std::fabs(fz) * rmass > R(0.5))
        raise(LBM_ERR_BIG_FORCE, "nve_lf] Big force (kernels_fluid_PART.CUF:1532)");
      R bx = vx + rmass * fx, by = vy + rmass * fy, bz = vz + rmass * fz;
      R vm = std::sqrt(bx * bx + by * by + bz * bz);
      if (vm > maxv) raise(LBM_ERR_EXTREME_VEL, "nve_lf] Extreme vel (kernels_fluid_PART.CUF:1545)");
      R x = xo + bx, y = yo + by, z = zo + bz;
      auto pbc = [](R xx, R c, int n) {
        R s = xx - c;
        R aaa = R(1) / R(n);
        return s - R(n) * R(nint_d(aaa * s)) + c;
      };
      x = pbc(x, cx, nx);
      if (x < R(0.5) || x > R(nx) + R(0.5)) raise(LBM_ERR_OUT_OF_BOX, "nve] x out of box (kernels_fluid_PART.CUF:1581)");
      y = pbc(y, cy, ny);
      if (y < R(0.5) || y > R(ny) + R(0.5)) raise(LBM_ERR_OUT_OF_BOX, "nve] y out of box (kernels_fluid_PART.CUF:1589)");
      z = pbc(z, cz, nz);
      if (z < R(0.5) || z > R(nz) + R(0.5)) raise(LBM_ERR_OUT_OF_BOX, "nve] z out of box (kernels_fluid_PART.CUF:1596)");
      /* Without rotation the reference leaves pos(7:13, flop) unwritten; the
         orientation is carried over so both time levels stay defined. */
      if (cfg.lrotate && rmass > 0) rot_midstep(a, rotin);
      else
        for (int c = 6; c < 13; ++c) POS(c, a, fo) = POS(c, a, fl);
      POS(0, a, fo) = x; POS(1, a, fo) = y; POS(2, a, fo) = z;
      POS(3, a, fo) = bx; POS(4, a, fo) = by; POS(5, a, fo) = bz;
      /* reset accumulators, PART:1117-1145 */
      for (int d = 0; d < 3; ++d) {
        FA(d, a, fo) = R(cfg.ext_force[d]);
        FA(3 + d, a, fo) = R(0);
        myf[3 * a + d] = 0.0;
        myf2[3 * a + d] = (double)R(cfg.ext_force[d]);
      }
      if (cfg.lrotate)
        for (int d = 0; d < 3; ++d) {
          FA(6 + d, a, fo) = R(cfg.ext_torque[d]);
          FA(9 + d, a, fo) = R(0);
          myt[3 * a + d] = 0.0;
          myt2[3 * a + d] = (double)R(cfg.ext_torque[d]);
        }
    }
  }

  /* rot_midstep_impl_lf, PART:1621-1827 */
  void rot_midstep(int a, R rotin) {
    int fl = flip, fo = 3 - flip;
    const int mxquat = 20;
    const R sqquattol = R(1.0e-24);
    R ox = POS(6, a, fl), oy = POS(7, a, fl), oz = POS(8, a, fl);
    R q0 = POS(9, a, fl), q1 = POS(10, a, fl), q2 = POS(11, a, fl), q3 = POS(12, a, fl);
    R tqx = FA(6, a, fl), tqy = FA(7, a, fl), tqz = FA(8, a, fl);
    R m[9];
    m[0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    m[1] = R(2) * (q1 * q2 - q0 * q3);
    m[2] = R(2) * (q1 * q3 + q0 * q2);
    m[3] = R(2) * (q1 * q2 + q0 * q3);
    m[4] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
    m[5] = R(2) * (q2 * q3 - q0 * q1);
    m[6] = R(2) * (q1 * q3 - q0 * q2);
    m[7] = R(2) * (q2 * q3 + q0 * q1);
    m[8] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
    R opx = ox, opy = oy, opz = oz;
    R ix = rotin, iy = rotin, iz = rotin;
    R trx = (tqx * m[0] + tqy * m[3] + tqz * m[6]) / ix + (iy - iz) * opy * opz / ix;
    R try_ = (tqx * m[1] + tqy * m[4] + tqz * m[7]) / iy + (iz - ix) * opz * opx / iy;
    R trz = (tqx * m[2] + tqy * m[5] + tqz * m[8]) / iz + (ix - iy) * opx * opy / iz;
    R dx = trx, dy = try_, dz = trz;
    opx = ox + dx * R(0.5); opy = oy + dy * R(0.5); opz = oz + dz * R(0.5);
    R dq0 = (-q1 * opx - q2 * opy - q3 * opz) * R(0.5);
    R dq1 = (q0 * opx - q3 * opy + q2 * opz) * R(0.5);
    R dq2 = (q3 * opx + q0 * opy - q1 * opz) * R(0.5);
    R dq3 = (-q2 * opx + q1 * opy + q0 * opz) * R(0.5);
    R qn0 = q0 + dq0 * R(0.5), qn1 = q1 + dq1 * R(0.5), qn2 = q2 + dq2 * R(0.5), qn3 = q3 + dq3 * R(0.5);
    R rn = R(1) / std::sqrt(qn0 * qn0 + qn1 * qn1 + qn2 * qn2 + qn3 * qn3);
    qn0 *= rn; qn1 *= rn; qn2 *= rn; qn3 *= rn;
    ox = ox + dx; oy = oy + dy; oz = oz + dz;
    POS(6, a, fo) = ox; POS(7, a, fo) = oy; POS(8, a, fo) = oz;
    opx = ox; opy = oy; opz = oz;
    int itq = 0;
    R eps = R(1.0e9);
    while (itq < mxquat && eps > sqquattol) {
      ++itq;
      dq0 = (-qn1 * opx - qn2 * opy - qn3 * opz) * R(0.5);
      dq1 = (qn0 * opx - qn3 * opy + qn2 * opz) * R(0.5);
      dq2 = (qn3 * opx + qn0 * opy - qn1 * opz) * R(0.5);
      dq3 = (-qn2 * opx + qn1 * opy + qn0 * opz) * R(0.5);
      R a0 = qn0 + dq0 * R(0.5), a1 = qn1 + dq1 * R(0.5), a2 = qn2 + dq2 * R(0.5), a3 = qn3 + dq3 * R(0.5);
      rn = R(1) / std::sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3);
      a0 *= rn; a1 *= rn; a2 *= rn; a3 *= rn;
      eps = (a0 - qn0) * (a0 - qn0) + (a1 - qn1) * (a1 - qn1) + (a2 - qn2) * (a2 - qn2) + (a3 - qn3) * (a3 - qn3);
      qn0 = a0; qn1 = a1; qn2 = a2; qn3 = a3;
    }
    q0 = q0 + dq0; q1 = q1 + dq1; q2 = q2 + dq2; q3 = q3 + dq3;
    rn = R(1) / std::sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    POS(9, a, fo) = q0 * rn; POS(10, a, fo) = q1 * rn; POS(11, a, fo) = q2 * rn; POS(12, a, fo) = q3 * rn;
  }

  /* one pass of the step loop body, main.CUF:127-278 (withCG, FLUIDSPLIT) */
  void one_step() {
    ++step_no;
    bool parts = cfg.with_particles && natoms > 0;
    if (parts) {
      init_isfluid(flip);
      fixPeriodic_isfluid(flip);
      del_fluid_pass();
    }
    init_rho_pass();
    fixPeriodic(false, true);
    if (parts) {
      make_fluid_pass();
      fixPeriodic(false, true);
      if (cfg.wetting == LBM_WET_JANUS) virtual_rho_pass();
      else if (cfg.wetting == LBM_WET_NEUTRAL) densities_wall_pass();
      fixPeriodic(false, true);
    }
    collide_pass();
    fixPeriodic(true, false);
    if (parts) {
      n2p_pass();
      p2n1_pass();
      p2n2_pass();
      fixPeriodic(true, false);
    }
    stream_pass();
    if (parts) {
      inter_force_pass();
      particle_update_pass();
      updateListAtoms(3 - flip);
    }
    flip = 3 - flip;
  }

  int step(int n) {
    if (!initialised) return LBM_ERR_STATE;
    for (int s = 0; s < n; ++s) one_step();
    return err;
  }

  /* ---------------- read-back ------------------------------------------ */
  int get_pops(int c, double* out) {
    for (int l = 0; l < 19; ++l)
      for (long k = 1; k <= nz; ++k)
        for (long j = 1; j <= ny; ++j)
          for (long i = 1; i <= nx; ++i)
            out[(i - 1) + nx * ((j - 1) + (long)ny * ((k - 1) + (long)nz * l))] = (double)P(c, flip, l, i, j, k);
    return 0;
  }
  int get_rho(double* rr, double* rb) {
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i) {
          long h = (i - 1) + nx * ((j - 1) + (long)ny * (k - 1));
          if (rr) rr[h] = (double)RH(0, i, j, k);
          if (rb) rb[h] = (double)RH(1, i, j, k);
        }
    return 0;
  }
  int get_flags(int8_t* out) {
    for (long k = 1; k <= nz; ++k)
      for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= nx; ++i)
          out[(i - 1) + nx * ((j - 1) + (long)ny * (k - 1))] = F(flip, i, j, k);
    return 0;
  }
  int get_particles(double* p13, double* f12) {
    for (int a = 0; a < natoms; ++a) {
      if (p13)
        for (int c = 0; c < 13; ++c) p13[13 * a + c] = (double)POS(c, a, flip);
      if (f12)
        for (int c = 0; c < 12; ++c) f12[12 * a + c] = (double)FA(c, a, 3 - flip);
    }
    return 0;
  }
};

struct OrcBox {
  int bits;
  Oracle<double>* d = nullptr;
  Oracle<float>* f = nullptr;
};

}  // namespace

extern "C" {

int orc_create(const lbm_config* cfg, int real_bits, orc_ctx** out) {
  if (!cfg || !out || (real_bits != 32 && real_bits != 64)) return LBM_ERR_ARG;
  OrcBox* b = new OrcBox;
  b->bits = real_bits;
  int rc;
  if (real_bits == 64) { b->d = new Oracle<double>; rc = b->d->create(cfg); }
  else { b->f = new Oracle<float>; rc = b->f->create(cfg); }
  if (rc) { delete b->d; delete b->f; delete b; return rc; }
  *out = reinterpret_cast<orc_ctx*>(b);
  return 0;
}

int orc_destroy(orc_ctx* c) {
  OrcBox* b = reinterpret_cast<OrcBox*>(c);
  if (!b) return 0;
  delete b->d; delete b->f; delete b;
  return 0;
}

#define ORC_DISPATCH(call) \
  OrcBox* b = reinterpret_cast<OrcBox*>(c); \
  if (!b) return LBM_ERR_ARG; \
  return b->bits == 64 ? b->d->call : b->f->call;

int orc_init_fields(orc_ctx* c, int type, const float u0[3], const float* rr, const float* rb, const float* vel) {
  ORC_DISPATCH(init_fields(type, u0, rr, rb, vel))
}
int orc_set_pops(orc_ctx* c, const float* pr, const float* pb) { ORC_DISPATCH(set_pops(pr, pb)) }
int orc_set_particles(orc_ctx* c, int n, const double* p13) { ORC_DISPATCH(set_particles(n, p13)) }
int orc_step(orc_ctx* c, int n) { ORC_DISPATCH(step(n)) }
int orc_get_pops(orc_ctx* c, int color, double* out) { ORC_DISPATCH(get_pops(color, out)) }
int orc_get_rho(orc_ctx* c, double* rr, double* rb) { ORC_DISPATCH(get_rho(rr, rb)) }
int orc_get_flags(orc_ctx* c, int8_t* out) { ORC_DISPATCH(get_flags(out)) }
int orc_get_particles(orc_ctx* c, double* p13, double* f12) { ORC_DISPATCH(get_particles(p13, f12)) }

int orc_status(orc_ctx* c, int* code, long* step, char* msg, int len) {
  OrcBox* b = reinterpret_cast<OrcBox*>(c);
  if (!b) return LBM_ERR_ARG;
  int e = b->bits == 64 ? b->d->err : b->f->err;
  long s = b->bits == 64 ? b->d->err_step : b->f->err_step;
  const char* m = b->bits == 64 ? b->d->err_msg : b->f->err_msg;
  if (code) *code = e;
  if (step) *step = s;
  if (msg && len > 0) std::snprintf(msg, (size_t)len, "%s", m);
  return 0;
}

int orc_template(orc_ctx* c, int* rmax, int8_t* out, int len) {
  OrcBox* b = reinterpret_cast<OrcBox*>(c);
  if (!b) return LBM_ERR_ARG;
  const std::vector<int8_t>& t = b->bits == 64 ? b->d->issub : b->f->issub;
  int r = b->bits == 64 ? b->d->rmax : b->f->rmax;
  if (rmax) *rmax = r;
  if (out) {
    if (len < (int)t.size()) return LBM_ERR_ARG;
    std::memcpy(out, t.data(), t.size());
  }
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  /* extern "C" */
