/*
 * lbm_fluid.hip -- fluid kernels of the colour-gradient step for gfx950.
 *
 * Layout in HBM: for each time level one buffer of 19 direction planes of N
 * colour pairs (f_R, f_B), x fastest, no halo (reference:
 * popsR_d/popsB_d(0:nx+1,0:ny+1,0:nz+1,0:18,1:2), kernels_fluid.CUF:9-12).  Periodicity is done by index wrap inside the
 * kernels; the reference's 24 halo-refresh/halo-stream launches per step
 * (main.CUF:577-1000, kernels_fluid_CG.CUF:1017-1520) have no counterpart.
 *
 * Per step the fluid needs two passes over the populations:
 *   k_density        -- rhoR, rhoB and the phase field phi on every node
 *                       (init_rhoR_CG/init_rhoB_CG, kernels_fluid_CG.CUF:398-509);
 *   k_collide_push   -- collision + perturbation + recolouring of fluid nodes
 *                       (time_step_CG :531-848) fused with the rest-population
 *                       copy and the push streaming into the other buffer
 *                       (flipflop*Pop0_CG :866-892, stream*_CG :929-1015),
 *                       plus the bounce-back copy of particle surface links
 *                       (particle_to_node_bounce_back_bc2_phase1,
 *                       kernels_fluid_PART.CUF:2153-2199).
 */
#include "lbm_common.h"

#include <math.h>

using namespace lbm;

namespace {

constexpr int BX = 64;   /* one wavefront per x-row segment */
constexpr int BY = 4;


/* fcut, kernels_fluid_CG.CUF:127-138 */
__device__ float fcut(float r, float r1, float r2) {
  if (r <= r1) return 1.0f;
  if (r > r2) return 0.0f;
  return 0.5f * cosf((r - r1) * 3.14159265358979323846f / (r2 - r1)) + 0.5f;
}

/* equil, kernels_fluid_CG.CUF:30-38 */
__device__ __forceinline__ float equil(float rho, float u, float v, float w, int l) {
  float uv = (1.0f / CSSQ) * (u * (float)EX(l) + v * (float)EY(l) + w * (float)EZ(l));
  return rho * W(l) * (1.0f + uv + 0.5f * (uv * uv) - (0.5f / CSSQ) * (u * u + v * v + w * w));
}

/* equilCG, kernels_fluid_CG.CUF:40-49: colour k's rest weight alpha_k */
__device__ __forceinline__ float equil_cg(float rho, float u, float v, float w, float alpha, int l) {
  float uv = (1.0f / CSSQ) * (u * (float)EX(l) + v * (float)EY(l) + w * (float)EZ(l));
  return rho * (PHI_DD(l) + VARPHI_DD(l) * alpha +
                W(l) * (uv + 0.5f * (uv * uv) - (0.5f / CSSQ) * (u * u + v * v + w * w)));
}

/* Initial equilibrium state: setup2 / setup2_zplanes / setupCyl / setupPops
 * (kernels_fluid_CG.CUF:140-346).  For host-supplied densities the rho arrays
 * are filled before the launch. */
template <class S>
__global__ __launch_bounds__(256) void k_init_eq(FluidArgs a, int type, float vx, float vy, float vz,
                                                 const float* vel, uint8_t* fl0, uint8_t* fl1) {
  size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  int x = (int)(n % a.nx);
  int y = (int)((n / a.nx) % a.ny);
  int z = (int)(n / ((size_t)a.nx * a.ny));
  int gi = x + 1, gj = y + 1, gk = z + 1;
  int glx = a.nx, gly = a.ny, glz = a.nz;
  float rr, rb;
  float u = vx, v = vy, w = vz;
  if (type == LBM_INIT_DROPLET || type == LBM_INIT_CYLINDER) {
    float radius = (float)(glz / 8);
    float dx = (float)gi - ((float)(glx / 2) + 0.5f);
    float dy = (float)gj - ((float)(gly / 2) + 0.5f);
    float dz = (float)gk - ((float)(glz / 2) + 0.5f);
    float rd = type == LBM_INIT_DROPLET ? sqrtf(dx * dx + dy * dy + dz * dz) : sqrtf(dx * dx + dy * dy);
    float t = fcut(rd, radius, radius + 0.1f);
    rr = t;
    rb = 1.0f - t;
    if (a.diffdens && type == LBM_INIT_DROPLET) {   /* setup2, :163-169 */
      rr = t * a.dens_r;
      rb = (1.0f - t) * a.dens_b;
    }
    a.rho_r[n] = rr;
    a.rho_b[n] = rb;
  } else if (type == LBM_INIT_ZPLANES) {
    float t = (glz * 1 / 8 < gk && gk < glz * 5 / 8) ? 0.0f : 1.0f;
    rr = t;
    rb = 1.0f - t;
    a.rho_r[n] = rr;
    a.rho_b[n] = rb;
  } else {
    rr = a.rho_r[n];
    rb = a.rho_b[n];
    if (type == LBM_INIT_VELFIELD) {
      u = vel[3 * n];
      v = vel[3 * n + 1];
      w = vel[3 * n + 2];
    }
  }
  void* dst = a.dst;                /* current level */
  void* oth = (void*)a.src;         /* other level: cleared */
  const bool dd = a.diffdens && type == LBM_INIT_DROPLET;   /* equilCG in setup2 only, :176-182 */
#pragma unroll
  for (int l = 0; l < 19; ++l) {
    const float er = dd ? equil_cg(rr, u, v, w, a.alpha_r, l) : equil(rr, u, v, w, l);
    const float eb = dd ? equil_cg(rb, u, v, w, a.alpha_b, l) : equil(rb, u, v, w, l);
    if (sizeof(S) == 2 && l == 0 && rest_out_of_range(er - W0 * a.rho0, eb - W0 * a.rho0, a.renc))
      raise_error(a.status, LBM_ERR_STORE_RANGE, a.step);
    st_pair<S>(dst, a.N, l, n, er, eb, a.rho0, a.renc);
    st_pair<S>(oth, a.N, l, n, 0.0f, 0.0f, a.rho0, a.renc);
  }
  fl0[n] = FL_FLUID;
  fl1[n] = FL_FLUID;
}

/* Densities and phase field.  init_rhoR_CG/init_rhoB_CG
 * (kernels_fluid_CG.CUF:398-509): fixed summation order 0..18, range guard,
 * clamp to MINDENS = 0, non-fluid nodes get MINDENS.  The phase field that
 * time_step_CG recomputes for each of 26 neighbours (:728-729) is formed once
 * per node here with the same expression. */
template <class S>
__global__ __launch_bounds__(256) void k_density(FluidArgs a) {
  size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  float rr = 0.0f, rb = 0.0f;
  if (a.flags[n] == FL_FLUID) {
    ld_pair<S>(a.src, a.N, 0, n, a.rho0, a.rdec, rr, rb);
#pragma unroll
    for (int l = 1; l < 19; ++l) {
      float r, b;
      ld_pair<S>(a.src, a.N, l, n, a.rho0, a.rdec, r, b);
      rr = rr + r;
      rb = rb + b;
    }
    if (rr < a.min_rho || rr > a.max_rho || rb < a.min_rho || rb > a.max_rho)
      raise_error(a.status, LBM_ERR_RHO_RANGE, a.step);
    if (rr < 0.0f) rr = 0.0f;
    if (rb < 0.0f) rb = 0.0f;
  }
  a.rho_r[n] = rr;
  a.rho_b[n] = rb;
  a.phi[n] = (rr - rb) / (rr + rb);
}

/* Fused collision + streaming (push).  One thread per node.
 *   fluid nodes   : time_step_CG body (kernels_fluid_CG.CUF:689-846), then the
 *                   19 post-collision values go to x+e_l in the other buffer;
 *                   with particles, a fluid node next to a surface link also
 *                   writes the bounce-back copy into its own opposite slot
 *                   (phase1, kernels_fluid_PART.CUF:2173-2197);
 *   other non-dead: their stored values are pushed unchanged (the reference
 *                   streams from every node that is not spheredead, :941),
 *                   except the slots a fluid neighbour claims for bounce-back;
 *   dead nodes    : only the rest population is carried (flipflop*Pop0_CG).
 */
template <class S, bool PART, bool DD>
__global__ __launch_bounds__(256) void k_collide_push(FluidArgs a) {
  const int x = blockIdx.x * BX + threadIdx.x;
  const int y = blockIdx.y * BY + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= a.nx || y >= a.ny) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t N = a.N;
  const size_t sxy = (size_t)nx * ny;
  const int xs[3] = {x == 0 ? nx - 1 : x - 1, x, x == nx - 1 ? 0 : x + 1};
  const size_t ys[3] = {(size_t)(y == 0 ? ny - 1 : y - 1) * nx, (size_t)y * nx,
                        (size_t)(y == ny - 1 ? 0 : y + 1) * nx};
  const size_t zs[3] = {(size_t)(z == 0 ? nz - 1 : z - 1) * sxy, (size_t)z * sxy,
                        (size_t)(z == nz - 1 ? 0 : z + 1) * sxy};
  const size_t n = (size_t)x + ys[1] + zs[1];
  auto nb = [&](int dx, int dy, int dz) -> size_t { return (size_t)xs[dx + 1] + ys[dy + 1] + zs[dz + 1]; };
  const void* src = a.src;
  void* dst = a.dst;
  const uint8_t f = a.flags[n];

  if (f == FL_FLUID) {
    float fr[19], fb[19];
#pragma unroll
    for (int l = 0; l < 19; ++l) ld_pair<S>(src, N, l, n, a.rho0, a.rdec, fr[l], fb[l]);
    const float rR = a.rho_r[n], rB = a.rho_b[n];
    const float rs = rR + rB;
    /* DIFFDENS: the phase field of normalised densities rho_k/dens_k, :691-692 */
    const float phf = DD ? (rR / a.dens_r - rB / a.dens_b) / (rR / a.dens_r + rB / a.dens_b) : a.phi[n];
    const float visc = 1.0f / (0.5f * (1.0f + phf) * a.inv_visc_r + 0.5f * (1.0f - phf) * a.inv_visc_b);
    const float omega = 1.0f / (visc / CSSQ + 0.5f);
    const float omm = 1.0f - omega;
    const float inv = 1.0f / rs;
    const float u = inv * (fr[1] - fr[2] + fr[7] - fr[8] - fr[9] + fr[10] + fr[11] - fr[12] - fr[13] + fr[14]) +
                    inv * (fb[1] - fb[2] + fb[7] - fb[8] - fb[9] + fb[10] + fb[11] - fb[12] - fb[13] + fb[14]);
    const float v = inv * (fr[3] - fr[4] + fr[7] - fr[8] + fr[9] - fr[10] + fr[15] - fr[16] - fr[17] + fr[18]) +
                    inv * (fb[3] - fb[4] + fb[7] - fb[8] + fb[9] - fb[10] + fb[15] - fb[16] - fb[17] + fb[18]);
    const float w = inv * (fr[5] - fr[6] + fr[11] - fr[12] + fr[13] - fr[14] + fr[15] - fr[16] + fr[17] - fr[18]) +
                    inv * (fb[5] - fb[6] + fb[11] - fb[12] + fb[13] - fb[14] + fb[15] - fb[16] + fb[17] - fb[18]);
    /* 26-point gradient in the D3Q27 order of kernels_fluid_CG.CUF:720-741;
       zero-component terms add nothing and are skipped */
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    float gR[3] = {0.0f, 0.0f, 0.0f}, gB[3] = {0.0f, 0.0f, 0.0f};   /* psi_R, psi_B (DIFFDENS) */
#pragma unroll
    for (int l = 1; l < 27; ++l) {
      const int dx = EX27(l), dy = EY27(l), dz = EZ27(l);
      float ps;
      if (DD) {
        const size_t m = nb(dx, dy, dz);
        const float ra = a.rho_r[m], rb = a.rho_b[m];
        ps = (ra / a.dens_r - rb / a.dens_b) / (ra / a.dens_r + rb / a.dens_b);
        if (dx) { gR[0] = gR[0] + A27(l) * (float)dx * ra; gB[0] = gB[0] + A27(l) * (float)dx * rb; }
        if (dy) { gR[1] = gR[1] + A27(l) * (float)dy * ra; gB[1] = gB[1] + A27(l) * (float)dy * rb; }
        if (dz) { gR[2] = gR[2] + A27(l) * (float)dz * ra; gB[2] = gB[2] + A27(l) * (float)dz * rb; }
      } else {
        ps = a.phi[nb(dx, dy, dz)];
      }
      if (dx) px = px + A27(l) * (float)dx * ps;
      if (dy) py = py + A27(l) * (float)dy * ps;
      if (dz) pz = pz + A27(l) * (float)dz * ps;
    }
    /* BGK towards f^eq(rho_k, u) per colour, :794-800 */
    if (DD) {
      /* DIFFDENS BGK (:767-792): equilCG plus visc (psi_l v.psi_k + xi_l
         sum_i G_k(i,i) e_li^2) with G_k = v (x) psi_k + its transpose */
      const float vR = u * gR[0] + v * gR[1] + w * gR[2];
      const float vB = u * gB[0] + v * gB[1] + w * gB[2];
      const float mR0 = u * gR[0], mR1 = v * gR[1], mR2 = w * gR[2];
      const float mB0 = u * gB[0], mB1 = v * gB[1], mB2 = w * gB[2];
      const float GR0 = mR0 + mR0, GR1 = mR1 + mR1, GR2 = mR2 + mR2;
      const float GB0 = mB0 + mB0, GB1 = mB1 + mB1, GB2 = mB2 + mB2;
#pragma unroll
      for (int l = 0; l < 19; ++l) {
        const float cx = (float)(EX(l) * EX(l)), cy = (float)(EY(l) * EY(l)), cz = (float)(EZ(l) * EZ(l));
        const float tR = GR0 * cx + GR1 * cy + GR2 * cz;
        const float tB = GB0 * cx + GB1 * cy + GB2 * cz;
        fr[l] = fr[l] * omm + (equil_cg(rR, u, v, w, a.alpha_r, l) + visc * (PSI_DD(l) * vR + XI_DD(l) * tR)) * omega;
        fb[l] = fb[l] * omm + (equil_cg(rB, u, v, w, a.alpha_b, l) + visc * (PSI_DD(l) * vB + XI_DD(l) * tB)) * omega;
      }
    } else {
      const float usq = (0.5f / CSSQ) * (u * u + v * v + w * w);
#pragma unroll
      for (int l = 0; l < 19; ++l) {
        const float uv = (1.0f / CSSQ) * (u * (float)EX(l) + v * (float)EY(l) + w * (float)EZ(l));
        const float brk = 1.0f + uv + 0.5f * (uv * uv) - usq;
        fr[l] = fr[l] * omm + rR * W(l) * brk * omega;
        fb[l] = fb[l] * omm + rB * W(l) * brk * omega;
      }
    }
    const float pn2 = px * px + py * py + pz * pz;
    if (pn2 > 1.e-20f) {
      /* perturbation (:803-815) and recolouring (:817-837) */
      const float pn = sqrtf(pn2);
      const float ipn2 = 1.0f / pn2;
      const float ipn = 1.0f / pn;
      const float ac = 9.0f / 4.0f * omega * a.sigma;
      const float bcoef = a.beta * rR * rB / (rs * rs);
#pragma unroll
      for (int l = 0; l < 19; ++l) {
        const float ed = (float)EX(l) * px + (float)EY(l) * py + (float)EZ(l) * pz;
        const float t = pn * (W(l) * (ed * ed) * ipn2 - BL(l));
        const float fd = fr[l] + fb[l] + ac * t;
        const float cph = ed * INV_REC(l) * ipn;
        const float tt = bcoef * cph;
        const float feq = DD ? equil_cg(rR, 0.0f, 0.0f, 0.0f, a.alpha_r, l) + equil_cg(rB, 0.0f, 0.0f, 0.0f, a.alpha_b, l)
                             : rR * W(l) + rB * W(l);
        fr[l] = fd * rR * inv + tt * feq;
        fb[l] = fd * rB * inv - tt * feq;
      }
    }
    /* rest population stays, the others move to x + e_l */
    if (sizeof(S) == 2 && rest_out_of_range(fr[0] - W0 * a.rho0, fb[0] - W0 * a.rho0, a.renc))
      raise_error(a.status, LBM_ERR_STORE_RANGE, a.step);
    st_pair<S>(dst, N, 0, n, fr[0], fb[0], a.rho0, a.renc);
#pragma unroll
    for (int l = 1; l < 19; ++l) st_pair<S>(dst, N, l, nb(EX(l), EY(l), EZ(l)), fr[l], fb[l], a.rho0, a.renc);
    if (PART) {
      /* bounce-back copy: surface node s = x + e_l whose far side s + e_l is
         not fluid sends f_l(x) back to x in direction opp(l) */
#pragma unroll
      for (int l = 1; l < 19; ++l) {
        if (a.flags[nb(EX(l), EY(l), EZ(l))] != FL_LIST) continue;
        const int x2 = wrapc(x + 2 * EX(l), nx), y2 = wrapc(y + 2 * EY(l), ny), z2 = wrapc(z + 2 * EZ(l), nz);
        if (a.flags[(size_t)x2 + (size_t)nx * (y2 + (size_t)ny * z2)] == FL_FLUID) continue;
        st_pair<S>(dst, N, OPP(l), n, fr[l], fb[l], a.rho0, a.renc);
      }
    }
  } else {
    /* non-fluid: raw copies of the stored values (no collision, :689) */
    cp_pair<S>(dst, n, src, n);
    if (f == FL_DEAD) return;
#pragma unroll
    for (int l = 1; l < 19; ++l) {
      const size_t t = nb(EX(l), EY(l), EZ(l));
      if (PART && f == FL_LIST) {
        if (a.flags[t] == FL_FLUID && a.flags[nb(-EX(l), -EY(l), -EZ(l))] != FL_FLUID) continue;
      }
      cp_pair<S>(dst, (size_t)l * N + t, src, (size_t)l * N + n);
    }
  }
}

/* moments2Fl, write_output.f90:975-1041 */
template <class S>
__global__ __launch_bounds__(256) void k_moments(FluidArgs a, float* rr_o, float* rb_o, float* vel_o, float* ph_o) {
  size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  float fr[19], fb[19];
#pragma unroll
  for (int l = 0; l < 19; ++l) ld_pair<S>(a.src, a.N, l, n, a.rho0, a.rdec, fr[l], fb[l]);
  float lr = fr[0], lb = fb[0];
#pragma unroll
  for (int l = 1; l < 19; ++l) { lr = lr + fr[l]; lb = lb + fb[l]; }
  float rs = lr + lb;
  float inv = 1.0f / rs;
  float u = inv * (fr[1] - fr[2] + fr[7] - fr[8] - fr[9] + fr[10] + fr[11] - fr[12] - fr[13] + fr[14] +
                   fb[1] - fb[2] + fb[7] - fb[8] - fb[9] + fb[10] + fb[11] - fb[12] - fb[13] + fb[14]);
  float v = inv * (fr[3] - fr[4] + fr[7] - fr[8] + fr[9] - fr[10] + fr[15] - fr[16] - fr[17] + fr[18] +
                   fb[3] - fb[4] + fb[7] - fb[8] + fb[9] - fb[10] + fb[15] - fb[16] - fb[17] + fb[18]);
  float w = inv * (fr[5] - fr[6] + fr[11] - fr[12] + fr[13] - fr[14] + fr[15] - fr[16] + fr[17] - fr[18] +
                   fb[5] - fb[6] + fb[11] - fb[12] + fb[13] - fb[14] + fb[15] - fb[16] + fb[17] - fb[18]);
  bool ok = a.flags[n] == FL_FLUID && rs > 0.0f;
  if (rr_o) rr_o[n] = ok ? lr : 0.0f;
  if (rb_o) rb_o[n] = ok ? lb : 0.0f;
  if (vel_o) {
    vel_o[3 * n] = ok ? u : 0.0f;
    vel_o[3 * n + 1] = ok ? v : 0.0f;
    vel_o[3 * n + 2] = ok ? w : 0.0f;
  }
  if (ph_o) ph_o[n] = ok ? (lr - lb) / (lr + lb) : 0.0f;
}

/* FP64 global sums over fluid nodes: mass per colour and momentum.  Block
 * partials are written to part[] and reduced by one block in a fixed order. */
template <class S>
__global__ __launch_bounds__(256) void k_sums_partial(FluidArgs a, double* part) {
  __shared__ double sh[5][256];
  double acc[5] = {0, 0, 0, 0, 0};
  for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < a.N; n += (size_t)gridDim.x * blockDim.x) {
    if (a.flags[n] != FL_FLUID) continue;
#pragma unroll
    for (int l = 0; l < 19; ++l) {
      float rf, bf;
      ld_pair<S>(a.src, a.N, l, n, a.rho0, a.rdec, rf, bf);
      const double r = rf, b = bf;
      acc[0] += r;
      acc[1] += b;
      acc[2] += (r + b) * EX(l);
      acc[3] += (r + b) * EY(l);
      acc[4] += (r + b) * EZ(l);
    }
  }
  for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 5; ++q) part[(size_t)blockIdx.x * 5 + q] = sh[q][0];
}

__global__ void k_sums_final(const double* part, int nblk, double* out) {
  __shared__ double sh[5][256];
  double acc[5] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x)
    for (int q = 0; q < 5; ++q) acc[q] += part[(size_t)b * 5 + q];
  for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 5; ++q) out[q] = sh[q][0];
}

/* densities as the reference's rhoR_d/rhoB_d hold them after a step of the
 * single-pass kernel: sums of the populations the step started from on its
 * fluid nodes, the stored (wetting / MINDENS) values elsewhere */
template <class S>
__global__ __launch_bounds__(256) void k_rho_from_pops(FluidArgs a, const uint8_t* fl, float* rr_o, float* rb_o) {
  size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  float rr, rb;
  if (fl[n] == FL_FLUID) {
    ld_pair<S>(a.src, a.N, 0, n, a.rho0, a.rdec, rr, rb);
#pragma unroll
    for (int l = 1; l < 19; ++l) {
      float r, b;
      ld_pair<S>(a.src, a.N, l, n, a.rho0, a.rdec, r, b);
      rr = rr + r;
      rb = rb + b;
    }
    if (rr < 0.0f) rr = 0.0f;
    if (rb < 0.0f) rb = 0.0f;
  } else {
    rr = a.rho_r[n];
    rb = a.rho_b[n];
  }
  rr_o[n] = rr;
  rb_o[n] = rb;
}

template <class S>
__global__ __launch_bounds__(256) void k_unpack(FluidArgs a, int color, float* out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N * 19) return;
  int l = (int)(i / a.N);
  float r, b;
  ld_pair<S>(a.src, a.N, l, i - (size_t)l * a.N, a.rho0, a.rdec, r, b);
  out[i] = color ? b : r;
}

template <class S>
__global__ __launch_bounds__(256) void k_pack(FluidArgs a, const float* pr, const float* pb) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N * 19) return;
  int l = (int)(i / a.N);
  if (sizeof(S) == 2 && l == 0 && rest_out_of_range(pr[i] - W0 * a.rho0, pb[i] - W0 * a.rho0, a.renc))
    raise_error(a.status, LBM_ERR_STORE_RANGE, a.step);
  st_pair<S>(a.dst, a.N, l, i - (size_t)l * a.N, pr[i], pb[i], a.rho0, a.renc);
}

/* plain device copy (read n float4, write n float4): the measured HBM
 * ceiling the fluid kernels are compared with */
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

/* The access pattern of the fused step without its arithmetic: each lane
 * reads the pairs of 19 direction planes of one node and writes 19 pairs,
 * the ones with e_x = +-1 shifted by one node along x (as the push does). */
__global__ __launch_bounds__(256) void k_stream_probe(const float2* __restrict__ src, float2* __restrict__ dst,
                                                      uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  float2 v[19];
#pragma unroll
  for (int l = 0; l < 19; ++l) v[l] = src[(size_t)l * n + i];
  float sx = 0.0f, sy = 0.0f;
#pragma unroll
  for (int l = 0; l < 19; ++l) {
    sx += v[l].x;
    sy += v[l].y;
  }
#pragma unroll
  for (int l = 0; l < 19; ++l) {
    uint32_t j = i;
    if (EX(l) > 0 && i + 1 < n) j = i + 1;
    if (EX(l) < 0 && i > 0) j = i - 1;
    dst[(size_t)l * n + j] = make_float2(v[l].x * 0.5f + sx * 1e-9f, v[l].y * 0.5f + sy * 1e-9f);
  }
}

FluidArgs make_args(lbm_ctx* c) {
  FluidArgs a;
  a.nx = c->nx; a.ny = c->ny; a.nz = c->nz;
  a.N = c->N;
  a.src = c->pop[c->cur];
  a.dst = c->pop[1 - c->cur];
  a.flags = c->flags[c->cur];
  a.rho_r = c->rho_r; a.rho_b = c->rho_b; a.phi = c->phi;
  float viscR = CSSQ * (c->cfg.tau_r - 0.5f);
  float viscB = CSSQ * (c->cfg.tau_b - 0.5f);
  a.inv_visc_r = 1.0f / viscR;
  a.inv_visc_b = 1.0f / viscB;
  a.sigma = c->cfg.sigma; a.beta = c->cfg.beta;
  a.rho0 = c->storage == LBM_STORE_F32 ? 0.0f : c->cfg.store_rho0;
  a.rdec = std::ldexp(1.0f, 118 + rest_offset(c->cfg));
  a.renc = std::ldexp(1.0f, -118 - rest_offset(c->cfg));
  a.min_rho = c->cfg.min_rho; a.max_rho = c->cfg.max_rho;
  a.status = c->d_status;
  a.step = (int)c->step;
  a.diffdens = c->cfg.diffdens;
  a.dens_r = 1.0f;
  a.dens_b = 1.0f;
  a.alpha_r = a.alpha_b = W0;
  if (c->cfg.diffdens) {
    /* SetConstants, main.CUF:1047-1085 (the else branch keeps the
       reference's expression as written) */
    const float dR = c->cfg.dens_r, dB = c->cfg.dens_b;
    a.dens_r = dR;
    a.dens_b = dB;
    if (dR == 1.0f) {
      a.alpha_r = W0;
      a.alpha_b = (-1.0f + a.alpha_r) * (dR / dB) + 1.0f;
    } else {
      a.alpha_b = W0;
      a.alpha_r = (1.0f / (-1.0f + a.alpha_b)) * (dB / dR) + 1.0f;
    }
  }
  return a;
}

inline unsigned nblocks(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

/* the fluid kernels' arguments of context c at its current step */
FluidArgs lbm_fluid_args(lbm_ctx* c) { return make_args(c); }

#define LBM_STORAGE_SWITCH(st, BODY)                 \
  switch (st) {                                      \
    case LBM_STORE_F32: { using S = float; BODY; } break;           \
    case LBM_STORE_F16: { using S = __half; BODY; } break;          \
    default: { using S = __hip_bfloat16; BODY; } break;             \
  }

int lbm_launch_init_eq(lbm_ctx* c, int type, const float u0[3], const float* d_vel) {
  FluidArgs a = make_args(c);
  a.dst = c->pop[c->cur];
  a.src = c->pop[1 - c->cur];
  LBM_STORAGE_SWITCH(c->storage,
    hipLaunchKernelGGL(k_init_eq<S>, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, a, type,
                       u0[0], u0[1], u0[2], d_vel, c->flags[0], c->flags[1]))
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

int lbm_launch_density(lbm_ctx* c) {
  FluidArgs a = make_args(c);
  LBM_STORAGE_SWITCH(c->storage,
    hipLaunchKernelGGL(k_density<S>, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, a))
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

int lbm_launch_collide(lbm_ctx* c) {
  FluidArgs a = make_args(c);
  dim3 grid((unsigned)((c->nx + BX - 1) / BX), (unsigned)((c->ny + BY - 1) / BY), (unsigned)c->nz);
  dim3 block(BX, BY, 1);
  bool part = c->np > 0;
  lbm_timing_begin(c);
  LBM_STORAGE_SWITCH(c->storage,
    if (a.diffdens) {
      if (part) hipLaunchKernelGGL((k_collide_push<S, true, true>), grid, block, 0, c->stream, a);
      else hipLaunchKernelGGL((k_collide_push<S, false, true>), grid, block, 0, c->stream, a);
    } else {
      if (part) hipLaunchKernelGGL((k_collide_push<S, true, false>), grid, block, 0, c->stream, a);
      else hipLaunchKernelGGL((k_collide_push<S, false, false>), grid, block, 0, c->stream, a);
    })
  lbm_timing_end(c);
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

int lbm_launch_moments(lbm_ctx* c, float* rr, float* rb, float* vel, float* ph) {
  FluidArgs a = make_args(c);
  LBM_STORAGE_SWITCH(c->storage,
    hipLaunchKernelGGL(k_moments<S>, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, a, rr, rb, vel, ph))
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

int lbm_launch_sums(lbm_ctx* c, double* d_out) {
  FluidArgs a = make_args(c);
  const int nblk = 1024;
  double* part = nullptr;
  if (hipMallocAsync((void**)&part, sizeof(double) * 5 * nblk, c->stream) != hipSuccess) return LBM_ERR_ALLOC;
  LBM_STORAGE_SWITCH(c->storage,
    hipLaunchKernelGGL(k_sums_partial<S>, dim3(nblk), dim3(256), 0, c->stream, a, part))
  hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(256), 0, c->stream, part, nblk, d_out);
  (void)hipFreeAsync(part, c->stream);
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

int lbm_launch_rho_from_pops(lbm_ctx* c, float* d_rr, float* d_rb) {
  FluidArgs a = make_args(c);
  a.src = c->pop[1 - c->cur];
  LBM_STORAGE_SWITCH(c->storage,
    hipLaunchKernelGGL(k_rho_from_pops<S>, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, a,
                       c->flags[1 - c->cur], d_rr, d_rb))
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

int lbm_launch_unpack_pops(lbm_ctx* c, int color, float* d_out) {
  FluidArgs a = make_args(c);
  LBM_STORAGE_SWITCH(c->storage,
    hipLaunchKernelGGL(k_unpack<S>, dim3(nblocks(c->N * 19, 256)), dim3(256), 0, c->stream, a, color, d_out))
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

int lbm_launch_pack_pops(lbm_ctx* c, const float* pr, const float* pb) {
  FluidArgs a = make_args(c);
  a.dst = c->pop[c->cur];
  LBM_STORAGE_SWITCH(c->storage,
    hipLaunchKernelGGL(k_pack<S>, dim3(nblocks(c->N * 19, 256)), dim3(256), 0, c->stream, a, pr, pb))
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

int lbm_launch_stream_probe(const void* a, void* b, uint32_t nodes, hipStream_t st) {
  hipLaunchKernelGGL(k_stream_probe, dim3((nodes + 255u) / 256u), dim3(256), 0, st, (const float2*)a, (float2*)b,
                     nodes);
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

int lbm_launch_copy_probe(const void* a, void* b, size_t n_float4, hipStream_t st) {
  const unsigned grid = 256 * 64;
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, st, (const float4*)a, (float4*)b, n_float4);
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}
