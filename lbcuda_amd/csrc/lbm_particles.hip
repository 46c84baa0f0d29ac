/*
 * lbm_particles.hip -- colloid coupling kernels for gfx950.
 *
 * The reference runs every particle kernel over the whole grid, one thread per
 * node, scanning a per-tile particle list (kernels_fluid_PART.CUF:85-180,
 * 916-1040 with the lists of :2315-2451).  Here each kernel is
 * particle-centric: one workgroup per particle walks the few hundred nodes
 * around its centre (template or distance ball), so the work scales with the
 * particle surface, not with the box.  Force and torque sums are reduced in
 * FP64 inside the workgroup and written once per particle, so the result does
 * not depend on scheduling.
 *
 * Node ownership: the reference gives a node to the first particle of the
 * tile list that covers it (setmyfluid, PART:230-264).  Here every particle
 * stamps key = (index<<3 | type) with atomicMin, so the lowest index wins.
 */
#include "lbm_common.h"

#include <math.h>

#include <vector>

using namespace lbm;

namespace {

constexpr int TPB = 256;
/* wider workgroups for the passes that walk a particle's whole ball or ring
 * (fewer iterations per thread on their dependent gathers) */
constexpr int NT_SHELL = 512;
constexpr int NT_BALL = 512;
constexpr int MAXNB = 32;   /* neighbour-list capacity per particle */

/* D3Q27 weights pd3q27 */
__device__ __forceinline__ float PD27(int l) { return l == 0 ? P0D : (l < 7 ? P1D : (l < 19 ? P2D : P3D)); }

/* second-belt offsets, main.CUF:1145-1161 (k, j, i loop order) */
__constant__ int8_t C_BELT2[98][3];

struct PartArgs {
  int nx, ny, nz;
  size_t N;
  int np;
  int cur;                 /* current level (flip)             */
  const double* pos_cur;   /* (13, np)                          */
  const double* pos_old;   /* (13, np), the other level         */
  double* pos_new;         /* written by the integrator         */
  double* f_cur;           /* forceAtoms(12, np) current level   */
  double* f_old;           /* forceAtoms other level             */
  double* myf;
  double* myf2;
  double* myt;
  double* myt2;
  uint8_t* fl_cur;
  const uint8_t* fl_old;
  uint32_t* key_cur;
  int* stamp_cur;          /* (3, np)                            */
  uint8_t* restamp;        /* (np) re-stamped by this flag pass  */
  int* claim;
  const int8_t* tmpl;      /* (4, ntmpl)                         */
  int ntmpl;
  const int8_t* ring;      /* (4, nring) D3Q27 ring of the surface */
  int nring;
  const int8_t* tlist;     /* (4, ntlist) surface nodes of the template */
  int ntlist;
  int* nbr;                /* (MAXNB, np) particles whose ball can reach
                              this particle's template, ascending       */
  int* nnbr;               /* (np) number found                         */
  double* contrib;         /* (6, MAXNB, np) link momentum a particle's
                              surface gives to each of its neighbours  */
  uint32_t* smask;         /* (ntlist, np) for each surface template node
                              the particle owns: bit l-1 set if the D3Q19
                              neighbour l is fluid, bit 31 set; 0 for nodes
                              it does not own.  Written by the surface
                              pass before the collision, read by the link
                              pass after it (node types do not change in
                              between)                                  */
  int rmax;
  const int8_t* ball;      /* (4, nball), [3] = 1 if r^2 < rmax^2 */
  int nball;
  float* rho_r;
  float* rho_b;
  float* phi;
  void* pop_src;           /* state buffer of this step          */
  void* pop_dst;           /* next buffer (after collide)        */
  float rho0;
  float rdec, renc;        /* rest-pair scales (16-bit storage)  */
  float radius;
  int lrotate;
  int wetting;
  const float* closure;    /* learned near-contact block or NULL        */
  int rho_from_pops;       /* single-pass step: densities away from the
                              particle shells are not stored, sum them    */
  int lag_from_pops;       /* erase: lagging density from the other level */
  DevStatus* status;
  int step;
};

__device__ __forceinline__ long nint(double x) { return lround(x); }   /* Fortran NINT */

/* x - n nint(x / n), the minimum image of an offset |x| < 1.5 n in a periodic
 * box of n nodes, without the division: nint(x / n) is 1 exactly when
 * x >= n/2 and -1 exactly when x <= -n/2, since the correctly rounded
 * quotient of a double by an integer that small cannot cross 1/2 (the
 * largest double below n/2 divided by n stays below the largest double
 * below 1/2) */
__device__ __forceinline__ double min_image(double x, int n) {
  const double h = 0.5 * (double)n;
  return x >= h ? x - (double)n : (x <= -h ? x + (double)n : x);
}

__device__ __forceinline__ size_t node_at(const PartArgs& a, int cx, int cy, int cz, int dx, int dy, int dz,
                                          int* gx = nullptr, int* gy = nullptr, int* gz = nullptr) {
  /* centres are 1-based lattice coordinates, nodes 0-based; centres lie in
     [1, n+1] and offsets within rmax + 2 <= n (checked at create), so one
     conditional wrap per axis suffices */
  int x = wrapc(cx - 1 + dx, a.nx), y = wrapc(cy - 1 + dy, a.ny), z = wrapc(cz - 1 + dz, a.nz);
  if (gx) { *gx = x; *gy = y; *gz = z; }
  return (size_t)x + (size_t)a.nx * ((size_t)y + (size_t)a.ny * z);
}

__device__ __forceinline__ size_t wrap_node(const PartArgs& a, int x, int y, int z) {
  x = wrapc(x, a.nx); y = wrapc(y, a.ny); z = wrapc(z, a.nz);
  return (size_t)x + (size_t)a.nx * ((size_t)y + (size_t)a.ny * z);
}

__device__ void centre(const double* pos, int p, int c[3]) {
  c[0] = (int)nint(pos[13 * p + 0]);
  c[1] = (int)nint(pos[13 * p + 1]);
  c[2] = (int)nint(pos[13 * p + 2]);
}

__device__ bool moved(const PartArgs& a, int p) {   /* particle_moved, PART:363-380 */
  int c1[3], c2[3];
  centre(a.pos_cur, p, c1);
  centre(a.pos_old, p, c2);
  return c1[0] != c2[0] || c1[1] != c2[1] || c1[2] != c2[2];
}

/* workgroup FP64 reduction of 6 values over NT threads; result valid in
 * thread 0 */
template <int NT = TPB>
__device__ void block_reduce6(double v[6]) {
  __shared__ double sh[6][NT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double x = v[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) sh[q][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      double s = 0.0;
      for (int w = 0; w < NT / 64; ++w) s += sh[q][w];
      v[q] = s;
    }
  __syncthreads();
}

__device__ void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

/* lever arm (R / |r|) r from the particle to node (gx,gy,gz) (0-based);
 * minimum image only when minimg (del_fluid/make_fluid skip it for
 * spheredead nodes, PART:537-545, 722-727). */
__device__ void lever(const PartArgs& a, const double* pos, int p, int gx, int gy, int gz, bool minimg,
                      double r[3]) {
  r[0] = (double)(gx + 1) - pos[13 * p + 0];
  r[1] = (double)(gy + 1) - pos[13 * p + 1];
  r[2] = (double)(gz + 1) - pos[13 * p + 2];
  if (minimg) {
    r[0] = min_image(r[0], a.nx);
    r[1] = min_image(r[1], a.ny);
    r[2] = min_image(r[2], a.nz);
  }
  double s = (double)a.radius / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  r[0] *= s; r[1] *= s; r[2] *= s;
}

/* Hamilton product o = a b (qmult, PART:1854-1860) */
__device__ void qmul(const double a[4], const double b[4], double o[4]) {
  o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  o[1] = a[0] * b[1] + b[0] * a[1] + (a[2] * b[3] - a[3] * b[2]);
  o[2] = a[0] * b[2] + b[0] * a[2] + (a[3] * b[1] - a[1] * b[3]);
  o[3] = a[0] * b[3] + b[0] * a[3] + (a[1] * b[2] - a[2] * b[1]);
}

/* lab-frame angular velocity q (0,omega_body) q*, PART:1900-1912 */
__device__ void omega_lab(const double* pos, int p, double o[3]) {
  const double q[4] = {pos[13 * p + 9], pos[13 * p + 10], pos[13 * p + 11], pos[13 * p + 12]};
  const double qc[4] = {q[0], -q[1], -q[2], -q[3]};
  const double w[4] = {0.0, pos[13 * p + 6], pos[13 * p + 7], pos[13 * p + 8]};
  double t[4], r[4];
  qmul(w, qc, t);
  qmul(q, t, r);
  o[0] = r[1];
  o[1] = r[2];
  o[2] = r[3];
}

/* rho_k = f_0 + ... + f_18 (init_rhoR_CG order, :427-436), clamped at
 * MINDENS = 0 (:443-446) */
template <class S>
__device__ void rho_sum(const void* pops, size_t N, size_t n, float rho0, float rdec, float* rr, float* rb) {
  float r, b;
  ld_pair<S>(pops, N, 0, n, rho0, rdec, r, b);
#pragma unroll
  for (int l = 1; l < 19; ++l) {
    float pr, pb;
    ld_pair<S>(pops, N, l, n, rho0, rdec, pr, pb);
    r = r + pr;
    b = b + pb;
  }
  *rr = r < 0.0f ? 0.0f : r;
  *rb = b < 0.0f ? 0.0f : b;
}

/* density of a node that is fluid on both levels (belt averages) */
template <class S>
__device__ __forceinline__ void belt_rho(const PartArgs& a, size_t m, float* rr, float* rb) {
  if (a.rho_from_pops) {
    rho_sum<S>(a.pop_src, a.N, m, a.rho0, a.rdec, rr, rb);
  } else {
    *rr = a.rho_r[m];
    *rb = a.rho_b[m];
  }
}

/* ---------------- flags ------------------------------------------------ */
/* Kernels without per-particle sums spread one particle's offset list over
 * gridDim.y workgroups (blockIdx.y = chunk), so large spheres (4x resolved
 * pair runs) do not serialise on one workgroup. */
/* template boxes (offsets within rmax per axis) centred at c1 and c2 share a
 * node only if they meet on every axis (minimum image) */
__device__ __forceinline__ bool boxes_meet(const PartArgs& a, const int c1[3], const int c2[3]) {
  const int n[3] = {a.nx, a.ny, a.nz};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int d = c1[k] - c2[k];
    if (2 * d > n[k]) d -= n[k];
    else if (2 * d < -n[k]) d += n[k];
    if (d > 2 * a.rmax || d < -2 * a.rmax) return false;
  }
  return true;
}

/* The flag level written this step was stamped two steps ago, at the centres
 * stamp_cur records.  A particle whose nearest node is the same keeps its
 * nodes, types and owner keys as they stand, unless a particle that moved
 * clears some of them (its old template) or claims some (its new one): this
 * pass clears the old template of each moved particle and marks for k_stamp
 * and k_resolve the moved particles and those whose template box meets a
 * moved particle's old or new box.  Any other particle's nodes hold what
 * re-stamping would write again (the owner is the lowest index either way). */
__device__ __forceinline__ void k_unstamp_body(const PartArgs& a, int p) {
  const int cx = a.stamp_cur[3 * p], cy = a.stamp_cur[3 * p + 1], cz = a.stamp_cur[3 * p + 2];
  int c[3];
  centre(a.pos_cur, p, c);
  const bool moved_p = cx != c[0] || cy != c[1] || cz != c[2];
  if (blockIdx.y == 0) {
    bool hit = false;
    if (!moved_p) {
      for (int q = threadIdx.x; q < a.np; q += TPB) {
        int cq[3];
        centre(a.pos_cur, q, cq);
        const int sq[3] = {a.stamp_cur[3 * q], a.stamp_cur[3 * q + 1], a.stamp_cur[3 * q + 2]};
        if (sq[0] == cq[0] && sq[1] == cq[1] && sq[2] == cq[2]) continue;   /* q kept its node */
        if (boxes_meet(a, c, cq) || (sq[0] != INT32_MIN && boxes_meet(a, c, sq))) hit = true;
      }
    }
    hit = __syncthreads_or(hit);
    if (threadIdx.x == 0) a.restamp[p] = (uint8_t)(moved_p || hit);
  }
  if (!moved_p || cx == INT32_MIN) return;
  for (int t = blockIdx.y * TPB + threadIdx.x; t < a.ntmpl; t += TPB * gridDim.y) {
    size_t n = node_at(a, cx, cy, cz, a.tmpl[4 * t], a.tmpl[4 * t + 1], a.tmpl[4 * t + 2]);
    a.fl_cur[n] = FL_FLUID;
    a.key_cur[n] = 0xffffffffu;
  }
}

__device__ __forceinline__ void k_stamp_body(const PartArgs& a, int p) {
  if (!a.restamp[p]) return;
  int c[3];
  centre(a.pos_cur, p, c);
  for (int t = blockIdx.y * TPB + threadIdx.x; t < a.ntmpl; t += TPB * gridDim.y) {
    size_t n = node_at(a, c[0], c[1], c[2], a.tmpl[4 * t], a.tmpl[4 * t + 1], a.tmpl[4 * t + 2]);
    atomicMin(&a.key_cur[n], ((uint32_t)p << 3) | (uint32_t)a.tmpl[4 * t + 3]);
  }
}

__device__ __forceinline__ void k_resolve_body(const PartArgs& a, int p) {
  if (!a.restamp[p]) return;
  int c[3];
  centre(a.pos_cur, p, c);
  for (int t = blockIdx.y * TPB + threadIdx.x; t < a.ntmpl; t += TPB * gridDim.y) {
    size_t n = node_at(a, c[0], c[1], c[2], a.tmpl[4 * t], a.tmpl[4 * t + 1], a.tmpl[4 * t + 2]);
    uint32_t k = a.key_cur[n];
    if ((k >> 3) == (uint32_t)p) a.fl_cur[n] = (uint8_t)(k & 7u);
  }
  if (threadIdx.x == 0 && blockIdx.y == 0) {
    a.stamp_cur[3 * p] = c[0];
    a.stamp_cur[3 * p + 1] = c[1];
    a.stamp_cur[3 * p + 2] = c[2];
  }
}

/* ---------------- erase: fluid -> solid (del_fluid, PART:383-561) --------
 * The same pass over the old ball claims the nodes that turn fluid for
 * make_fluid (k_fill_apply). */
template <class S, int NT>
__device__ __forceinline__ void k_erase_body(const PartArgs& a, int p) {
  if (!moved(a, p)) return;
  int co[3];
  centre(a.pos_old, p, co);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < a.nball; t += NT) {
    int gx, gy, gz;
    size_t n = node_at(a, co[0], co[1], co[2], a.ball[4 * t], a.ball[4 * t + 1], a.ball[4 * t + 2], &gx, &gy, &gz);
    const uint8_t fnew = a.fl_cur[n], fold = a.fl_old[n];
    /* a node that turns fluid is claimed for make_fluid: the highest index
       among the moving particles whose old ball holds it writes it */
    if ((fold == FL_LIST || fold == FL_DEAD) && fnew == FL_FLUID) atomicMax(&a.claim[n], p);
    if (fold != FL_FLUID || !(fnew == FL_LIST || fnew == FL_DEAD)) continue;
    /* rhoR_d/rhoB_d still hold the previous step's densities here (del_fluid
       runs before init_rho*_CG, main.CUF:130-144); the single-pass step does
       not store them, so they are summed from the previous level */
    float rR, rB;
    if (a.lag_from_pops) {
      rho_sum<S>(a.pop_dst, a.N, n, a.rho0, a.rdec, &rR, &rB);
    } else {
      rR = a.rho_r[n];
      rB = a.rho_b[n];
    }
    const float inv = 1.0f / (rR + rB);
    float fr[19], fb[19];
#pragma unroll
    for (int l = 0; l < 19; ++l) ld_pair<S>(a.pop_src, a.N, l, n, a.rho0, a.rdec, fr[l], fb[l]);
    const float u = inv * (fr[1] - fr[2] + fr[7] - fr[8] - fr[9] + fr[10] + fr[11] - fr[12] - fr[13] + fr[14]) +
                    inv * (fb[1] - fb[2] + fb[7] - fb[8] - fb[9] + fb[10] + fb[11] - fb[12] - fb[13] + fb[14]);
    const float v = inv * (fr[3] - fr[4] + fr[7] - fr[8] + fr[9] - fr[10] + fr[15] - fr[16] - fr[17] + fr[18]) +
                    inv * (fb[3] - fb[4] + fb[7] - fb[8] + fb[9] - fb[10] + fb[15] - fb[16] - fb[17] + fb[18]);
    const float w = inv * (fr[5] - fr[6] + fr[11] - fr[12] + fr[13] - fr[14] + fr[15] - fr[16] + fr[17] - fr[18]) +
                    inv * (fb[5] - fb[6] + fb[11] - fb[12] + fb[13] - fb[14] + fb[15] - fb[16] + fb[17] - fb[18]);
    double ft[3] = {(double)((rR + rB) * u), (double)((rR + rB) * v), (double)((rR + rB) * w)};
    acc[0] += ft[0]; acc[1] += ft[1]; acc[2] += ft[2];
    if (a.lrotate) {
      double r[3], tq[3];
      lever(a, a.pos_cur, p, gx, gy, gz, fnew == FL_LIST, r);
      cross3(r, ft, tq);
      acc[3] += tq[0]; acc[4] += tq[1]; acc[5] += tq[2];
    }
  }
  block_reduce6<NT>(acc);
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; ++d) a.myf2[3 * p + d] += acc[d];
    if (a.lrotate)
      for (int d = 0; d < 3; ++d) a.myt2[3 * p + d] += acc[3 + d];
  }
}

/* ---------------- fill: solid -> fluid (make_fluid, PART:563-747) -------- */
__device__ bool fill_candidate(const PartArgs& a, size_t n) {
  const uint8_t fo = a.fl_old[n];
  return (fo == FL_LIST || fo == FL_DEAD) && a.fl_cur[n] == FL_FLUID;
}

template <class S, int NT>
__device__ __forceinline__ void k_fill_apply_body(const PartArgs& a, int p) {
  if (!moved(a, p)) return;
  int co[3];
  centre(a.pos_old, p, co);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const float vx = (float)a.pos_cur[13 * p + 3], vy = (float)a.pos_cur[13 * p + 4], vz = (float)a.pos_cur[13 * p + 5];
  for (int t = threadIdx.x; t < a.nball; t += NT) {
    int gx, gy, gz;
    size_t n = node_at(a, co[0], co[1], co[2], a.ball[4 * t], a.ball[4 * t + 1], a.ball[4 * t + 2], &gx, &gy, &gz);
    if (!fill_candidate(a, n)) continue;
    const uint8_t fo = a.fl_old[n];
    /* compute_onebelt_density_twofluids, PART:783-813: the neighbours' types
       are loaded in one batch, then the densities of the neighbours fluid on
       both levels, two at a time, summed in link order */
    float Rs = 0.0f, Bs = 0.0f;
    int D = 0;
    {
      uint32_t mn[26];
      uint8_t fc[26], fp[26];
#pragma unroll
      for (int l = 1; l < 27; ++l) mn[l - 1] = (uint32_t)wrap_node(a, gx + EX27(l), gy + EY27(l), gz + EZ27(l));
#pragma unroll
      for (int l = 0; l < 26; ++l) {
        fc[l] = a.fl_cur[mn[l]];
        fp[l] = a.fl_old[mn[l]];
      }
      uint32_t mask = 0;
#pragma unroll
      for (int l = 0; l < 26; ++l)
        if (fc[l] == FL_FLUID && fp[l] == FL_FLUID) mask |= 1u << l;
      while (mask) {
        const int l1 = __ffs(mask) - 1;
        mask &= mask - 1;
        const int l2 = mask ? __ffs(mask) - 1 : -1;
        if (l2 >= 0) mask &= mask - 1;
        float r1, b1, r2 = 0.0f, b2 = 0.0f;
        belt_rho<S>(a, mn[l1], &r1, &b1);
        if (l2 >= 0) belt_rho<S>(a, mn[l2], &r2, &b2);
        Rs = Rs + r1;
        Bs = Bs + b1;
        D = D + 1;
        if (l2 >= 0) {
          Rs = Rs + r2;
          Bs = Bs + b2;
          D = D + 1;
        }
      }
    }
    if (D == 0) {
      /* compute_secbelt_density_twofluids, PART:815-852 */
      for (int l = 0; l < 98; ++l) {
        size_t m = wrap_node(a, gx + C_BELT2[l][0], gy + C_BELT2[l][1], gz + C_BELT2[l][2]);
        if (a.fl_cur[m] == FL_FLUID && a.fl_old[m] == FL_FLUID) {
          float mr, mb;
          belt_rho<S>(a, m, &mr, &mb);
          Rs = Rs + mr;
          Bs = Bs + mb;
          D = D + 1;
        }
      }
    }
    if (D != 0) { Rs = Rs / (float)D; Bs = Bs / (float)D; }
    /* the highest-index particle writes the node, as the last list entry does */
    if (a.claim[n] == p) {
      a.rho_r[n] = Rs;
      a.rho_b[n] = Bs;
      a.phi[n] = (Rs - Bs) / (Rs + Bs);
      const float usq = (0.5f / CSSQ) * (vx * vx + vy * vy + vz * vz);
#pragma unroll
      for (int l = 0; l < 19; ++l) {
        const float uv = (1.0f / CSSQ) * (vx * (float)EX(l) + vy * (float)EY(l) + vz * (float)EZ(l));
        const float brk = 1.0f + uv + 0.5f * (uv * uv) - usq;
        st_pair<S>(a.pop_src, a.N, l, n, Rs * W(l) * brk, Bs * W(l) * brk, a.rho0, a.renc);
      }
    }
    double ft[3] = {(double)(-(Rs + Bs) * vx), (double)(-(Rs + Bs) * vy), (double)(-(Rs + Bs) * vz)};
    acc[0] += ft[0]; acc[1] += ft[1]; acc[2] += ft[2];
    if (a.lrotate) {
      double r[3], tq[3];
      lever(a, a.pos_cur, p, gx, gy, gz, fo == FL_LIST, r);
      cross3(r, ft, tq);
      acc[3] += tq[0]; acc[4] += tq[1]; acc[5] += tq[2];
    }
  }
  block_reduce6<NT>(acc);
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; ++d) a.myf2[3 * p + d] += acc[d];
    if (a.lrotate)
      for (int d = 0; d < 3; ++d) a.myt2[3 * p + d] += acc[3 + d];
  }
}

template <int NT = TPB>
__device__ __forceinline__ void k_fill_release_body(const PartArgs& a, int p) {
  if (!moved(a, p)) return;
  int co[3];
  centre(a.pos_old, p, co);
  for (int t = blockIdx.y * NT + threadIdx.x; t < a.nball; t += NT * gridDim.y) {
    size_t n = node_at(a, co[0], co[1], co[2], a.ball[4 * t], a.ball[4 * t + 1], a.ball[4 * t + 2]);
    a.claim[n] = -1;
  }
}

/* ---------------- wetting of surface nodes -------------------------------- */
/* surface node t of the template list of particle p (centre c): if p owns
 * it, its D3Q19 fluid-neighbour mask for the link pass and, unless wetting is
 * off, its Janus or neutral wetting densities */
template <class S>
__device__ __forceinline__ void surface_node(const PartArgs& a, int p, const int c[3], int t) {
  int gx, gy, gz;
  size_t n = node_at(a, c[0], c[1], c[2], a.tlist[4 * t], a.tlist[4 * t + 1], a.tlist[4 * t + 2], &gx, &gy, &gz);
  uint32_t* mslot = a.smask + (size_t)p * a.ntlist + t;
  const uint8_t fn = a.fl_cur[n];
  const uint32_t kn = a.key_cur[n];
  if ((fn != FL_LIST) | ((kn >> 3) != (uint32_t)p)) {   /* both loads issued together */
    *mslot = 0u;
    return;
  }
  /* the 26 neighbours' types (and densities for neutral wetting) in one batch */
  uint8_t f27[27];
  float r27[27], b27[27];
  if (a.wetting == LBM_WET_NEUTRAL) {
#pragma unroll
    for (int l = 1; l < 27; ++l) {
      const size_t m = wrap_node(a, gx + EX27(l), gy + EY27(l), gz + EZ27(l));
      f27[l] = a.fl_cur[m];
      r27[l] = a.rho_r[m];
      b27[l] = a.rho_b[m];
    }
  } else {
#pragma unroll
    for (int l = 1; l < 27; ++l) {
      f27[l] = a.fl_cur[wrap_node(a, gx + EX27(l), gy + EY27(l), gz + EZ27(l))];
      r27[l] = 0.0f;
      b27[l] = 0.0f;
    }
  }
  uint32_t mask = 1u << 31;
#pragma unroll
  for (int l = 1; l < 19; ++l)
    if (f27[l] == FL_FLUID) mask |= 1u << (l - 1);
  *mslot = mask;
  if (a.wetting == LBM_WET_NONE) return;
  float rr, rb;
  if (a.wetting == LBM_WET_JANUS) {
    /* compute_virtual_rho, PART:299-352: sign of r.(q x q*) */
    double r[3] = {(double)(gx + 1) - a.pos_cur[13 * p], (double)(gy + 1) - a.pos_cur[13 * p + 1],
                   (double)(gz + 1) - a.pos_cur[13 * p + 2]};
    r[0] = min_image(r[0], a.nx);
    r[1] = min_image(r[1], a.ny);
    r[2] = min_image(r[2], a.nz);
    const double q0 = a.pos_cur[13 * p + 9], q1 = a.pos_cur[13 * p + 10], q2 = a.pos_cur[13 * p + 11],
                 q3 = a.pos_cur[13 * p + 12];
    /* body x axis in the lab frame: first column of the rotation matrix */
    const double ax = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    const double ay = 2.0 * (q1 * q2 + q0 * q3);
    const double az = 2.0 * (q1 * q3 - q0 * q2);
    const double ac = ax * r[0] + ay * r[1] + az * r[2];
    if (ac < 0.0) { rr = 0.0f; rb = 1.0f; } else { rr = 1.0f; rb = 0.0f; }
  } else {
    /* compute_densities_wall + compute_onebelt_density_2fl_weights, PART:267-297, 864-894 */
    float Rs = 0.0f, Bs = 0.0f, Ds = 0.0f;
#pragma unroll
    for (int l = 1; l < 27; ++l) {
      if (f27[l] == FL_FLUID) {
        Rs = fmaf(PD27(l), r27[l], Rs);   /* fused in both shell passes alike */
        Bs = fmaf(PD27(l), b27[l], Bs);
        Ds = Ds + PD27(l);
      }
    }
    if (Ds != 0.0f) {
      rr = Rs / Ds;
      rb = Bs / Ds;
    } else {
      float R2 = 0.0f, B2 = 0.0f;
      int D = 0;
      for (int l = 0; l < 98; ++l) {
        size_t m = wrap_node(a, gx + C_BELT2[l][0], gy + C_BELT2[l][1], gz + C_BELT2[l][2]);
        if (a.fl_cur[m] == FL_FLUID && a.fl_old[m] == FL_FLUID) {
          float mr, mb;
          belt_rho<S>(a, m, &mr, &mb);
          R2 = R2 + mr;
          B2 = B2 + mb;
          D = D + 1;
        }
      }
      if (D == 0) return;   /* fix_onebelt_density_twofluids is empty */
      rr = R2 / (float)D;
      rb = B2 / (float)D;
    }
  }
  a.rho_r[n] = rr;
  a.rho_b[n] = rb;
  a.phi[n] = (rr - rb) / (rr + rb);
}

template <class S>
__device__ __forceinline__ void k_wetting_body(const PartArgs& a, int p) {
  int c[3];
  centre(a.pos_cur, p, c);
  for (int t = blockIdx.y * TPB + threadIdx.x; t < a.ntlist; t += TPB * gridDim.y) surface_node<S>(a, p, c, t);
}

/* ---------------- densities around each particle (single-pass step) -----
 * The single-pass fluid kernel sums densities of fluid nodes itself and reads
 * rhoR_d/rhoB_d only on non-fluid nodes.  The particle kernels that follow
 * the density pass in the reference (make_fluid, compute_*_rho, partBB_CG*)
 * read densities of fluid nodes next to particle surfaces, so this kernel
 * forms them there from the populations of this step, and gives the
 * particle's own non-fluid nodes MINDENS = 0 (init_rhoR_CG / init_rhoB_CG,
 * kernels_fluid_CG.CUF:398-509).  Offsets: the template and the D3Q27 ring
 * of its surface nodes. */
/* the particle's own template node t: MINDENS (every template node is
 * non-fluid) */
/* an entry (dx, dy, dz, value) of a (4, n) int8 offset table as one word */
__device__ __forceinline__ uint32_t off_word(const int8_t* tbl, int t) {
  return reinterpret_cast<const uint32_t*>(tbl)[t];
}
__device__ __forceinline__ int off_c(uint32_t w, int k) { return (int)(int8_t)((w >> (8 * k)) & 0xffu); }

__device__ __forceinline__ void shell_zero(const PartArgs& a, int p, const int c[3], uint32_t w) {
  const size_t n = node_at(a, c[0], c[1], c[2], off_c(w, 0), off_c(w, 1), off_c(w, 2));
  if ((a.key_cur[n] >> 3) != (uint32_t)p) return;
  a.rho_r[n] = 0.0f;
  a.rho_b[n] = 0.0f;
  a.phi[n] = (0.0f - 0.0f) / (0.0f + 0.0f);
}

/* ring node (offset word w) around the particle's surface: densities of a
 * fluid node from this step's populations.  A node filled this step (fluid
 * now, not on the previous level) keeps the belt densities make_fluid gave
 * it. */
template <class S>
__device__ __forceinline__ void shell_ring(const PartArgs& a, const int c[3], uint32_t w) {
  const size_t n = node_at(a, c[0], c[1], c[2], off_c(w, 0), off_c(w, 1), off_c(w, 2));
  const uint8_t f = a.fl_cur[n], fo = a.fl_old[n];
  float rr, rb;
  rho_sum<S>(a.pop_src, a.N, n, a.rho0, a.rdec, &rr, &rb);   /* issued with the type loads */
  if (f != FL_FLUID || fo != FL_FLUID) return;
  a.rho_r[n] = rr;
  a.rho_b[n] = rb;
  a.phi[n] = (rr - rb) / (rr + rb);
}

template <int NT = TPB>
__device__ __forceinline__ void k_neigh_body(const PartArgs& a, int i);

/* shell densities, wetting and the link masks of the surface nodes in one
 * pass, one workgroup per particle (runs after make_fluid).  Every fluid
 * D3Q27 neighbour of a surface node p owns is in p's ring, so the wetting
 * after the barrier reads only densities this workgroup wrote (or make_fluid
 * wrote before the launch). */
template <class S>
__device__ __forceinline__ void k_shellwet_body(const PartArgs& a, int p) {
  /* the neighbour list the link pass uses (positions do not change before
     it) and the release of make_fluid's claims (every particle's fill is done) */
  k_neigh_body(a, p);
  k_fill_release_body(a, p);
  int c[3];
  centre(a.pos_cur, p, c);
  /* template and ring in one loop, so their loads are in flight together;
     the offset words of the next iteration are fetched ahead */
  const int nmax = max(a.ntmpl, a.nring);
  int t = threadIdx.x;
  uint32_t wt = t < a.ntmpl ? off_word(a.tmpl, t) : 0u, wr = t < a.nring ? off_word(a.ring, t) : 0u;
  for (; t < nmax; t += TPB) {
    const int tn = t + TPB;
    const uint32_t wtn = tn < a.ntmpl ? off_word(a.tmpl, tn) : 0u, wrn = tn < a.nring ? off_word(a.ring, tn) : 0u;
    if (t < a.ntmpl) shell_zero(a, p, c, wt);
    if (t < a.nring) shell_ring<S>(a, c, wr);
    wt = wtn;
    wr = wrn;
  }
  __syncthreads();
  for (int u = threadIdx.x; u < a.ntlist; u += TPB) surface_node<S>(a, p, c, u);
}

/* ---------------- shell pass with the particle's neighbourhood in LDS ----
 * The same work as k_shellwet_body, but the types and densities the surface
 * pass reads around each surface node (26 D3Q27 neighbours, all of them in
 * the particle's template or ring) come from a box of side 2 rmax + 1 in LDS
 * that the template and ring loop fill as they go, instead of being gathered
 * again from global memory: template nodes are non-fluid (their type is not
 * needed beyond that), ring nodes carry their type and, if fluid, the
 * densities this pass forms (or make_fluid's, for a node filled this step). */
struct ShellBox {
  float* rr;
  float* rb;
  uint8_t* fl;
  int h, w;
  __device__ __forceinline__ int at(int dx, int dy, int dz) const { return (dx + h) + w * ((dy + h) + w * (dz + h)); }
};

__device__ __forceinline__ ShellBox shell_box(const PartArgs& a) {
  extern __shared__ float shell_lds[];
  ShellBox b;
  b.h = a.rmax;
  b.w = 2 * a.rmax + 1;
  const int n = b.w * b.w * b.w;
  b.rr = shell_lds;
  b.rb = shell_lds + n;
  b.fl = reinterpret_cast<uint8_t*>(shell_lds + 2 * n);
  return b;
}

__device__ __forceinline__ void shell_zero_box(const PartArgs& a, int p, const int c[3], uint32_t w, const ShellBox& bx) {
  bx.fl[bx.at(off_c(w, 0), off_c(w, 1), off_c(w, 2))] = FL_DEAD;
  shell_zero(a, p, c, w);
}

template <class S>
__device__ __forceinline__ void shell_ring_box(const PartArgs& a, const int c[3], uint32_t w, const ShellBox& bx) {
  const size_t n = node_at(a, c[0], c[1], c[2], off_c(w, 0), off_c(w, 1), off_c(w, 2));
  const uint8_t f = a.fl_cur[n], fo = a.fl_old[n];
  float rr, rb;
  rho_sum<S>(a.pop_src, a.N, n, a.rho0, a.rdec, &rr, &rb);   /* issued with the type loads */
  const int b = bx.at(off_c(w, 0), off_c(w, 1), off_c(w, 2));
  bx.fl[b] = f;
  if (f != FL_FLUID) return;
  if (fo != FL_FLUID) {
    rr = a.rho_r[n];   /* filled this step: make_fluid's densities */
    rb = a.rho_b[n];
  } else {
    a.rho_r[n] = rr;
    a.rho_b[n] = rb;
    a.phi[n] = (rr - rb) / (rr + rb);
  }
  bx.rr[b] = rr;
  bx.rb[b] = rb;
}

/* surface_node with the neighbours' types and densities from the box */
template <class S>
__device__ __forceinline__ void surface_node_box(const PartArgs& a, int p, const int c[3], int t, const ShellBox& bx) {
  const int ox = a.tlist[4 * t], oy = a.tlist[4 * t + 1], oz = a.tlist[4 * t + 2];
  int gx, gy, gz;
  size_t n = node_at(a, c[0], c[1], c[2], ox, oy, oz, &gx, &gy, &gz);
  uint32_t* mslot = a.smask + (size_t)p * a.ntlist + t;
  const uint8_t fn = a.fl_cur[n];
  const uint32_t kn = a.key_cur[n];
  if ((fn != FL_LIST) | ((kn >> 3) != (uint32_t)p)) {
    *mslot = 0u;
    return;
  }
  uint32_t mask = 1u << 31;
#pragma unroll
  for (int l = 1; l < 19; ++l)
    if (bx.fl[bx.at(ox + EX(l), oy + EY(l), oz + EZ(l))] == FL_FLUID) mask |= 1u << (l - 1);
  *mslot = mask;
  if (a.wetting == LBM_WET_NONE) return;
  float rr, rb;
  if (a.wetting == LBM_WET_JANUS) {
    /* compute_virtual_rho, PART:299-352: sign of r.(q x q*) */
    double r[3] = {(double)(gx + 1) - a.pos_cur[13 * p], (double)(gy + 1) - a.pos_cur[13 * p + 1],
                   (double)(gz + 1) - a.pos_cur[13 * p + 2]};
    r[0] = min_image(r[0], a.nx);
    r[1] = min_image(r[1], a.ny);
    r[2] = min_image(r[2], a.nz);
    const double q0 = a.pos_cur[13 * p + 9], q1 = a.pos_cur[13 * p + 10], q2 = a.pos_cur[13 * p + 11],
                 q3 = a.pos_cur[13 * p + 12];
    const double ax = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    const double ay = 2.0 * (q1 * q2 + q0 * q3);
    const double az = 2.0 * (q1 * q3 - q0 * q2);
    const double ac = ax * r[0] + ay * r[1] + az * r[2];
    if (ac < 0.0) { rr = 0.0f; rb = 1.0f; } else { rr = 1.0f; rb = 0.0f; }
  } else {
    /* compute_densities_wall + compute_onebelt_density_2fl_weights, PART:267-297, 864-894 */
    float Rs = 0.0f, Bs = 0.0f, Ds = 0.0f;
#pragma unroll
    for (int l = 1; l < 27; ++l) {
      const int b = bx.at(ox + EX27(l), oy + EY27(l), oz + EZ27(l));
      if (bx.fl[b] == FL_FLUID) {
        Rs = fmaf(PD27(l), bx.rr[b], Rs);
        Bs = fmaf(PD27(l), bx.rb[b], Bs);
        Ds = Ds + PD27(l);
      }
    }
    if (Ds != 0.0f) {
      rr = Rs / Ds;
      rb = Bs / Ds;
    } else {
      float R2 = 0.0f, B2 = 0.0f;
      int D = 0;
      for (int l = 0; l < 98; ++l) {
        size_t m = wrap_node(a, gx + C_BELT2[l][0], gy + C_BELT2[l][1], gz + C_BELT2[l][2]);
        if (a.fl_cur[m] == FL_FLUID && a.fl_old[m] == FL_FLUID) {
          float mr, mb;
          belt_rho<S>(a, m, &mr, &mb);
          R2 = R2 + mr;
          B2 = B2 + mb;
          D = D + 1;
        }
      }
      if (D == 0) return;   /* fix_onebelt_density_twofluids is empty */
      rr = R2 / (float)D;
      rb = B2 / (float)D;
    }
  }
  a.rho_r[n] = rr;
  a.rho_b[n] = rb;
  a.phi[n] = (rr - rb) / (rr + rb);
}

template <class S, int NT>
__device__ __forceinline__ void k_shellbox_body(const PartArgs& a, int p) {
  k_neigh_body<NT>(a, p);
  k_fill_release_body<NT>(a, p);
  const ShellBox bx = shell_box(a);
  int c[3];
  centre(a.pos_cur, p, c);
  const int nmax = max(a.ntmpl, a.nring);
  int t = threadIdx.x;
  uint32_t wt = t < a.ntmpl ? off_word(a.tmpl, t) : 0u, wr = t < a.nring ? off_word(a.ring, t) : 0u;
  for (; t < nmax; t += NT) {
    const int tn = t + NT;
    const uint32_t wtn = tn < a.ntmpl ? off_word(a.tmpl, tn) : 0u, wrn = tn < a.nring ? off_word(a.ring, tn) : 0u;
    if (t < a.ntmpl) shell_zero_box(a, p, c, wt, bx);
    if (t < a.nring) shell_ring_box<S>(a, c, wr, bx);
    wt = wtn;
    wr = wrn;
  }
  __syncthreads();
  for (int u = threadIdx.x; u < a.ntlist; u += NT) surface_node_box<S>(a, p, c, u, bx);
}

/* ---------------- bounce-back on surface links ---------------------------
 * After the collide-stream kernel the post-collision value f_hi(ii) that the
 * fluid node ii pushed into the surface node s sits in dst(s, hi).  One pass
 * over the surface nodes, each handled by the particle that owns it:
 *
 *  - bounce-back (particle_to_node_bounce_back_bc2_phase1 + phase2,
 *    kernels_fluid_PART.CUF:2153-2312): dst(ii, lo) = dst(s, hi) minus the
 *    moving-wall term 6 w rho_k(ii) (e.v_b) of every particle whose centre
 *    lies within r^2 < rmax^2 of s (partBB_CG_p2n2 visits each of them), in
 *    ascending particle order; every slot has one writer, no atomics;
 *  - momentum exchange (node_to_particle_bounce_back_bc2, PART:1874-2113) on
 *    the same links for every such particle q: the owner's own share is
 *    summed in FP64 over the workgroup into myf/myt; the share of another
 *    particle q (two particles in contact) is summed the same way into
 *    contrib(slot of q) and added to q's force, in ascending owner order, by
 *    the rigid-body update.  The sums never depend on scheduling.
 */

/* integer minimum-image squared distance of node (gx,gy,gz) (0-based) from
 * the centre c (1-based lattice coordinates) */
__device__ __forceinline__ int ball_r2(const PartArgs& a, const int c[3], int gx, int gy, int gz) {
  int d[3] = {gx + 1 - c[0], gy + 1 - c[1], gz + 1 - c[2]};
  const int n[3] = {a.nx, a.ny, a.nz};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (2 * d[k] > n[k]) d[k] -= n[k];
    else if (2 * d[k] < -n[k]) d[k] += n[k];
  }
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

/* what the link pass reads around one surface node s after the collide, for
 * each link, i.e. each opposite pair il = (lo, hi) = (2 il - 1, 2 il) with
 * exactly one fluid neighbour ii = s + e_lin (from the mask the surface pass
 * recorded): the densities of ii and the value ii pushed into s,
 * dst(s, lout).  All of them are issued in one batch. */
template <class S>
struct LinkData {
  uint32_t mask;           /* D3Q19 fluid-neighbour mask of s */
  uint32_t nb[10];         /* ii of link il                   */
  float rr[10], rb[10], pr[10], pb[10];
  __device__ static bool is_link(uint32_t m, int il) {
    return ((m >> (2 * il - 2)) & 1u) != ((m >> (2 * il - 1)) & 1u);
  }
  __device__ static bool lo_is_fluid(uint32_t m, int il) { return (m >> (2 * il - 2)) & 1u; }
  __device__ __forceinline__ void load(const PartArgs& a, size_t s, int gx, int gy, int gz, uint32_t m) {
    mask = m;
#pragma unroll
    for (int il = 1; il <= 9; ++il) {
      if (!is_link(m, il)) continue;
      const bool fi = lo_is_fluid(m, il);
      const int lin = fi ? 2 * il - 1 : 2 * il;
      const int lout = fi ? 2 * il : 2 * il - 1;
      nb[il] = (uint32_t)wrap_node(a, gx + EX(lin), gy + EY(lin), gz + EZ(lin));
      rr[il] = a.rho_r[nb[il]];
      rb[il] = a.rho_b[nb[il]];
      ld_pair<S>(a.pop_dst, a.N, lout, s, a.rho0, a.rdec, pr[il], pb[il]);
    }
  }
};

/* neighbour-list entry cached in LDS */
struct NbrCache {
  int q[MAXNB];
  int c[MAXNB][3];
  double v[MAXNB][3];
  double o[MAXNB][3];
};

/* wall velocity v_b of particle k (cache slot) at the link lin of s, with
 * the particle's lever R r/|r| to s (rv, formed once per node): v + omega x
 * (rv + e_lin / 2) when rotating (PART:1900-1927, 1954-1962); rt is the
 * torque arm rv + e_lin / 2 */
__device__ __forceinline__ void link_velocity(const PartArgs& a, const NbrCache& nc, int k, const double rv[3],
                                              int lin, double v[3], double rt[3]) {
  v[0] = nc.v[k][0]; v[1] = nc.v[k][1]; v[2] = nc.v[k][2];
  rt[0] = rt[1] = rt[2] = 0.0;
  if (a.lrotate) {
    double cr[3];
    rt[0] = rv[0] + 0.5 * EX(lin); rt[1] = rv[1] + 0.5 * EY(lin); rt[2] = rv[2] + 0.5 * EZ(lin);
    const double ot[3] = {nc.o[k][0], nc.o[k][1], nc.o[k][2]};
    cross3(ot, rt, cr);
    v[0] += cr[0]; v[1] += cr[1]; v[2] += cr[2];
  }
}

/* momentum given to particle k by link il of s (addForce at :1964-1972 /
 * :2049-2057, torque :1999-2001), from the pushed values before the
 * bounce-back */
template <class S>
__device__ __forceinline__ void link_force_one(const PartArgs& a, const LinkData<S>& ld, int il, const double v[3],
                                               const double rt[3], double acc[6]) {
  const bool fi = LinkData<S>::lo_is_fluid(ld.mask, il);
  const int lout = fi ? 2 * il : 2 * il - 1;
  const float ev = (float)(EX(lout) * v[0] + EY(lout) * v[1] + EZ(lout) * v[2]);
  const float wr = W(lout) * BUZ * ld.rr[il] * ev;
  const float wb = W(lout) * BUZ * ld.rb[il] * ev;
  const double f2p = (double)((2.0f * ld.pr[il] - wr) + (2.0f * ld.pb[il] - wb));
  const double ft[3] = {f2p * EX(lout), f2p * EY(lout), f2p * EZ(lout)};
  acc[0] += ft[0]; acc[1] += ft[1]; acc[2] += ft[2];
  if (a.lrotate) {
    double tq[3];
    cross3(rt, ft, tq);
    acc[3] += tq[0] * abs(EX(lout));
    acc[4] += tq[1] * abs(EY(lout));
    acc[5] += tq[2] * abs(EZ(lout));
  }
}

/* lever of cache slot k to node (gx, gy, gz), when rotating */
__device__ __forceinline__ void slot_lever(const PartArgs& a, const NbrCache& nc, int k, int gx, int gy, int gz,
                                           double rv[3]) {
  rv[0] = rv[1] = rv[2] = 0.0;
  if (a.lrotate) lever(a, a.pos_cur, nc.q[k], gx, gy, gz, true, rv);
}

template <class S>
__device__ __forceinline__ void k_links_body(const PartArgs& a, int p) {
  __shared__ NbrCache nc;
  __shared__ int self_slot;
  __shared__ unsigned shared_slots;   /* slots of the other particles covering some surface node */
  const int nnb = min(a.nnbr[p], MAXNB);
  if (threadIdx.x == 0) {
    self_slot = -1;
    shared_slots = 0u;
  }
  __syncthreads();
  if ((int)threadIdx.x < nnb) {
    const int k = threadIdx.x;
    const int q = a.nbr[MAXNB * p + k];
    nc.q[k] = q;
    centre(a.pos_cur, q, nc.c[k]);
    for (int d = 0; d < 3; ++d) nc.v[k][d] = a.pos_cur[13 * q + 3 + d];
    double ot[3] = {0.0, 0.0, 0.0};
    if (a.lrotate) omega_lab(a.pos_cur, q, ot);
    for (int d = 0; d < 3; ++d) nc.o[k][d] = ot[d];
    if (q == p) self_slot = k;
  }
  __syncthreads();
  int c[3];
  centre(a.pos_cur, p, c);
  const int r2max = a.rmax * a.rmax;
  const int kp = self_slot;
  /* a full neighbour list (LBM_ERR_LIST_OVERFLOW raised by k_neigh) may miss
     the particle itself: leave its link sums untouched then */
  if (kp < 0) return;
  const uint32_t* msk = a.smask + (size_t)p * a.ntlist;
  double own[6] = {0, 0, 0, 0, 0, 0};
  unsigned others = 0u;   /* slots of the other particles covering a surface node of this thread */
  for (int t = threadIdx.x; t < a.ntlist; t += TPB) {
    const uint32_t m = msk[t];
    if (!(m >> 31)) continue;
    int gx, gy, gz;
    const size_t s = node_at(a, c[0], c[1], c[2], a.tlist[4 * t], a.tlist[4 * t + 1], a.tlist[4 * t + 2], &gx, &gy, &gz);
    LinkData<S> ld;
    ld.load(a, s, gx, gy, gz, m);
    unsigned cover = 0;   /* cache slots of the particles whose ball holds s */
    for (int k = 0; k < nnb; ++k)
      if (ball_r2(a, nc.c[k], gx, gy, gz) < r2max) cover |= 1u << k;
    others |= cover & ~(1u << kp);
    /* bounce-back: dst(ii, lin) = dst(s, lout) minus the moving-wall term of
       every covering particle, in ascending order; the owner's link momentum
       on the way */
    float fR[10], fB[10];
#pragma unroll
    for (int il = 1; il <= 9; ++il) {
      fR[il] = ld.pr[il];
      fB[il] = ld.pb[il];
    }
    for (int k = 0; k < nnb; ++k) {
      if (!(cover >> k & 1u)) continue;
      double rv[3];
      slot_lever(a, nc, k, gx, gy, gz, rv);
#pragma unroll
      for (int il = 1; il <= 9; ++il) {
        if (!LinkData<S>::is_link(m, il)) continue;
        const bool fi = LinkData<S>::lo_is_fluid(m, il);
        const int lin = fi ? 2 * il - 1 : 2 * il;
        const int lout = fi ? 2 * il : 2 * il - 1;
        double v[3], rt[3];
        link_velocity(a, nc, k, rv, lin, v, rt);
        const float ev = (float)(EX(lout) * v[0] + EY(lout) * v[1] + EZ(lout) * v[2]);
        fR[il] = fR[il] - W(lout) * BUZ * ld.rr[il] * ev;
        fB[il] = fB[il] - W(lout) * BUZ * ld.rb[il] * ev;
        if (k == kp) link_force_one<S>(a, ld, il, v, rt, own);
      }
    }
#pragma unroll
    for (int il = 1; il <= 9; ++il) {
      if (!LinkData<S>::is_link(m, il)) continue;
      const int lin = LinkData<S>::lo_is_fluid(m, il) ? 2 * il - 1 : 2 * il;
      st_pair<S>(a.pop_dst, a.N, lin, ld.nb[il], fR[il], fB[il], a.rho0, a.renc);
    }
  }
  block_reduce6(own);
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; ++d) a.myf[3 * p + d] = own[d];
    if (a.lrotate)
      for (int d = 0; d < 3; ++d) a.myt[3 * p + d] = own[3 + d];
  }
  /* shares of the other particles in contact with this one: a pass over the
     surface for each particle whose ball covers one of its nodes, zero for
     the rest of the list */
  if (others) atomicOr(&shared_slots, others);
  __syncthreads();
  const unsigned sh = shared_slots;
  for (int k = 0; k < nnb; ++k) {
    if (k == kp) continue;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (sh >> k & 1u) {
      for (int t = threadIdx.x; t < a.ntlist; t += TPB) {
        const uint32_t m = msk[t];
        if (!(m >> 31)) continue;
        int gx, gy, gz;
        const size_t s =
            node_at(a, c[0], c[1], c[2], a.tlist[4 * t], a.tlist[4 * t + 1], a.tlist[4 * t + 2], &gx, &gy, &gz);
        if (ball_r2(a, nc.c[k], gx, gy, gz) >= r2max) continue;
        LinkData<S> ld;
        ld.load(a, s, gx, gy, gz, m);
        double rv[3];
        slot_lever(a, nc, k, gx, gy, gz, rv);
#pragma unroll
        for (int il = 1; il <= 9; ++il) {
          if (!LinkData<S>::is_link(m, il)) continue;
          const int lin = LinkData<S>::lo_is_fluid(m, il) ? 2 * il - 1 : 2 * il;
          double v[3], rt[3];
          link_velocity(a, nc, k, rv, lin, v, rt);
          link_force_one<S>(a, ld, il, v, rt, acc);
        }
      }
      block_reduce6(acc);
    }
    if (threadIdx.x == 0)
      for (int d = 0; d < 6; ++d) a.contrib[6 * ((size_t)MAXNB * p + k) + d] = acc[d];
  }
}

/* neighbour list: particles j (i included) whose integer centres lie within
 * 2 rmax of particle i, in ascending index order (wave ballots keep the
 * order), so k_links finds every particle whose ball reaches i's surface.
 * More than MAXNB raises LBM_ERR_LIST_OVERFLOW, as makeListAtomsGPU raises
 * on a full tile list (kernels_fluid_PART.CUF:2436-2439). */
template <int NT>
__device__ __forceinline__ void k_neigh_body(const PartArgs& a, int i) {
  int ci[3];
  centre(a.pos_cur, i, ci);
  __shared__ int wcount[NT / 64];
  __shared__ int base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int lim = 4 * a.rmax * a.rmax;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int j0 = 0; j0 < a.np; j0 += NT) {
    const int j = j0 + threadIdx.x;
    bool hit = false;
    if (j < a.np) {
      int cj[3];
      centre(a.pos_cur, j, cj);
      int d[3] = {cj[0] - ci[0], cj[1] - ci[1], cj[2] - ci[2]};
      const int n[3] = {a.nx, a.ny, a.nz};
      for (int k = 0; k < 3; ++k) {
        d[k] = ((d[k] % n[k]) + n[k]) % n[k];
        if (2 * d[k] > n[k]) d[k] -= n[k];
      }
      hit = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < lim;
    }
    const unsigned long long m = __ballot(hit);
    if (lane == 0) wcount[wid] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wid; ++w) off += wcount[w];
    if (hit) {
      const int pos = off + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < MAXNB) a.nbr[MAXNB * i + pos] = j;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 0; w < NT / 64; ++w) base += wcount[w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.nnbr[i] = base;
    if (base > MAXNB) raise_error(a.status, LBM_ERR_LIST_OVERFLOW, a.step);
  }
}

/* ---------------- pair forces (inter_force_particles, PART:1052-1068) ---- */
struct PairParams {
  double hz_k, hz_rmin, hz_rcap, gmin;
  double rpar, lubfactor, rparcut, rparcap, mxrsqcut, visc;
  double closure_r;   /* centre distance below which the closure may act */
  int hertz, lub;
};

/* learned near-contact correction of pair (i, j) (the closure block layout
 * of LBM_CLOSURE_NFLOAT): adds the force and torque on i and returns true
 * when the pair's features lie inside the training box */
__device__ bool closure_pair(const PartArgs& a, int i, int j, double r, const double n[3], double radius,
                             double f[6]) {
  const float* cb = a.closure;
  const float* mean = cb + LBM_MLP_NPARAM;
  const float* scale = mean + 4;
  const float* oscale = scale + 4;
  const float* lo = oscale + 3;
  const float* hi = lo + 4;
  double oi[3] = {0, 0, 0}, oj[3] = {0, 0, 0};
  if (a.lrotate) {
    omega_lab(a.pos_cur, i, oi);
    omega_lab(a.pos_cur, j, oj);
  }
  const double os[3] = {oi[0] + oj[0], oi[1] + oj[1], oi[2] + oj[2]};
  double ocn[3];
  cross3(os, n, ocn);
  double u[3];
  for (int d = 0; d < 3; ++d) u[d] = a.pos_cur[13 * i + 3 + d] - a.pos_cur[13 * j + 3 + d] + radius * ocn[d];
  const double vn = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
  double t[3] = {u[0] - vn * n[0], u[1] - vn * n[1], u[2] - vn * n[2]};
  const double vt = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  if (vt > 1e-300)
    for (int d = 0; d < 3; ++d) t[d] /= vt;
  else
    t[0] = t[1] = t[2] = 0.0;
  const float feat[4] = {(float)((r - 2.0 * radius) / radius), (float)vn, (float)vt, (float)radius};
  for (int k = 0; k < 4; ++k)
    if (!(feat[k] >= lo[k] && feat[k] <= hi[k])) return false;
  float in[4], out[3];
  for (int k = 0; k < 4; ++k) in[k] = (feat[k] - mean[k]) / scale[k];
  lbm_mlp_eval(cb, in, out);
  const double Fn = (double)(out[0] * oscale[0]), Ft = (double)(out[1] * oscale[1]), Tq = (double)(out[2] * oscale[2]);
  double nt[3];
  cross3(n, t, nt);
  for (int d = 0; d < 3; ++d) {
    f[d] -= Fn * n[d] + Ft * t[d];
    f[3 + d] += Tq * nt[d];
  }
  return true;
}

__device__ __forceinline__ void k_pairs_body(const PartArgs& a, int i, const PairParams& q) {
  /* one workgroup per particle i, its threads over the partners j; FP64
     partial forces reduced in a fixed order */
  double f[6] = {0, 0, 0, 0, 0, 0};
  const double xi = a.pos_cur[13 * i], yi = a.pos_cur[13 * i + 1], zi = a.pos_cur[13 * i + 2];
  const double cutsq = a.closure ? fmax(q.mxrsqcut, q.closure_r * q.closure_r) : q.mxrsqcut;
  for (int j = threadIdx.x; j < a.np; j += TPB) {
    if (j == i) continue;
    /* addHertzian, PART:1385-1503, with the pair vector from i to j */
    double xdf = a.pos_cur[13 * j] - xi, ydf = a.pos_cur[13 * j + 1] - yi, zdf = a.pos_cur[13 * j + 2] - zi;
    xdf = min_image(xdf, a.nx);
    ydf = min_image(ydf, a.ny);
    zdf = min_image(zdf, a.nz);
    const double rsq = xdf * xdf + ydf * ydf + zdf * zdf;
    if (rsq > cutsq) continue;
    const double rrr = sqrt(rsq);
    if (q.hertz && rrr <= q.hz_rmin) {
      const double g = rrr <= q.hz_rcap ? q.gmin : 2.5 * q.hz_k * pow(q.hz_rmin - rrr, 1.5);
      f[0] -= g * xdf / rrr; f[1] -= g * ydf / rrr; f[2] -= g * zdf / rrr;
    }
    /* near contact: the learned correction where it was trained, the
       reference's lubrication term elsewhere */
    bool learned = false;
    if (a.closure && rrr < q.closure_r) {
      const double n[3] = {xdf / rrr, ydf / rrr, zdf / rrr};
      learned = closure_pair(a, i, j, rrr, n, q.rpar * 0.5, f);
    }
    if (!learned && q.lub && rrr <= q.rparcut) {
      const double ux = xdf / rrr, uy = ydf / rrr, uz = zdf / rrr;
      const double rr = rrr <= q.rparcap ? q.rparcap : rrr;
      const double dotuv = ux * (a.pos_cur[13 * i + 3] - a.pos_cur[13 * j + 3]) +
                           uy * (a.pos_cur[13 * i + 4] - a.pos_cur[13 * j + 4]) +
                           uz * (a.pos_cur[13 * i + 5] - a.pos_cur[13 * j + 5]);
      const double s = q.visc * q.lubfactor * dotuv * (1.0 / (rr - q.rpar) - 1.0);
      f[0] -= s * ux; f[1] -= s * uy; f[2] -= s * uz;
    }
  }
  block_reduce6(f);
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; ++d) a.myf2[3 * i + d] += f[d];
    if (a.lrotate && a.closure)
      for (int d = 0; d < 3; ++d) a.myt2[3 * i + d] += f[3 + d];
  }
}

/* ---------------- rigid-body update --------------------------------------
 * time_step_force_particles -> addForce_particle_bounce_back -> nve_lf ->
 * rot_midstep_impl_lf (PART:1098-1166, 1266-1382, 1506-1614, 1621-1827),
 * in FP64. */
struct IntParams {
  double rmass, rotin, max_vel;
  double ext_f[3], ext_t[3];
};

__device__ __forceinline__ void k_integrate_body(const PartArgs& a, int i, const IntParams& q) {
  const double* P = a.pos_cur + 13 * i;
  double* Pn = a.pos_new + 13 * i;
  double* Fc = a.f_cur + 12 * i;
  double* Fo = a.f_old + 12 * i;
  /* link momentum: own surface, then the surfaces of particles in contact
     (their owners' shares for i, ascending owner index) */
  double fb[6] = {a.myf[3 * i], a.myf[3 * i + 1], a.myf[3 * i + 2], 0.0, 0.0, 0.0};
  if (a.lrotate)
    for (int d = 0; d < 3; ++d) fb[3 + d] = a.myt[3 * i + d];
  const int nni = min(a.nnbr[i], MAXNB);
  for (int k = 0; k < nni; ++k) {
    const int p = a.nbr[MAXNB * i + k];
    if (p == i) continue;
    const int nnp = min(a.nnbr[p], MAXNB);
    for (int j = 0; j < nnp; ++j)
      if (a.nbr[MAXNB * p + j] == i)
        for (int d = 0; d < 6; ++d) fb[d] += a.contrib[6 * ((size_t)MAXNB * p + j) + d];
  }
  /* addForce_particle_bounce_back: F(t) = F_other + (F_bb(t+1/2)+F_bb(t-1/2))/2 */
  for (int d = 0; d < 3; ++d) {
    Fc[3 + d] = fb[d];
    Fc[d] = a.myf2[3 * i + d] + (Fc[3 + d] + Fo[3 + d]) * 0.5;
  }
  if (a.lrotate)
    for (int d = 0; d < 3; ++d) {
      Fc[9 + d] = fb[3 + d];
      Fc[6 + d] = a.myt2[3 * i + d] + (Fc[9 + d] + Fo[9 + d]) * 0.5;
    }
  /* nve_lf */
  const double fx = Fc[0], fy = Fc[1], fz = Fc[2];
  if (fabs(fx) * q.rmass > 0.5 || fabs(fy) * q.rmass > 0.5 || fabs(fz) * q.rmass > 0.5)
    raise_error(a.status, LBM_ERR_BIG_FORCE, a.step);
  const double bx = P[3] + q.rmass * fx, by = P[4] + q.rmass * fy, bz = P[5] + q.rmass * fz;
  if (sqrt(bx * bx + by * by + bz * bz) > q.max_vel) raise_error(a.status, LBM_ERR_EXTREME_VEL, a.step);
  double xn[3] = {P[0] + bx, P[1] + by, P[2] + bz};
  const int nn[3] = {a.nx, a.ny, a.nz};
  for (int d = 0; d < 3; ++d) {
    /* pbc_images_centered about the box centre (n+1)/2, PART:2124-2151 */
    const double cc = (double)(nn[d] + 1) * 0.5;
    const double s = xn[d] - cc;
    xn[d] = s - (double)nn[d] * (double)nint(s / (double)nn[d]) + cc;
    if (xn[d] < 0.5 || xn[d] > (double)nn[d] + 0.5) raise_error(a.status, LBM_ERR_OUT_OF_BOX, a.step);
  }
  if (a.lrotate && q.rmass > 0.0) {
    const double I = q.rotin;
    double ox = P[6], oy = P[7], oz = P[8];
    double q0 = P[9], q1 = P[10], q2 = P[11], q3 = P[12];
    const double tqx = Fc[6], tqy = Fc[7], tqz = Fc[8];
    /* quat_2_rotmat, PART:1830-1844 */
    const double m1 = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, m2 = 2.0 * (q1 * q2 - q0 * q3),
                 m3 = 2.0 * (q1 * q3 + q0 * q2), m4 = 2.0 * (q1 * q2 + q0 * q3),
                 m5 = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, m6 = 2.0 * (q2 * q3 - q0 * q1),
                 m7 = 2.0 * (q1 * q3 - q0 * q2), m8 = 2.0 * (q2 * q3 + q0 * q1),
                 m9 = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
    double opx = ox, opy = oy, opz = oz;
    const double trx = (tqx * m1 + tqy * m4 + tqz * m7) / I + (I - I) * opy * opz / I;
    const double tr_y = (tqx * m2 + tqy * m5 + tqz * m8) / I + (I - I) * opz * opx / I;
    const double trz = (tqx * m3 + tqy * m6 + tqz * m9) / I + (I - I) * opx * opy / I;
    opx = ox + trx * 0.5; opy = oy + tr_y * 0.5; opz = oz + trz * 0.5;
    double dq0 = (-q1 * opx - q2 * opy - q3 * opz) * 0.5;
    double dq1 = (q0 * opx - q3 * opy + q2 * opz) * 0.5;
    double dq2 = (q3 * opx + q0 * opy - q1 * opz) * 0.5;
    double dq3 = (-q2 * opx + q1 * opy + q0 * opz) * 0.5;
    double qn0 = q0 + dq0 * 0.5, qn1 = q1 + dq1 * 0.5, qn2 = q2 + dq2 * 0.5, qn3 = q3 + dq3 * 0.5;
    double rn = 1.0 / sqrt(qn0 * qn0 + qn1 * qn1 + qn2 * qn2 + qn3 * qn3);
    qn0 *= rn; qn1 *= rn; qn2 *= rn; qn3 *= rn;
    ox += trx; oy += tr_y; oz += trz;
    Pn[6] = ox; Pn[7] = oy; Pn[8] = oz;
    opx = ox; opy = oy; opz = oz;
    int itq = 0;
    double eps = 1.0e9;
    while (itq < 20 && eps > 1.0e-24) {
      ++itq;
      dq0 = (-qn1 * opx - qn2 * opy - qn3 * opz) * 0.5;
      dq1 = (qn0 * opx - qn3 * opy + qn2 * opz) * 0.5;
      dq2 = (qn3 * opx + qn0 * opy - qn1 * opz) * 0.5;
      dq3 = (-qn2 * opx + qn1 * opy + qn0 * opz) * 0.5;
      double a0 = qn0 + dq0 * 0.5, a1 = qn1 + dq1 * 0.5, a2 = qn2 + dq2 * 0.5, a3 = qn3 + dq3 * 0.5;
      rn = 1.0 / sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3);
      a0 *= rn; a1 *= rn; a2 *= rn; a3 *= rn;
      eps = (a0 - qn0) * (a0 - qn0) + (a1 - qn1) * (a1 - qn1) + (a2 - qn2) * (a2 - qn2) + (a3 - qn3) * (a3 - qn3);
      qn0 = a0; qn1 = a1; qn2 = a2; qn3 = a3;
    }
    q0 += dq0; q1 += dq1; q2 += dq2; q3 += dq3;
    rn = 1.0 / sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    Pn[9] = q0 * rn; Pn[10] = q1 * rn; Pn[11] = q2 * rn; Pn[12] = q3 * rn;
  } else {
    for (int c = 6; c < 13; ++c) Pn[c] = P[c];
  }
  Pn[0] = xn[0]; Pn[1] = xn[1]; Pn[2] = xn[2];
  Pn[3] = bx; Pn[4] = by; Pn[5] = bz;
  /* reset accumulators, PART:1117-1145 */
  for (int d = 0; d < 3; ++d) {
    Fo[d] = q.ext_f[d];
    Fo[3 + d] = 0.0;
    a.myf[3 * i + d] = 0.0;
    a.myf2[3 * i + d] = q.ext_f[d];
  }
  if (a.lrotate)
    for (int d = 0; d < 3; ++d) {
      Fo[6 + d] = q.ext_t[d];
      Fo[9 + d] = 0.0;
      a.myt[3 * i + d] = 0.0;
      a.myt2[3 * i + d] = q.ext_t[d];
    }
}

/* ---------------- kernels: one box, and many boxes per launch ------------
 * Single box: workgroup (or thread) p handles particle p of the box.
 * Batched: item g = (box, particle) over all boxes; the per-box argument
 * blocks sit in device memory for both time-level parities. */
struct BatchStep {
  int step;
  int lag_from_pops;
};

__device__ __forceinline__ PartArgs batch_args(const PartArgs* __restrict__ args, int box, BatchStep b) {
  PartArgs a = args[box];
  a.step = b.step;
  a.lag_from_pops = b.lag_from_pops;
  return a;
}

#define LBM_PART_KERNEL(NAME)                                                                         \
  __global__ __launch_bounds__(TPB) void NAME(PartArgs a) { NAME##_body(a, blockIdx.x); }          \
  __global__ __launch_bounds__(TPB) void NAME##_batch(const PartArgs* __restrict__ args,           \
                                                      const int2* __restrict__ items, BatchStep b) { \
    const int2 it = items[blockIdx.x];                                                                \
    NAME##_body(batch_args(args, it.x, b), it.y);                                                     \
  }
#define LBM_PART_KERNEL_S(NAME)                                                                       \
  template <class S>                                                                                  \
  __global__ __launch_bounds__(TPB) void NAME(PartArgs a) { NAME##_body<S>(a, blockIdx.x); }       \
  template <class S>                                                                                  \
  __global__ __launch_bounds__(TPB) void NAME##_batch(const PartArgs* __restrict__ args,           \
                                                      const int2* __restrict__ items, BatchStep b) { \
    const int2 it = items[blockIdx.x];                                                                \
    NAME##_body<S>(batch_args(args, it.x, b), it.y);                                                  \
  }

#define LBM_PART_KERNEL_SN(NAME, NT)                                                                  \
  template <class S>                                                                                  \
  __global__ __launch_bounds__(NT) void NAME(PartArgs a) { NAME##_body<S, NT>(a, blockIdx.x); }    \
  template <class S>                                                                                  \
  __global__ __launch_bounds__(NT) void NAME##_batch(const PartArgs* __restrict__ args,            \
                                                     const int2* __restrict__ items, BatchStep b) {  \
    const int2 it = items[blockIdx.x];                                                                \
    NAME##_body<S, NT>(batch_args(args, it.x, b), it.y);                                              \
  }

LBM_PART_KERNEL(k_unstamp)
LBM_PART_KERNEL(k_stamp)
LBM_PART_KERNEL(k_resolve)
LBM_PART_KERNEL(k_fill_release)
LBM_PART_KERNEL(k_neigh)
LBM_PART_KERNEL_SN(k_erase, NT_BALL)
LBM_PART_KERNEL_SN(k_fill_apply, NT_BALL)
LBM_PART_KERNEL_S(k_wetting)
LBM_PART_KERNEL_S(k_shellwet)
LBM_PART_KERNEL_SN(k_shellbox, NT_SHELL)
LBM_PART_KERNEL_S(k_links)

/* pair forces of particle i, then (thread 0) its rigid-body update: the
 * update needs only i's own sums and the link shares k_links wrote */
__device__ __forceinline__ void k_forces_body(const PartArgs& a, int i, const PairParams& pp, const IntParams& ip) {
  k_pairs_body(a, i, pp);
  if (threadIdx.x == 0) k_integrate_body(a, i, ip);
}
__global__ __launch_bounds__(TPB) void k_forces(PartArgs a, PairParams pp, IntParams ip) {
  k_forces_body(a, blockIdx.x, pp, ip);
}
__global__ __launch_bounds__(TPB) void k_forces_batch(const PartArgs* __restrict__ args,
                                                      const PairParams* __restrict__ pps,
                                                      const IntParams* __restrict__ ips,
                                                      const int2* __restrict__ items, BatchStep b) {
  const int2 it = items[blockIdx.x];
  k_forces_body(batch_args(args, it.x, b), it.y, pps[it.x], ips[it.x]);
}

PartArgs make_pargs(lbm_ctx* c) {
  PartArgs a;
  a.nx = c->nx; a.ny = c->ny; a.nz = c->nz;
  a.N = c->N;
  a.np = c->np;
  a.cur = c->cur;
  a.pos_cur = c->d_pos[c->cur];
  a.pos_old = c->d_pos[1 - c->cur];
  a.pos_new = c->d_pos[1 - c->cur];
  a.f_cur = c->d_force[c->cur];
  a.f_old = c->d_force[1 - c->cur];
  a.myf = c->d_myf; a.myf2 = c->d_myf2; a.myt = c->d_myt; a.myt2 = c->d_myt2;
  a.fl_cur = c->flags[c->cur];
  a.fl_old = c->flags[1 - c->cur];
  a.key_cur = c->d_key[c->cur];
  a.stamp_cur = c->d_stamp[c->cur];
  a.restamp = c->d_restamp;
  a.claim = c->d_claim;
  a.tmpl = c->d_tmpl; a.ntmpl = c->ntmpl;
  a.ring = c->d_ring; a.nring = c->nring;
  a.tlist = c->d_tlist; a.ntlist = c->ntlist;
  a.nbr = c->d_nbr; a.nnbr = c->d_nnbr; a.contrib = c->d_contrib; a.smask = c->d_smask;
  a.rmax = c->rmax;
  a.ball = c->d_ball; a.nball = c->nball;
  a.rho_r = c->rho_r; a.rho_b = c->rho_b; a.phi = c->phi;
  a.pop_src = c->pop[c->cur];
  a.pop_dst = c->pop[1 - c->cur];
  a.rho0 = c->storage == LBM_STORE_F32 ? 0.0f : c->cfg.store_rho0;
  a.rdec = std::ldexp(1.0f, 118 + rest_offset(c->cfg));
  a.renc = std::ldexp(1.0f, -118 - rest_offset(c->cfg));
  a.radius = c->cfg.radius;
  a.lrotate = c->cfg.lrotate;
  a.wetting = c->cfg.wetting;
  a.closure = c->d_closure;
  a.rho_from_pops = c->variant == LBM_KERNEL_FUSED ? 1 : 0;
  a.lag_from_pops = (c->variant == LBM_KERNEL_FUSED && c->step > 1) ? 1 : 0;
  a.status = c->d_status;
  a.step = (int)c->step;
  return a;
}

int last_err() { return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP; }

}  // namespace

#define LBM_STORAGE_SWITCH(st, BODY)                                  \
  switch (st) {                                                       \
    case LBM_STORE_F32: { using S = float; BODY; } break;           \
    case LBM_STORE_F16: { using S = __half; BODY; } break;          \
    default: { using S = __hip_bfloat16; BODY; } break;             \
  }

int lbm_particles_upload_belt2() {
  int8_t t[98][3];
  int l = 0;
  for (int k = -2; k <= 2; ++k)
    for (int j = -2; j <= 2; ++j)
      for (int i = -2; i <= 2; ++i)
        if (i == -2 || i == 2 || j == -2 || j == 2 || k == -2 || k == 2) {
          t[l][0] = (int8_t)i; t[l][1] = (int8_t)j; t[l][2] = (int8_t)k;
          ++l;
        }
  return hipMemcpyToSymbol(HIP_SYMBOL(C_BELT2), t, sizeof t) == hipSuccess ? 0 : LBM_ERR_HIP;
}

/* workgroups per particle for the chunked kernels: about 4 entries of the
 * largest offset list (template or ball) per lane */
unsigned lbm_particles_chunks(int ntmpl, int nball) {
  const int n = ntmpl > nball ? ntmpl : nball;
  const int c = (n + 4 * TPB - 1) / (4 * TPB);
  return (unsigned)(c < 1 ? 1 : c);
}

int lbm_particles_flags(lbm_ctx* c, int level) {
  PartArgs a = make_pargs(c);
  a.pos_cur = c->d_pos[level];
  a.fl_cur = c->flags[level];
  a.key_cur = c->d_key[level];
  a.stamp_cur = c->d_stamp[level];
  const dim3 g((unsigned)c->np, lbm_particles_chunks(c->ntmpl, c->nball));
  hipLaunchKernelGGL(k_unstamp, g, dim3(TPB), 0, c->stream, a);
  hipLaunchKernelGGL(k_stamp, g, dim3(TPB), 0, c->stream, a);
  hipLaunchKernelGGL(k_resolve, g, dim3(TPB), 0, c->stream, a);
  return last_err();
}

int lbm_particles_erase(lbm_ctx* c) {
  PartArgs a = make_pargs(c);
  LBM_STORAGE_SWITCH(c->storage, hipLaunchKernelGGL(k_erase<S>, dim3(c->np), dim3(NT_BALL), 0, c->stream, a))
  return last_err();
}

/* make_fluid; the nodes were claimed by the erase pass, and the claims are
 * released here or, in the single-pass step, by the shell pass */
int lbm_particles_fill(lbm_ctx* c) {
  PartArgs a = make_pargs(c);
  const dim3 g((unsigned)c->np, lbm_particles_chunks(c->ntmpl, c->nball));
  LBM_STORAGE_SWITCH(c->storage, hipLaunchKernelGGL(k_fill_apply<S>, dim3(c->np), dim3(NT_BALL), 0, c->stream, a))
  if (c->variant != LBM_KERNEL_FUSED) hipLaunchKernelGGL(k_fill_release, g, dim3(TPB), 0, c->stream, a);
  return last_err();
}

int lbm_particles_wetting(lbm_ctx* c) {
  PartArgs a = make_pargs(c);
  const dim3 g((unsigned)c->np, lbm_particles_chunks(c->ntmpl, c->nball));
  LBM_STORAGE_SWITCH(c->storage, hipLaunchKernelGGL(k_wetting<S>, g, dim3(TPB), 0, c->stream, a))
  hipLaunchKernelGGL(k_neigh, dim3(c->np), dim3(TPB), 0, c->stream, a);
  return last_err();
}

int lbm_particles_shell_wetting(lbm_ctx* c) {
  PartArgs a = make_pargs(c);
  const size_t lds = lbm_particles_shell_lds(c);
  if (lds) {
    LBM_STORAGE_SWITCH(c->storage, hipLaunchKernelGGL(k_shellbox<S>, dim3(c->np), dim3(NT_SHELL), lds, c->stream, a))
  } else {
    LBM_STORAGE_SWITCH(c->storage, hipLaunchKernelGGL(k_shellwet<S>, dim3(c->np), dim3(TPB), 0, c->stream, a))
  }
  return last_err();
}

/* LDS bytes of the boxed shell pass, 0 when the box is too large (then the
 * pass gathers from global memory): every template and ring offset must lie
 * within rmax of the centre */
size_t lbm_particles_shell_lds(const lbm_ctx* c) {
  if (!c->shell_box_ok) return 0;
  const size_t w = 2 * (size_t)c->rmax + 1, n = w * w * w;
  const size_t bytes = ((2 * sizeof(float) + 1) * n + 15) & ~(size_t)15;
  return bytes <= 24 * 1024 ? bytes : 0;
}

int lbm_particles_bounce(lbm_ctx* c) {
  PartArgs a = make_pargs(c);
  LBM_STORAGE_SWITCH(c->storage,
    hipLaunchKernelGGL(k_links<S>, dim3(c->np), dim3(TPB), 0, c->stream, a))
  return last_err();
}

int lbm_particles_max_neighbours() { return MAXNB; }

namespace {
PairParams make_pair_params(const lbm_ctx* c) {
  const lbm_config& g = c->cfg;
  PairParams pp;
  pp.hz_k = g.hz_k; pp.hz_rmin = g.hz_rmin; pp.hz_rcap = g.hz_rcap;
  pp.gmin = 2.5 * g.hz_k * pow((double)g.hz_rmin - (double)g.hz_rcap, 1.5);
  pp.rpar = 2.0 * (double)g.radius;
  const double R = g.radius;
  pp.lubfactor = (double)g.lub_k * 6.0 * 3.14159265358979323846 * (R * R) * (R * R) / (pp.rpar * pp.rpar);
  pp.rparcut = pp.rpar + (double)g.lub_rmin;
  pp.rparcap = pp.rpar + (double)g.lub_rcap;
  pp.mxrsqcut = fmax(pp.rparcut * pp.rparcut, (double)g.hz_rmin * (double)g.hz_rmin);
  pp.visc = ((double)g.tau_r - 0.5) / 3.0;
  pp.hertz = g.hertz;
  pp.lub = g.lubrication;
  pp.closure_r = pp.rpar + (double)c->closure_gap_max;
  return pp;
}

IntParams make_int_params(const lbm_ctx* c) {
  const lbm_config& g = c->cfg;
  const double R = g.radius;
  IntParams ip;
  ip.rmass = g.rmass;
  ip.rotin = g.rmass > 0 ? 2.0 / 5.0 * (1.0 / (double)g.rmass) * R * R : 0.0;
  ip.max_vel = g.max_vel;
  for (int d = 0; d < 3; ++d) { ip.ext_f[d] = g.ext_force[d]; ip.ext_t[d] = g.lrotate ? g.ext_torque[d] : 0.0; }
  return ip;
}
}  // namespace

int lbm_particles_integrate(lbm_ctx* c) {
  PartArgs a = make_pargs(c);
  const PairParams pp = make_pair_params(c);
  const IntParams ip = make_int_params(c);
  hipLaunchKernelGGL(k_forces, dim3(c->np), dim3(TPB), 0, c->stream, a, pp, ip);
  return last_err();
}

/* ---------------------------------------------------------------------- */
/* Particle phases of many boxes, one launch per kernel for all of them.   */
struct lbm_part_batch {
  int nbox;
  int storage;
  PartArgs* d_args[2];     /* per box, for the two time-level parities */
  PairParams* d_pp;
  IntParams* d_ip;
  int2* d_items;           /* (box, particle) */
  int nitems;
  unsigned chunks;         /* workgroups per particle of the chunked kernels */
};

lbm_part_batch* lbm_part_batch_create(lbm_ctx* const* boxes, int nbox, int* rc) {
  *rc = 0;
  lbm_part_batch* b = new lbm_part_batch();
  b->nbox = nbox;
  b->chunks = 1;
  b->storage = boxes[0]->storage;
  std::vector<PartArgs> args[2];
  std::vector<PairParams> pp(nbox);
  std::vector<IntParams> ip(nbox);
  std::vector<int2> items;
  for (int i = 0; i < nbox; ++i) {
    lbm_ctx* c = boxes[i];
    const int cur = c->cur;
    for (int par = 0; par < 2; ++par) {
      c->cur = par;
      args[par].push_back(make_pargs(c));
      args[par].back().rho_from_pops = 1;   /* the batch always runs the single-pass fluid kernel */
    }
    c->cur = cur;
    pp[i] = make_pair_params(c);
    ip[i] = make_int_params(c);
    for (int q = 0; q < c->np; ++q) items.push_back(make_int2(i, q));
    if (c->np > 0) {
      const unsigned ch = lbm_particles_chunks(c->ntmpl, c->nball);
      if (ch > b->chunks) b->chunks = ch;
    }
  }
  b->nitems = (int)items.size();
  b->d_args[0] = b->d_args[1] = nullptr;
  b->d_pp = nullptr; b->d_ip = nullptr; b->d_items = nullptr;
  auto up = [&](void** dptr, const void* h, size_t bytes) {
    if (*rc) return;
    if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) { *rc = LBM_ERR_ALLOC; return; }
    if (bytes && hipMemcpy(*dptr, h, bytes, hipMemcpyHostToDevice) != hipSuccess) *rc = LBM_ERR_HIP;
  };
  for (int par = 0; par < 2; ++par) up((void**)&b->d_args[par], args[par].data(), sizeof(PartArgs) * nbox);
  up((void**)&b->d_pp, pp.data(), sizeof(PairParams) * nbox);
  up((void**)&b->d_ip, ip.data(), sizeof(IntParams) * nbox);
  up((void**)&b->d_items, items.data(), sizeof(int2) * items.size());
  return b;
}

void lbm_part_batch_destroy(lbm_part_batch* b) {
  if (!b) return;
  for (int par = 0; par < 2; ++par) (void)hipFree(b->d_args[par]);
  (void)hipFree(b->d_pp);
  (void)hipFree(b->d_ip);
  (void)hipFree(b->d_items);
  delete b;
}

int lbm_part_batch_count(const lbm_part_batch* b) { return b->nitems; }

/* flags, erase, fill, shell densities + wetting (the particle work before the
 * collide-stream kernel, in lbm_step's order) */
int lbm_part_batch_pre(lbm_part_batch* b, int parity, long step, hipStream_t st) {
  if (b->nitems == 0) return 0;
  const PartArgs* A = b->d_args[parity];
  BatchStep bs;
  bs.step = (int)step;
  bs.lag_from_pops = step > 1 ? 1 : 0;
  const dim3 g((unsigned)b->nitems), gc((unsigned)b->nitems, b->chunks), t(TPB), tb(NT_BALL);
  hipLaunchKernelGGL(k_unstamp_batch, gc, t, 0, st, A, b->d_items, bs);
  hipLaunchKernelGGL(k_stamp_batch, gc, t, 0, st, A, b->d_items, bs);
  hipLaunchKernelGGL(k_resolve_batch, gc, t, 0, st, A, b->d_items, bs);
  LBM_STORAGE_SWITCH(b->storage,
    hipLaunchKernelGGL(k_erase_batch<S>, g, tb, 0, st, A, b->d_items, bs);
    hipLaunchKernelGGL(k_fill_apply_batch<S>, g, tb, 0, st, A, b->d_items, bs);
    hipLaunchKernelGGL(k_shellwet_batch<S>, g, t, 0, st, A, b->d_items, bs))
  return last_err();
}

/* neighbour lists, bounce-back links, pair forces, rigid-body update */
int lbm_part_batch_post(lbm_part_batch* b, int parity, long step, hipStream_t st) {
  if (b->nitems == 0) return 0;
  const PartArgs* A = b->d_args[parity];
  BatchStep bs;
  bs.step = (int)step;
  bs.lag_from_pops = step > 1 ? 1 : 0;
  const dim3 g((unsigned)b->nitems), t(TPB);
  LBM_STORAGE_SWITCH(b->storage, hipLaunchKernelGGL(k_links_batch<S>, g, t, 0, st, A, b->d_items, bs))
  hipLaunchKernelGGL(k_forces_batch, g, t, 0, st, A, b->d_pp, b->d_ip, b->d_items, bs);
  return last_err();
}
