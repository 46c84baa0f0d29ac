/*
 * lbm_fused.hip -- single-pass colour-gradient step for gfx950.
 *
 * One kernel does what the reference spreads over init_rhoR_CG/init_rhoB_CG,
 * the periodic refreshes, time_step_CG and the push streaming
 * (kernels_fluid_CG.CUF:398-509, 531-848, 866-1141; main.CUF:142-220):
 * densities, phase field, 26-point gradient, BGK + perturbation +
 * recolouring, and push streaming into the other time level.
 *
 * Work decomposition.  A workgroup owns a TX x TY column of the box over a
 * chunk of z planes and marches up the chunk.  For plane z+1 it loads the
 * populations of its own TX x TY nodes into registers and sums them into
 * rhoR, rhoB and phi; it also sums the populations of the one-node ring of
 * halo nodes around the column (their phi is needed by the gradient, their
 * populations are not).  phi of planes z-1, z, z+1 lives in a three-plane
 * ring in LDS, so the 26-point gradient of plane z reads LDS only.  Then
 * plane z collides from registers and pushes.  When TX equals nx the x
 * direction wraps inside the workgroup and there is no x halo at all.
 *
 * HBM traffic per interior node: 2 colours x 19 populations read once and
 * written once (the density pre-pass of a two-kernel step is gone); the halo
 * ring and the two extra planes of each z chunk are re-reads that neighbour
 * workgroups fetch too, so they are served by L2 / MALL.  Consecutive
 * workgroup ids are dealt round-robin over the 8 XCDs, so tiles are renumbered
 * to give each XCD a contiguous range of neighbouring columns (shared halos
 * stay in one L2).
 *
 * With particles the bounce-back copies of surface links are not done here:
 * fluid nodes push as usual and non-fluid nodes carry their stored values,
 * and the particle link kernel then writes every bounced slot (copy plus
 * moving-wall term) once, so this kernel needs no neighbour-type tests.
 *
 * Non-fluid nodes (particle surface and interior) take rhoR/rhoB from the
 * density arrays, which the particle kernels fill before this kernel runs
 * (zero inside particles, wetting values on the surface), as init_rho*_CG and
 * compute_*_rho leave them in the reference.
 */
#include "lbm_common.h"

#include <math.h>

#include <algorithm>
#include <type_traits>
#include <vector>

using namespace lbm;

namespace {

/* Population access.  A time level is addressed through five buffer
 * resources, one per group of four direction planes (base = level + 4 g *
 * plane bytes), a uniform offset (l mod 4) * plane bytes and a 32-bit byte
 * offset per lane, so a load or store needs no vector address arithmetic.
 * Offsets within a group stay below 4 GiB because four planes do (checked at
 * create).  A pair is 8 bytes (float) or 4 bytes (16-bit, kept packed in one
 * register until it is converted). */
using Rsrc = __amdgpu_buffer_rsrc_t;
using V2U = __attribute__((__vector_size__(2 * sizeof(unsigned int)))) unsigned int;

template <class S>
struct LevelIO {
  static constexpr uint32_t PS = 2 * sizeof(S);   /* bytes of one pair */
  Rsrc r[5];
  uint32_t pb;                                    /* bytes of one direction plane */
  float rdec, renc;                               /* rest-pair scales (16-bit)    */
  __device__ __forceinline__ LevelIO(const void* level, uint32_t plane_bytes, float rest_dec_scale, float rest_enc_scale)
      : pb(plane_bytes), rdec(rest_dec_scale), renc(rest_enc_scale) {
    char* base = const_cast<char*>(static_cast<const char*>(level));
#pragma unroll
    for (int g = 0; g < 5; ++g)
      r[g] = __builtin_amdgcn_make_buffer_rsrc(base + (size_t)(4 * g) * plane_bytes, (short)0, -1, 0x00020000);
  }
  __device__ __forceinline__ typename Pair<S>::Raw ld(int l, uint32_t node) const {
    const uint32_t so = (uint32_t)(l & 3) * pb;
    if constexpr (sizeof(S) == 4)
      return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r[l >> 2], node * PS, so, 0));
    else
      return __builtin_bit_cast(typename Pair<S>::Raw,
                                __builtin_amdgcn_raw_buffer_load_b32(r[l >> 2], node * PS, so, 0));
  }
  /* a loaded pair of direction l as floats (no shift) */
  __device__ __forceinline__ void cvt(int l, typename Pair<S>::Raw q, float& vr, float& vb) const;
  /* store a pair given as floats (already shifted for 16-bit storage); the
   * 16-bit conversions round to nearest even, two values per instruction; the
   * rest pair of 16-bit storage to its 3-exponent, 12-mantissa-bit format */
  __device__ __forceinline__ void st(int l, uint32_t node, float vr, float vb) const {
    const uint32_t so = (uint32_t)(l & 3) * pb;
    if constexpr (sizeof(S) == 4) {
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2U, make_float2(vr, vb)), r[l >> 2], node * PS, so, 0);
    } else if (l == 0) {
      __builtin_amdgcn_raw_buffer_store_b32(rest_enc(vr, vb, renc), r[0], node * PS, 0, 0);
    } else if constexpr (std::is_same<S, __half>::value) {
      typedef _Float16 H2 __attribute__((ext_vector_type(2)));
      const H2 h = {(_Float16)vr, (_Float16)vb};
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, h), r[l >> 2], node * PS, so, 0);
    } else {
      typedef __bf16 B2 __attribute__((ext_vector_type(2)));
      const B2 h = {(__bf16)vr, (__bf16)vb};
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, h), r[l >> 2], node * PS, so, 0);
    }
  }
};

/* convert a loaded pair to float (no shift) */
template <class S>
__device__ __forceinline__ void cvt_raw(typename Pair<S>::Raw q, float& r, float& b) {
  if constexpr (sizeof(S) == 4) {
    r = q.x;
    b = q.y;
  } else if constexpr (std::is_same<S, __half>::value) {
    r = __low2float(q);
    b = __high2float(q);
  } else {
    r = __uint_as_float(q << 16);
    b = __uint_as_float(q & 0xffff0000u);
  }
}

template <class S>
__device__ __forceinline__ void LevelIO<S>::cvt(int l, typename Pair<S>::Raw q, float& vr, float& vb) const {
  if constexpr (sizeof(S) == 2) {
    if (l == 0) {
      rest_dec(__builtin_bit_cast(uint32_t, q), rdec, vr, vb);
      return;
    }
  }
  cvt_raw<S>(q, vr, vb);
}

/* reciprocal and reciprocal square root to 1 ulp (v_rcp_f32 / v_rsq_f32)
 * instead of the correctly rounded multi-instruction sequences */
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

/* e_l . (x, y, z) with no multiplications by zero components (0 * x is not
 * folded away in IEEE arithmetic) */
__device__ __forceinline__ float edot(int l, float x, float y, float z) {
  float r = 0.0f;
  bool any = false;
  if (EX(l)) { r = EX(l) > 0 ? x : -x; any = true; }
  if (EY(l)) { r = any ? (EY(l) > 0 ? r + y : r - y) : (EY(l) > 0 ? y : -y); any = true; }
  if (EZ(l)) { r = any ? (EZ(l) > 0 ? r + z : r - z) : (EZ(l) > 0 ? z : -z); }
  return r;
}

/* 16-bit storage holds h = f - w_l rho0 (the shift makes the stored
 * deviations small); the fused kernel works on h directly and folds the
 * shift into its per-node constants, so no per-value add or subtract is spent
 * on it.  Float storage holds f (no shift). */
template <class S>
constexpr bool kShifted = sizeof(S) == 2;

/* rho = f_0 + f_1 + ... + f_18 in this order (init_rhoR_CG, :427-436); for
 * shifted storage the sum of the h, plus sum_l w_l rho0 = rho0 */
template <class S>
__device__ __forceinline__ void load_sum(const LevelIO<S>& src, uint32_t n, float rho0, float (&fr)[19],
                                         float (&fb)[19], float& rr, float& rb) {
#pragma unroll
  for (int l = 0; l < 19; ++l) src.cvt(l, src.ld(l, n), fr[l], fb[l]);
  rr = fr[0];
  rb = fb[0];
#pragma unroll
  for (int l = 1; l < 19; ++l) {
    rr = rr + fr[l];
    rb = rb + fb[l];
  }
  if (kShifted<S>) {
    rr = rr + rho0;
    rb = rb + rho0;
  }
}

template <class S>
__device__ __forceinline__ void sum_only(const LevelIO<S>& src, uint32_t n, float rho0, float& rr, float& rb) {
  float fr[19], fb[19];
  load_sum<S>(src, n, rho0, fr, fb, rr, rb);
}

/* The populations of the next plane, kept in registers across the barrier:
 * floats for float storage; for 16-bit storage the packed pairs as loaded
 * (19 registers instead of 38), converted to float h values when used. */
template <class S, bool PACKED = (sizeof(S) == 2)>
struct Ahead;
template <class S>
struct Ahead<S, false> {
  float r[19], b[19];
  __device__ __forceinline__ void load(const LevelIO<S>& src, uint32_t n, float rho0, float& rr, float& rb) {
    load_sum<S>(src, n, rho0, r, b, rr, rb);
  }
  __device__ __forceinline__ void get(const LevelIO<S>&, float (&fr)[19], float (&fb)[19]) const {
#pragma unroll
    for (int l = 0; l < 19; ++l) {
      fr[l] = r[l];
      fb[l] = b[l];
    }
  }
};
template <class S>
struct Ahead<S, true> {
  typename Pair<S>::Raw q[19];
  __device__ __forceinline__ void load(const LevelIO<S>& src, uint32_t n, float rho0, float& rr, float& rb) {
    float fr[19], fb[19];
#pragma unroll
    for (int l = 0; l < 19; ++l) {
      q[l] = src.ld(l, n);
      src.cvt(l, q[l], fr[l], fb[l]);
    }
    rr = fr[0];
    rb = fb[0];
#pragma unroll
    for (int l = 1; l < 19; ++l) {
      rr = rr + fr[l];
      rb = rb + fb[l];
    }
    rr = rr + rho0;
    rb = rb + rho0;
  }
  __device__ __forceinline__ void get(const LevelIO<S>& src, float (&fr)[19], float (&fb)[19]) const {
#pragma unroll
    for (int l = 0; l < 19; ++l) src.cvt(l, q[l], fr[l], fb[l]);
  }
};

/* weight class of direction l: 0 rest, 1 axis, 2 diagonal */
__device__ constexpr int WCLASS(int l) { return l == 0 ? 0 : (l < 7 ? 1 : 2); }
__device__ constexpr float WK(int k) { return k == 0 ? W0 : (k == 1 ? W1 : W2); }
__device__ constexpr float BK(int k) { return k == 0 ? B0 : (k == 1 ? B1 : B2); }

/* Collision of one fluid node, time_step_CG (kernels_fluid_CG.CUF:689-846),
 * in place on fr/fb: BGK towards f^eq(rho_k, u) per colour (:794-800), and
 * where |grad phi|^2 > 1e-20 the Leclaire perturbation (:803-815) and the
 * d'Ortona recolouring (:817-837).  The formulas are rearranged so that
 * everything that depends only on the node is formed once and each
 * direction costs about fifteen operations:
 *   f^eq_l(rho_k) = rho_k w_l (c1 + uv (1 + uv / 2)), uv = 3 e_l.u, c1 = 1 - 3/2 |u|^2
 *   perturbation  = (A w_l / |g|) (e_l.g)^2 - A |g| B_l, A = 9/4 omega sigma
 *   recolouring   = fd rho_k / rho +- T_l (e_l.g), T_l = beta rho_R rho_B / rho w_l / (|e_l| |g|)
 * (tt * feq of the reference, beta rho_R rho_B / rho^2 cos(phi_l) w_l rho, is
 * the same product).  Shifted storage: fr/fb hold h on entry and the shifted
 * outputs on exit; the shift enters the constants (BGK: -w_l rho0 omega;
 * perturbation: +2 w_l rho0; recolouring: -w_l rho0).  The momentum is formed
 * from the differences of opposite directions, in which the shift cancels.
 * 1/rho, 1/nu, 1/tau and 1/|g| use the 1-ulp hardware reciprocals; the colour
 * fractions rho_k / rho are correctly rounded quotients, since they carry each
 * colour's mass through the recolouring. */
template <bool SH>
__device__ __forceinline__ void collide_node(float (&fr)[19], float (&fb)[19], float rR, float rB, float phf,
                                             float px, float py, float pz, const FluidArgs& a) {
  const float rs = rR + rB;
  const float visc = frcp(0.5f * (1.0f + phf) * a.inv_visc_r + 0.5f * (1.0f - phf) * a.inv_visc_b);
  const float omega = frcp(visc * 3.0f + 0.5f);
  const float omm = 1.0f - omega;
  const float inv = frcp(rs);
  float d[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) d[i] = (fr[2 * i + 1] - fr[2 * i + 2]) + (fb[2 * i + 1] - fb[2 * i + 2]);
  const float u = inv * (d[0] + d[3] - d[4] + d[5] - d[6]);
  const float v = inv * (d[1] + d[3] + d[4] + d[7] - d[8]);
  const float wv = inv * (d[2] + d[5] + d[6] + d[7] + d[8]);
  const float u3 = 3.0f * u, v3 = 3.0f * v, w3 = 3.0f * wv;
  const float c1 = 1.0f - (0.5f / CSSQ) * (u * u + v * v + wv * wv);
  const float rRo = rR * omega, rBo = rB * omega;
  const float swn = SH ? a.rho0 * omega : 0.0f;
  /* BGK towards f^eq(rho_k, u) per colour (:794-800) */
#pragma unroll
  for (int l = 0; l < 19; ++l) {
    const int k = WCLASS(l);
    float brk = c1;
    if (l > 0) {
      const float uv = edot(l, u3, v3, w3);
      brk = fmaf(uv, fmaf(0.5f, uv, 1.0f), c1);
    }
    const float bw = WK(k) * brk;
    fr[l] = fmaf(fr[l], omm, SH ? fmaf(rRo, bw, -WK(k) * swn) : rRo * bw);
    fb[l] = fmaf(fb[l], omm, SH ? fmaf(rBo, bw, -WK(k) * swn) : rBo * bw);
  }
  const float pn2 = px * px + py * py + pz * pz;
  if (pn2 > 1.e-20f) {
    /* perturbation (:803-815) and recolouring (:817-837) */
    const float ipn = frsq(pn2);
    const float pn = pn2 * ipn;
    const float ac = 9.0f / 4.0f * omega * a.sigma;
    const float acp = ac * ipn, acq = ac * pn;
    const float tb = a.beta * rR * rB * inv * ipn;
    const float rRi = rR / rs, rBi = rB / rs;
    const float shn = SH ? a.rho0 : 0.0f;
#pragma unroll
    for (int l = 0; l < 19; ++l) {
      const int k = WCLASS(l);
      const float sk = WK(k) * shn;
      const float q = SH ? fmaf(acq, BK(k), -2.0f * sk) : acq * BK(k);
      if (l == 0) {
        const float fd = (fr[0] + fb[0]) - q;
        fr[0] = SH ? fmaf(fd, rRi, -sk) : fd * rRi;
        fb[0] = SH ? fmaf(fd, rBi, -sk) : fd * rBi;
      } else {
        const float ed = edot(l, px, py, pz);
        const float fd = (fr[l] + fb[l]) + fmaf(acp * WK(k), ed * ed, -q);
        const float te = (tb * (k == 1 ? W1 : W2 * INV_SQRT2)) * ed;
        fr[l] = fmaf(fd, rRi, SH ? te - sk : te);
        fb[l] = fmaf(fd, rBi, SH ? -te - sk : -te);
      }
    }
  }
}

template <int TX, int TY, bool WX>
struct TileGeom {
  static constexpr int NT = TX * TY;
  static constexpr int LW = WX ? TX : TX + 2;   /* LDS row length */
  static constexpr int MAXHALO = WX ? 2 * TX : 2 * (TX + 2) + 2 * TY;
  static_assert(MAXHALO <= NT, "one halo node per thread at most");
  __device__ static constexpr int col(int c) { return WX ? (c < 0 ? c + TX : (c >= TX ? c - TX : c)) : c + 1; }
};

/* one TX x TY column of the box over the z chunk izc (planes [izc zc, izc zc + zc)) */
template <class S, bool PART, int TX, int TY, bool WX>
__device__ __forceinline__ void fused_tile(const FluidArgs& a, int itx, int ity, int izc, int zc,
                                           float (&sphi)[3][TY + 2][TileGeom<TX, TY, WX>::LW]) {
  using G = TileGeom<TX, TY, WX>;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t pb = (uint32_t)(a.N * (2 * sizeof(S)));   /* bytes of one direction plane */
  const uint32_t sxy = (uint32_t)nx * (uint32_t)ny;
  const float rho0 = a.rho0;
  const LevelIO<S> src(a.src, pb, a.rdec, a.renc);
  const LevelIO<S> dst(a.dst, pb, a.rdec, a.renc);
  const int x0 = itx * TX, y0 = ity * TY, z0 = izc * zc;
  const int z1 = min(z0 + zc, nz);
  const int w = WX ? TX : min(TX, nx - x0);
  const int h = min(TY, ny - y0);
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = tx + TX * ty;
  const bool inside = tx < w && ty < h;
  const int x = x0 + tx, y = y0 + ty;
  const int cm = G::col(tx - 1), c0 = G::col(tx), cp = G::col(tx + 1);
  const uint32_t nxy = inside ? (uint32_t)x + (uint32_t)y * (uint32_t)nx : 0u;

  /* the halo node of this thread (fixed in x, y) */
  const int nhalo = WX ? 2 * TX : 2 * (w + 2) + 2 * h;
  const bool has_halo = tid < nhalo;
  int hrow = 0, hcol = 0;
  uint32_t hxy = 0;
  if (has_halo) {
    int c, r;
    if (WX) {
      r = tid < TX ? -1 : h;
      c = tid < TX ? tid : tid - TX;
    } else {
      const int W2 = w + 2;
      if (tid < W2) { r = -1; c = tid - 1; }
      else if (tid < 2 * W2) { r = h; c = tid - W2 - 1; }
      else { const int k2 = tid - 2 * W2; r = k2 >> 1; c = (k2 & 1) ? w : -1; }
    }
    hxy = (uint32_t)wrapc(x0 + c, nx) + (uint32_t)wrapc(y0 + r, ny) * (uint32_t)nx;
    hrow = r + 1;
    hcol = G::col(c);
  }
  /* x and y neighbours of the interior node */
  const uint32_t xs[3] = {(uint32_t)(x == 0 ? nx - 1 : x - 1), (uint32_t)x, (uint32_t)(x == nx - 1 ? 0 : x + 1)};
  const uint32_t ys[3] = {(uint32_t)(y == 0 ? ny - 1 : y - 1) * (uint32_t)nx, (uint32_t)y * (uint32_t)nx,
                          (uint32_t)(y == ny - 1 ? 0 : y + 1) * (uint32_t)nx};

  /* state of the interior node on the next plane (kept in registers) */
  Ahead<S> ahead;
  float qR = 0.0f, qB = 0.0f;
  uint8_t qF = FL_FLUID;

  /* phi of plane zl (wrapped) into ring slot s; the halo node first, then the
   * interior node, whose populations are kept in ahead for the next
   * iteration (after the last plane of the chunk they are not used).  With
   * particles the node type and the stored phase field (zero densities inside
   * particles, wetting values on their surface) are loaded beside the
   * populations, and the phase field of a halo node is formed whatever its
   * type and selected afterwards, so no population load waits on a type. */
  auto plane = [&](int zl, int s) {
    const int zg = zl < 0 ? zl + nz : (zl >= nz ? zl - nz : zl);
    const uint32_t zoff = (uint32_t)zg * sxy;
    if (has_halo) {
      const uint32_t n = hxy + zoff;
      const uint8_t hf = PART ? a.flags[n] : FL_FLUID;
      const float hp = PART ? a.phi[n] : 0.0f;
      float rr, rb;
      sum_only<S>(src, n, rho0, rr, rb);
      if (rr < 0.0f) rr = 0.0f;
      if (rb < 0.0f) rb = 0.0f;
      const float ph = (rr - rb) * frcp(rr + rb);
      sphi[s][hrow][hcol] = hf == FL_FLUID ? ph : hp;
    }
    if (inside) {
      const uint32_t n = nxy + zoff;
      qF = PART ? a.flags[n] : FL_FLUID;
      const float qp = PART ? a.phi[n] : 0.0f;
      float rr, rb;
      ahead.load(src, n, rho0, rr, rb);
      float ph = qp;
      if (qF == FL_FLUID) {
        if (rr < a.min_rho || rr > a.max_rho || rb < a.min_rho || rb > a.max_rho)
          raise_error(a.status, LBM_ERR_RHO_RANGE, a.step);
        if (rr < 0.0f) rr = 0.0f;
        if (rb < 0.0f) rb = 0.0f;
        ph = (rr - rb) * frcp(rr + rb);
      }
      qR = rr;
      qB = rb;
      sphi[s][ty + 1][c0] = ph;
    }
  };

  bool rest_bad = false;   /* a rest pair beyond its 16-bit range */
  plane(z0 - 1, 0);
  plane(z0, 1);
  int sm = 0, s0 = 1, sp = 2;
  for (int z = z0; z < z1; ++z) {
    float fr[19], fb[19];
    ahead.get(src, fr, fb);
    const float rR = qR, rB = qB;
    const uint8_t f = qF;
    plane(z + 1, sp);
    __syncthreads();
    if (inside) {
      const uint32_t zs[3] = {(uint32_t)(z == 0 ? nz - 1 : z - 1) * sxy, (uint32_t)z * sxy,
                              (uint32_t)(z == nz - 1 ? 0 : z + 1) * sxy};
      const uint32_t n = xs[1] + ys[1] + zs[1];
      auto nb3 = [&](int dx, int dy, int dz) -> uint32_t { return xs[dx + 1] + ys[dy + 1] + zs[dz + 1]; };
      if (f == FL_FLUID) {
        /* 26-point gradient in the D3Q27 order of kernels_fluid_CG.CUF:720-741;
           zero-component terms add nothing and are skipped */
        float px = 0.0f, py = 0.0f, pz = 0.0f;
#pragma unroll
        for (int l = 1; l < 27; ++l) {
          const int dx = EX27(l), dy = EY27(l), dz = EZ27(l);
          const int slot = dz < 0 ? sm : (dz > 0 ? sp : s0);
          const int cc = dx < 0 ? cm : (dx > 0 ? cp : c0);
          const float ps = sphi[slot][ty + 1 + dy][cc];
          if (dx) px = px + A27(l) * (float)dx * ps;
          if (dy) py = py + A27(l) * (float)dy * ps;
          if (dz) pz = pz + A27(l) * (float)dz * ps;
        }
        collide_node<kShifted<S>>(fr, fb, rR, rB, sphi[s0][ty + 1][c0], px, py, pz, a);
      }
      if (kShifted<S>) rest_bad |= rest_out_of_range(fr[0], fb[0], a.renc);
      /* push: every node but spheredead sends its 18 moving populations to
         x + e_l and keeps the rest population (flipflop*Pop0_CG, stream*_CG,
         kernels_fluid_CG.CUF:866-1015); non-fluid nodes send their stored
         values unchanged (no collision there, :689).  The slots a bounce-back
         link claims are overwritten afterwards by the particle link kernel.
         Stored values of non-fluid nodes are re-stored from the registers
         they were loaded into, exactly (16-bit values are held as floats
         without the shift, so the conversion back is exact). */
      dst.st(0, n, fr[0], fb[0]);
      if (!PART || f != FL_DEAD) {
#pragma unroll
        for (int l = 1; l < 19; ++l) dst.st(l, nb3(EX(l), EY(l), EZ(l)), fr[l], fb[l]);
      }
    }
    __syncthreads();
    const int so = sm;
    sm = s0;
    s0 = sp;
    sp = so;
  }
  if (rest_bad) raise_error(a.status, LBM_ERR_STORE_RANGE, a.step);
}

template <class S, bool PART, int TX, int TY, bool WX>
__global__ __launch_bounds__(TX * TY) void k_fused(FluidArgs a, int ntx, int nty, int nzc, int zc) {
  __shared__ float sphi[3][TY + 2][TileGeom<TX, TY, WX>::LW];
  /* tile id: XCD k (= blockIdx.x mod 8) gets the contiguous tiles [k q, (k+1) q) */
  const int nb = ntx * nty * nzc;
  int t = blockIdx.x;
  {
    const int q = nb >> 3;
    if (t < (q << 3)) t = (t & 7) * q + (t >> 3);
  }
  const int itx = t % ntx;
  const int ity = (t / ntx) % nty;
  const int izc = t / (ntx * nty);
  fused_tile<S, PART, TX, TY, WX>(a, itx, ity, izc, zc, sphi);
}

/* Batched step: many independent boxes in one persistent launch.  Work items
 * are (box, tile x, tile y, z chunk), largest boxes first; each workgroup
 * takes the next item from a device counter until the list is exhausted, so
 * uneven boxes balance across the CUs.  The counter is never reset: launch L
 * owns the range [base, base + nitems + gridDim.x), every workgroup ending
 * with exactly one failed fetch. */
template <class S, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_fused_batch(const FluidArgs* __restrict__ boxes,
                                                         const int4* __restrict__ items, int nitems, int zc,
                                                         unsigned long long* counter, unsigned long long base,
                                                         int step) {
  __shared__ float sphi[3][TY + 2][TileGeom<TX, TY, false>::LW];
  __shared__ int next;
  const int tid = threadIdx.x + TX * threadIdx.y;
  for (;;) {
    if (tid == 0) next = (int)(atomicAdd(counter, 1ull) - base);
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(next);
    if (i >= nitems) break;
    const int4 it = items[i];
    FluidArgs a = boxes[it.x];
    a.step = step;
    fused_tile<S, true, TX, TY, false>(a, it.y, it.z, it.w, zc, sphi);
    __syncthreads();
  }
}

/* ------------------------------------------------------------------ */
/* tile table: the autotuner picks one entry and a z-chunk length */
struct TileEntry {
  int tx, ty;
  bool wx;
};
constexpr TileEntry kTiles[] = {
    {64, 4, false}, {64, 8, false}, {64, 16, false}, {128, 4, false}, {128, 8, false},
    {256, 2, true}, {256, 4, true}, {128, 4, true},  {128, 8, true},  {64, 8, true},
};
constexpr int kNumTiles = (int)(sizeof(kTiles) / sizeof(kTiles[0]));

FluidArgs make_fargs(lbm_ctx* c) { return lbm_fluid_args(c); }

template <class S, bool PART, int TX, int TY, bool WX>
void launch_one(const FluidArgs& a, int zc, hipStream_t st) {
  const int ntx = (a.nx + TX - 1) / TX, nty = (a.ny + TY - 1) / TY, nzc = (a.nz + zc - 1) / zc;
  hipLaunchKernelGGL((k_fused<S, PART, TX, TY, WX>), dim3((unsigned)(ntx * nty * nzc)), dim3(TX, TY), 0, st, a,
                     ntx, nty, nzc, zc);
}

template <class S, bool PART>
void launch_tile(int tile, const FluidArgs& a, int zc, hipStream_t st) {
  switch (tile) {
    case 0: launch_one<S, PART, 64, 4, false>(a, zc, st); break;
    case 1: launch_one<S, PART, 64, 8, false>(a, zc, st); break;
    case 2: launch_one<S, PART, 64, 16, false>(a, zc, st); break;
    case 3: launch_one<S, PART, 128, 4, false>(a, zc, st); break;
    case 4: launch_one<S, PART, 128, 8, false>(a, zc, st); break;
    case 5: launch_one<S, PART, 256, 2, true>(a, zc, st); break;
    case 6: launch_one<S, PART, 256, 4, true>(a, zc, st); break;
    case 7: launch_one<S, PART, 128, 4, true>(a, zc, st); break;
    case 8: launch_one<S, PART, 128, 8, true>(a, zc, st); break;
    default: launch_one<S, PART, 64, 8, true>(a, zc, st); break;
  }
}

}  // namespace

int lbm_fused_num_tiles() { return kNumTiles; }

bool lbm_fused_tile_valid(const lbm_ctx* c, int tile, int zc) {
  if (tile < 0 || tile >= kNumTiles || zc < 1) return false;
  const TileEntry& e = kTiles[tile];
  if (e.wx && c->nx != e.tx) return false;
  if (!e.wx && c->nx < 3) return false;
  return true;
}

void lbm_fused_tile_shape(int tile, int* tx, int* ty, int* wx) {
  const TileEntry& e = kTiles[tile < 0 || tile >= kNumTiles ? 0 : tile];
  *tx = e.tx; *ty = e.ty; *wx = e.wx ? 1 : 0;
}

/* launch the fused step with the given tile / z chunk; dst overrides the
 * destination buffer (the autotuner times into a scratch buffer) */
int lbm_launch_fused_with(lbm_ctx* c, int tile, int zc, void* dst) {
  if (!lbm_fused_tile_valid(c, tile, zc)) return LBM_ERR_ARG;
  FluidArgs a = make_fargs(c);
  if (dst) a.dst = dst;
  const bool part = c->np > 0;
#define LBM_FUSED_GO(S)                                       \
  if (part) launch_tile<S, true>(tile, a, zc, c->stream);     \
  else launch_tile<S, false>(tile, a, zc, c->stream);
  switch (c->storage) {
    case LBM_STORE_F32: { LBM_FUSED_GO(float) } break;
    case LBM_STORE_F16: { LBM_FUSED_GO(__half) } break;
    default: { LBM_FUSED_GO(__hip_bfloat16) } break;
  }
#undef LBM_FUSED_GO
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

int lbm_launch_fused(lbm_ctx* c) {
  lbm_timing_begin(c);
  int rc = lbm_launch_fused_with(c, c->tile, c->zchunk, nullptr);
  lbm_timing_end(c);
  return rc;
}

/* ---------------------------------------------------------------------- */
/* Fluid step of many boxes in one persistent launch.                      */
struct lbm_fluid_batch {
  int nbox;
  int storage;
  FluidArgs* d_args[2];           /* per box, for the two time-level parities */
  std::vector<FluidArgs> h_args[2];
  int4* d_items;
  int nitems;
  int tile, zc, grid;
  unsigned long long* d_counter;
  unsigned long long base;
  std::vector<int> vol;           /* nodes per box, to order the work items */
  std::vector<int> dims;          /* nx, ny, nz per box */
};

namespace {

template <class S, int TX, int TY>
int batch_grid() {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused_batch<S, TX, TY>, TX * TY, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
  return per_cu * cus;
}

template <class S, int TX, int TY>
void batch_launch(const lbm_fluid_batch* b, const FluidArgs* args, int step, hipStream_t st) {
  hipLaunchKernelGGL((k_fused_batch<S, TX, TY>), dim3((unsigned)b->grid), dim3(TX, TY), 0, st, args, b->d_items,
                     b->nitems, b->zc, b->d_counter, b->base, step);
}

template <class S>
int batch_dispatch(lbm_fluid_batch* b, const FluidArgs* args, int step, hipStream_t st, bool grid_only) {
  switch (b->tile) {
    case 0: if (grid_only) return batch_grid<S, 64, 4>(); batch_launch<S, 64, 4>(b, args, step, st); break;
    case 1: if (grid_only) return batch_grid<S, 64, 8>(); batch_launch<S, 64, 8>(b, args, step, st); break;
    case 2: if (grid_only) return batch_grid<S, 64, 16>(); batch_launch<S, 64, 16>(b, args, step, st); break;
    case 3: if (grid_only) return batch_grid<S, 128, 4>(); batch_launch<S, 128, 4>(b, args, step, st); break;
    default: if (grid_only) return batch_grid<S, 128, 8>(); batch_launch<S, 128, 8>(b, args, step, st); break;
  }
  return 0;
}

int batch_call(lbm_fluid_batch* b, const FluidArgs* args, int step, hipStream_t st, bool grid_only) {
  switch (b->storage) {
    case LBM_STORE_F32: return batch_dispatch<float>(b, args, step, st, grid_only);
    case LBM_STORE_F16: return batch_dispatch<__half>(b, args, step, st, grid_only);
    default: return batch_dispatch<__hip_bfloat16>(b, args, step, st, grid_only);
  }
}

}  // namespace

/* tiles the batched kernel is built for: the entries of the tile table
 * without x wrap (boxes of one batch differ in nx) */
bool lbm_fluid_batch_tile_ok(int tile) { return tile >= 0 && tile <= 4; }

lbm_fluid_batch* lbm_fluid_batch_create(lbm_ctx* const* boxes, int nbox, int* rc) {
  *rc = 0;
  lbm_fluid_batch* b = new lbm_fluid_batch();
  b->nbox = nbox;
  b->storage = boxes[0]->storage;
  b->d_args[0] = b->d_args[1] = nullptr;
  b->d_items = nullptr;
  b->d_counter = nullptr;
  b->nitems = 0;
  b->base = 0;
  b->tile = -1;
  for (int i = 0; i < nbox; ++i) {
    lbm_ctx* c = boxes[i];
    const int cur = c->cur;
    for (int par = 0; par < 2; ++par) {
      c->cur = par;
      b->h_args[par].push_back(make_fargs(c));
    }
    c->cur = cur;
    b->vol.push_back((int)c->N);
    b->dims.push_back(c->nx); b->dims.push_back(c->ny); b->dims.push_back(c->nz);
  }
  for (int par = 0; par < 2 && !*rc; ++par) {
    if (hipMalloc(&b->d_args[par], sizeof(FluidArgs) * nbox) != hipSuccess) { *rc = LBM_ERR_ALLOC; break; }
    if (hipMemcpy(b->d_args[par], b->h_args[par].data(), sizeof(FluidArgs) * nbox, hipMemcpyHostToDevice) != hipSuccess)
      *rc = LBM_ERR_HIP;
  }
  if (!*rc && hipMalloc(&b->d_counter, sizeof(unsigned long long)) != hipSuccess) *rc = LBM_ERR_ALLOC;
  if (!*rc && hipMemset(b->d_counter, 0, sizeof(unsigned long long)) != hipSuccess) *rc = LBM_ERR_HIP;
  return b;
}

void lbm_fluid_batch_destroy(lbm_fluid_batch* b) {
  if (!b) return;
  for (int par = 0; par < 2; ++par) (void)hipFree(b->d_args[par]);
  (void)hipFree(b->d_items);
  (void)hipFree(b->d_counter);
  delete b;
}

/* work items of tile `tile` and z chunk zc: (box, tile x, tile y, chunk),
 * boxes in decreasing size so the big ones start first */
int lbm_fluid_batch_config(lbm_fluid_batch* b, int tile, int zc) {
  if (!lbm_fluid_batch_tile_ok(tile) || zc < 1) return LBM_ERR_ARG;
  const TileEntry& e = kTiles[tile];
  std::vector<int> order(b->nbox);
  for (int i = 0; i < b->nbox; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int u, int v) { return b->vol[u] > b->vol[v]; });
  std::vector<int4> items;
  for (int i : order) {
    const int nx = b->dims[3 * i], ny = b->dims[3 * i + 1], nz = b->dims[3 * i + 2];
    const int ntx = (nx + e.tx - 1) / e.tx, nty = (ny + e.ty - 1) / e.ty, nzc = (nz + zc - 1) / zc;
    for (int k = 0; k < nzc; ++k)
      for (int j = 0; j < nty; ++j)
        for (int t = 0; t < ntx; ++t) items.push_back(make_int4(i, t, j, k));
  }
  (void)hipFree(b->d_items);
  b->d_items = nullptr;
  if (hipMalloc(&b->d_items, sizeof(int4) * items.size()) != hipSuccess) return LBM_ERR_ALLOC;
  if (hipMemcpy(b->d_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice) != hipSuccess)
    return LBM_ERR_HIP;
  b->nitems = (int)items.size();
  b->tile = tile;
  b->zc = zc;
  b->grid = batch_call(b, nullptr, 0, nullptr, true);
  return LBM_OK;
}

/* one launch for the fluid of every box; dst_override (per box, or NULL)
 * redirects the output level (the autotuner times into scratch levels) */
int lbm_fluid_batch_launch(lbm_fluid_batch* b, int parity, long step, hipStream_t st, FluidArgs* d_override) {
  if (b->tile < 0) return LBM_ERR_STATE;
  batch_call(b, d_override ? d_override : b->d_args[parity], (int)step, st, false);
  b->base += (unsigned long long)b->nitems + (unsigned long long)b->grid;
  return hipGetLastError() == hipSuccess ? 0 : LBM_ERR_HIP;
}

/* per-box argument blocks with the output level replaced (autotuning) */
FluidArgs* lbm_fluid_batch_scratch_args(lbm_fluid_batch* b, int parity, void* const* scratch) {
  std::vector<FluidArgs> a = b->h_args[parity];
  for (int i = 0; i < b->nbox; ++i) a[i].dst = scratch[i];
  FluidArgs* d = nullptr;
  if (hipMalloc(&d, sizeof(FluidArgs) * b->nbox) != hipSuccess) return nullptr;
  if (hipMemcpy(d, a.data(), sizeof(FluidArgs) * b->nbox, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  return d;
}

int lbm_fluid_batch_items(const lbm_fluid_batch* b) { return b->nitems; }
