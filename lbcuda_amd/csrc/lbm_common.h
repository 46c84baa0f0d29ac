/*
 * lbm_common.h -- lattice constants, storage traits and the internal context
 * shared by the HIP translation units of liblbcuda_amd.
 *
 * Lattice tables follow dimensions_m.CUF:201-300 (D3Q19 order, Leclaire B_l
 * with xi = 0.5, D3Q27 weights for the phase-field gradient).  All constants
 * are formed in float the way the reference's single-precision parameters are.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <stddef.h>
#include <algorithm>
#include <cmath>

#include "../../include/lbcuda_amd.h"

namespace lbm {

/* node types, dimensions_m.CUF:305-309 */
constexpr uint8_t FL_WALL = 0, FL_FLUID = 1, FL_LIST = 2, FL_DEAD = 4, FL_CM = 5;

/* D3Q19 velocity set in the reference order (dimensions_m.CUF:212-219) */
__host__ __device__ constexpr int EX(int l) {
  return l == 1 || l == 7 || l == 10 || l == 11 || l == 14 ? 1
       : (l == 2 || l == 8 || l == 9 || l == 12 || l == 13 ? -1 : 0);
}
__host__ __device__ constexpr int EY(int l) {
  return l == 3 || l == 7 || l == 9 || l == 15 || l == 18 ? 1
       : (l == 4 || l == 8 || l == 10 || l == 16 || l == 17 ? -1 : 0);
}
__host__ __device__ constexpr int EZ(int l) {
  return l == 5 || l == 11 || l == 13 || l == 15 || l == 17 ? 1
       : (l == 6 || l == 12 || l == 14 || l == 16 || l == 18 ? -1 : 0);
}
__host__ __device__ constexpr int OPP(int l) { return l == 0 ? 0 : (l % 2 == 1 ? l + 1 : l - 1); }

/* weights p(l) and recolouring factors, dimensions_m.CUF:206-209, main.CUF:1102 */
constexpr float CSSQ = 1.0f / 3.0f;
/* The floats nearest 1/3, 1/18, 1/36 sum to 1 + 1.5e-8, so every BGK step
 * would add 1.5e-8 of each node's mass (7.5e-5 after 5000 steps, as the
 * reference's FP32 build does).  W0 two ulps below, W1 four ulps above and
 * W2 two ulps below the nearest floats sum to exactly 1 and keep the second
 * moment 3 (2 W1 + 8 W2) = 1 + 7.5e-9 of the nearest floats (it sets the
 * momentum each BGK step returns). */
constexpr float W0 = 0x1.555552p-2f, W1 = 0x1.c71c7ap-5f, W2 = 0x1.c71c6ep-6f;
static_assert((double)W0 + 6.0 * (double)W1 + 12.0 * (double)W2 == 1.0, "D3Q19 weights must sum to 1");
static_assert(3.0 * (2.0 * (double)W1 + 8.0 * (double)W2) - 1.0 < 1e-8, "second moment of the weights");
__host__ __device__ constexpr float W(int l) { return l == 0 ? W0 : (l < 7 ? W1 : W2); }
/* Leclaire B_l, dimensions_m.CUF:222-228 */
constexpr float BXI = 0.5f;
constexpr float B0 = -(2.0f + 2.0f * BXI) / (3.0f * BXI + 12.0f);
constexpr float B1 = BXI / (6.0f * BXI + 24.0f);
constexpr float B2 = 1.0f / (6.0f * BXI + 24.0f);
__host__ __device__ constexpr float BL(int l) { return l == 0 ? B0 : (l < 7 ? B1 : B2); }
/* 1/|e_l| (rec_fact_d = |e_l|, main.CUF:1102; l = 0 gives cos(phi) = 0) */
constexpr float INV_SQRT2 = 0.70710678118654752440f;
__host__ __device__ constexpr float INV_REC(int l) { return l == 0 ? 0.0f : (l < 7 ? 1.0f : INV_SQRT2); }

/* density-ratio model tables phi, varphi, psi, xi (dimensions_m.CUF:230-262):
 * equilCG = rho (phi_l + varphi_l alpha_k + p_l (u-terms)); psi_l and xi_l
 * weigh the v.grad(rho_k) and diagonal (v grad rho_k + transpose) terms */
__host__ __device__ constexpr float PHI_DD(int l) { return l == 0 ? 0.0f : (l < 7 ? 1.0f / 12.0f : 1.0f / 24.0f); }
__host__ __device__ constexpr float VARPHI_DD(int l) { return l == 0 ? 1.0f : (l < 7 ? -1.0f / 12.0f : -1.0f / 24.0f); }
__host__ __device__ constexpr float PSI_DD(int l) { return l == 0 ? -2.5f : (l < 7 ? -1.0f / 6.0f : 1.0f / 24.0f); }
__host__ __device__ constexpr float XI_DD(int l) { return l == 0 ? 0.0f : (l < 7 ? 0.25f : 0.125f); }

/* D3Q27 gradient weights ad3q27 = pd3q27/cssq, dimensions_m.CUF:283-300 */
constexpr float P1D = (2.0f / 3.0f) * (2.0f / 3.0f) * (1.0f / 6.0f);
constexpr float P2D = (1.0f / 6.0f) * (1.0f / 6.0f) * (2.0f / 3.0f);
constexpr float P3D = (1.0f / 6.0f) * (1.0f / 6.0f) * (1.0f / 6.0f);
constexpr float P0D = (2.0f / 3.0f) * (2.0f / 3.0f) * (2.0f / 3.0f);
constexpr float A1D = P1D / CSSQ, A2D = P2D / CSSQ, A3D = P3D / CSSQ;

/* D3Q27 links in the reference order (dimensions_m.CUF:275-282): 0..18 are the
 * D3Q19 links, 19..26 the corners. */
__host__ __device__ constexpr int EX27(int l) {
  return l < 19 ? EX(l) : (l == 19 || l == 21 || l == 24 || l == 25 ? 1 : -1);
}
__host__ __device__ constexpr int EY27(int l) {
  return l < 19 ? EY(l) : (l == 19 || l == 21 || l == 23 || l == 26 ? 1 : -1);
}
__host__ __device__ constexpr int EZ27(int l) {
  return l < 19 ? EZ(l) : (l == 19 || l == 22 || l == 23 || l == 25 ? 1 : -1);
}
__host__ __device__ constexpr float A27(int l) { return l < 7 ? A1D : (l < 19 ? A2D : A3D); }

/* pref_bouzidi = 2/cssq, dimensions_m.CUF:99 */
constexpr float BUZ = 2.0f / CSSQ;

/* ------------------------------------------------------------------ */
/* storage traits: populations are kept either as float or as the shifted
 * value f - w_l*rho0 in a 16-bit type; arithmetic is always float. */
template <class S> struct Store;
template <> struct Store<float> {
  __device__ static float ld(const float* p, size_t i, float) { return p[i]; }
  __device__ static void st(float* p, size_t i, float v, float) { p[i] = v; }
};
template <> struct Store<__half> {
  __device__ static float ld(const __half* p, size_t i, float s) { return __half2float(p[i]) + s; }
  __device__ static void st(__half* p, size_t i, float v, float s) { p[i] = __float2half_rn(v - s); }
};
template <> struct Store<__hip_bfloat16> {
  __device__ static float ld(const __hip_bfloat16* p, size_t i, float s) { return __bfloat162float(p[i]) + s; }
  __device__ static void st(__hip_bfloat16* p, size_t i, float v, float s) { p[i] = __float2bfloat16(v - s); }
};

/* Population layout.  Each time level holds, for every direction l, a plane of
 * N colour pairs (f_R, f_B) of node n, x fastest: the pair (l, n) is element
 * l N + n of an array of 2-value pairs (8 bytes for float, 4 bytes for the
 * 16-bit types).  One vector access moves both colours of a node and
 * direction, and a gather of whole nodes touches 19 cache lines per row, not
 * 38.  Pair<S>::ld/st convert one pair with the shift s = w_l rho0 of the
 * direction, element by element exactly as Store<S> does. */
template <class S> struct Pair;
template <> struct Pair<float> {
  using Raw = float2;   /* a pair as loaded, before conversion */
  __device__ static Raw raw(const float* p, size_t i) { return reinterpret_cast<const float2*>(p)[i]; }
  __device__ static void cvt(Raw v, float, float& r, float& b) {
    r = v.x;
    b = v.y;
  }
  __device__ static void ld(const float* p, size_t i, float, float& r, float& b) {
    const float2 v = reinterpret_cast<const float2*>(p)[i];
    r = v.x;
    b = v.y;
  }
  __device__ static void st(float* p, size_t i, float r, float b, float) {
    reinterpret_cast<float2*>(p)[i] = make_float2(r, b);
  }
};
template <> struct Pair<__half> {
  using Raw = __half2;
  __device__ static Raw raw(const __half* p, size_t i) { return reinterpret_cast<const __half2*>(p)[i]; }
  __device__ static void cvt(Raw v, float s, float& r, float& b) {
    r = __low2float(v) + s;
    b = __high2float(v) + s;
  }
  __device__ static void ld(const __half* p, size_t i, float s, float& r, float& b) {
    const __half2 v = reinterpret_cast<const __half2*>(p)[i];
    r = __low2float(v) + s;
    b = __high2float(v) + s;
  }
  __device__ static void st(__half* p, size_t i, float r, float b, float s) {
    reinterpret_cast<__half2*>(p)[i] = __halves2half2(__float2half_rn(r - s), __float2half_rn(b - s));
  }
};
template <> struct Pair<__hip_bfloat16> {
  using Raw = uint32_t;
  __device__ static Raw raw(const __hip_bfloat16* p, size_t i) { return reinterpret_cast<const uint32_t*>(p)[i]; }
  __device__ static void cvt(Raw u, float s, float& r, float& b) {
    r = __uint_as_float(u << 16) + s;
    b = __uint_as_float(u & 0xffff0000u) + s;
  }
  __device__ static void ld(const __hip_bfloat16* p, size_t i, float s, float& r, float& b) {
    const uint32_t u = reinterpret_cast<const uint32_t*>(p)[i];
    r = __uint_as_float(u << 16) + s;
    b = __uint_as_float(u & 0xffff0000u) + s;
  }
  __device__ static void st(__hip_bfloat16* p, size_t i, float r, float b, float s) {
    const uint32_t lo = __bfloat16_as_ushort(__float2bfloat16(r - s));
    const uint32_t hi = __bfloat16_as_ushort(__float2bfloat16(b - s));
    reinterpret_cast<uint32_t*>(p)[i] = lo | (hi << 16);
  }
};

/* The rest pair (l = 0) of 16-bit storage.  The shifted rest value
 * h_0 = f_0 - w_0 rho0 is the largest stored deviation: about w_0 |rho_k -
 * rho0| (0.017 in a 50/50 mixture, 0.17 where a colour is absent or at full
 * density), where FP16 spacing is 2^-16 and 2^-13.  The rest pair is kept
 * as two 16-bit floats with 3 exponent and 12 mantissa bits instead
 * (sign | e | m): |h| = 2^(e - 9 + s) (1 + m / 4096) for e = 1..7 and
 * m 2^(-20 + s) for e = 0, 4x finer than FP16 from 2^(-8 + s) up to the
 * largest value, 2^(s - 1) (1 - 2^-13).  s comes from the configuration
 * (rest_offset below); rounding is to nearest even, values beyond the range
 * saturate (the collide passes report LBM_ERR_STORE_RANGE).
 *
 * The code of a value is the FP32 bit pattern of h 2^(-118 - s), whose
 * exponent field is e (a denormal for e = 0), shifted right by 11 bits with
 * rounding; the decoder shifts back and scales by 2^(118 + s), exact for
 * every code (FP32 denormals are kept: denorm mode 3 in the kernel
 * descriptors).  The two scales are rdec = 2^(118 + s), renc = 2^(-118 - s). */
constexpr uint32_t REST_MAXCODE = 0x7fffu;
__device__ __forceinline__ float rest_dec1(uint32_t code16, float rdec) {
  const uint32_t bits = ((code16 & 0x7fffu) << 11) | ((code16 & 0x8000u) << 16);
  return __uint_as_float(bits) * rdec;
}
__device__ __forceinline__ uint32_t rest_enc1(float h, float renc) {
  const uint32_t u = __float_as_uint(h * renc);
  const uint32_t mag = u & 0x7fffffffu;
  const uint32_t code = min((mag + 0x3ffu + ((mag >> 11) & 1u)) >> 11, REST_MAXCODE);
  return code | ((u >> 16) & 0x8000u);
}
__device__ __forceinline__ void rest_dec(uint32_t u, float rdec, float& r, float& b) {
  r = rest_dec1(u & 0xffffu, rdec);
  b = rest_dec1(u >> 16, rdec);
}
__device__ __forceinline__ uint32_t rest_enc(float r, float b, float renc) {
  return rest_enc1(r, renc) | (rest_enc1(b, renc) << 16);
}
/* |h| rounds beyond the largest code: |h| 2^(-118 - s) at or above the
 * midpoint between code 0x7fff and the next pattern, 2^-120 (2 - 2^-13) */
__device__ __forceinline__ bool rest_out_of_range(float r, float b, float renc) {
  return !(fmaxf(fabsf(r), fabsf(b)) * renc < 0x1.fff8p-120f);
}

/* Rest-pair exponent offset s of a configuration: the smallest s >= 0 whose
 * range 2^(s - 1) holds a rest population of twice the nominal colour
 * density: f_0 = w_0 rho_k with rho_k up to max(1, rho0) in the plain model
 * (s = 0 stores colour densities up to rho0 + 1.5), f_0 up to alpha_k rho_k
 * <= rho_k with rho_k up to max(densR, densB) in the density-ratio model. */
inline int rest_offset(const lbm_config& cfg) {
  const double rho0 = cfg.store_rho0;
  const double fmax = cfg.diffdens ? 2.0 * std::max(1.0, (double)std::max(cfg.dens_r, cfg.dens_b))
                                   : 2.0 * (double)W0 * std::max(1.0, rho0);
  int s = 0;
  while (s < 8 && std::ldexp(1.0, s - 1) < fmax - (double)W0 * rho0) ++s;
  return s;
}

/* both colours of direction l at node n of a population level; rdec and
 * renc are the rest-pair scales (16-bit storage) */
template <class S>
__device__ __forceinline__ void ld_pair(const void* pops, size_t N, int l, size_t n, float rho0, float rdec, float& r,
                                        float& b) {
  if (sizeof(S) == 2 && l == 0) {
    rest_dec(reinterpret_cast<const uint32_t*>(pops)[n], rdec, r, b);
    r = r + W0 * rho0;
    b = b + W0 * rho0;
    return;
  }
  Pair<S>::ld((const S*)pops, (size_t)l * N + n, W(l) * rho0, r, b);
}
template <class S>
__device__ __forceinline__ void st_pair(void* pops, size_t N, int l, size_t n, float r, float b, float rho0, float renc) {
  if (sizeof(S) == 2 && l == 0) {
    reinterpret_cast<uint32_t*>(pops)[n] = rest_enc(r - W0 * rho0, b - W0 * rho0, renc);
    return;
  }
  Pair<S>::st((S*)pops, (size_t)l * N + n, r, b, W(l) * rho0);
}
/* raw copy of one pair (no conversion) */
template <class S>
__device__ __forceinline__ void cp_pair(void* dst, size_t di, const void* src, size_t si) {
  if constexpr (sizeof(S) == 4)
    reinterpret_cast<uint64_t*>(dst)[di] = reinterpret_cast<const uint64_t*>(src)[si];
  else
    reinterpret_cast<uint32_t*>(dst)[di] = reinterpret_cast<const uint32_t*>(src)[si];
}
/* one colour c of direction l at node n (host transfers) */
__host__ __device__ constexpr size_t pop_index(size_t N, int c, int l, size_t n) { return 2 * ((size_t)l * N + n) + c; }

/* device status word: first error wins */
struct DevStatus {
  int code;
  int step;
};
__device__ inline void raise_error(DevStatus* st, int code, int step) {
  if (atomicCAS(&st->code, 0, code) == 0) atomicExch(&st->step, step);
}

/* periodic wrap of a coordinate that is at most one box length off */
__host__ __device__ inline int wrapc(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

/* kernel arguments of the fluid path */
struct FluidArgs {
  int nx, ny, nz;
  size_t N;              /* nx*ny*nz                                          */
  const void* src;       /* state buffer (post-streaming), 2 colours x 19 x N */
  void* dst;             /* next state buffer                                  */
  const uint8_t* flags;  /* node types of this step                            */
  float* rho_r;
  float* rho_b;
  float* phi;            /* phase field (rhoR-rhoB)/(rhoR+rhoB) per node       */
  float inv_visc_r, inv_visc_b;
  float sigma, beta;
  float rho0;            /* storage shift                                      */
  float rdec, renc;      /* rest-pair decode / encode scales (16-bit storage)  */
  float min_rho, max_rho;
  DevStatus* status;
  int step;
  /* density-ratio model (diffdens != 0): densR, densB, alphaCG(1:2) */
  int diffdens;
  float dens_r, dens_b;
  float alpha_r, alpha_b;
};

}  // namespace lbm

/* Near-contact closure network 4 -> 64 -> 64 -> 3 (tanh, tanh, linear),
 * parameters W1[64][4] b1[64] W2[64][64] b2[64] W3[3][64] b3[3].  The same
 * fused multiply-add order as the training forward pass in lbm_mlp.hip. */
__device__ inline void lbm_mlp_eval(const float* __restrict__ prm, const float in[4], float out[3]) {
  constexpr int NH = 64, OB1 = NH * 4, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + 3 * NH;
  float a1[NH];
#pragma unroll 8
  for (int j = 0; j < NH; ++j) {
    float z = prm[OB1 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i) z = fmaf(prm[4 * j + i], in[i], z);
    a1[j] = tanhf(z);
  }
  float y0 = prm[OB3], y1 = prm[OB3 + 1], y2 = prm[OB3 + 2];
  float a2[NH];
#pragma unroll 4
  for (int k = 0; k < NH; ++k) {
    float z = prm[OB2 + k];
#pragma unroll 16
    for (int j = 0; j < NH; ++j) z = fmaf(prm[OW2 + NH * k + j], a1[j], z);
    a2[k] = tanhf(z);
  }
#pragma unroll 16
  for (int k = 0; k < NH; ++k) y0 = fmaf(prm[OW3 + k], a2[k], y0);
#pragma unroll 16
  for (int k = 0; k < NH; ++k) y1 = fmaf(prm[OW3 + NH + k], a2[k], y1);
#pragma unroll 16
  for (int k = 0; k < NH; ++k) y2 = fmaf(prm[OW3 + 2 * NH + k], a2[k], y2);
  out[0] = y0;
  out[1] = y1;
  out[2] = y2;
}

/* ---------------------------------------------------------------------- */
/* internal context (host side) */
struct lbm_ctx {
  lbm_config cfg;
  int nx, ny, nz;
  size_t N;
  hipStream_t stream;
  int storage;
  size_t esize;           /* bytes per stored population                      */
  void* pop[2];           /* two time levels, each 2 x 19 x N                  */
  float* rho_r;
  float* rho_b;
  float* phi;
  uint8_t* flags[2];      /* flags of the two time levels                      */
  lbm::DevStatus* d_status;
  int cur;                /* index of the state buffer / current time level    */
  long step;
  bool initialised;
  long generation;        /* bumped when device buffers are (re)allocated    */
  int host_err;
  long host_err_step;
  char host_msg[160];
  /* particles */
  int np;
  double* d_pos[2];       /* (13, np) per time level                           */
  double* d_force[2];     /* (12, np) per time level                           */
  double* d_myf;          /* (3, np) FP64 sums: bounce-back force              */
  double* d_myf2;         /* (3, np) fill/erase + pair forces                   */
  double* d_myt;
  double* d_myt2;
  int* d_stamp[2];        /* (3, np) centres stamped into flags[b]             */
  uint8_t* d_restamp;     /* (np) particles the flag pass re-stamps this step  */
  uint32_t* d_key[2];     /* (N) owner key per time level                      */
  int* d_claim;           /* (N) make_fluid writer resolution                  */
  int* d_nbr;             /* (32, np) neighbour lists of the link kernel        */
  int* d_nnbr;            /* (np)                                               */
  double* d_contrib;      /* (6, 32, np) link momentum shares between particles */
  uint32_t* d_smask;      /* (ntlist, np) link masks of owned surface nodes      */
  int8_t* d_tmpl;         /* (4, ntmpl) template offsets + value               */
  int ntmpl;
  int8_t* d_ring;         /* (4, nring) D3Q27 ring of the template's surface */
  int nring;
  int8_t* d_tlist;        /* (4, ntlist) surface (spherelist) template nodes */
  int ntlist;
  int8_t* d_ball;         /* (4, nball) offsets with r^2 <= rmax^2; [3]=1 if <   */
  int nball;
  int rmax;
  bool shell_box_ok;      /* template and ring offsets within rmax (LDS shell pass) */
  /* near-contact closure: parameters + normalisation (LBM_CLOSURE_NFLOAT) */
  float* d_closure;       /* NULL when disabled                                */
  float closure_gap_max;
  /* collide-stream kernel: LBM_KERNEL_*, fused tile index and z chunk */
  int variant;
  int tile;
  int zchunk;
  /* timing: one event pair per collide-stream launch, read back lazily */
  hipEvent_t* ev_pairs;   /* 2 * ev_cap events                                 */
  int ev_cap;
  int ev_used;
  bool timing;
  double t_total_ms;
  long t_launches;
};

/* batched (ensemble) step: opaque per-phase state, lbm_fused.hip / lbm_particles.hip */
struct lbm_fluid_batch;
struct lbm_part_batch;
size_t lbm_particles_shell_lds(const lbm_ctx* c);
lbm_fluid_batch* lbm_fluid_batch_create(lbm_ctx* const* boxes, int nbox, int* rc);
void lbm_fluid_batch_destroy(lbm_fluid_batch* b);
bool lbm_fluid_batch_tile_ok(int tile);
int lbm_fluid_batch_config(lbm_fluid_batch* b, int tile, int zc);
int lbm_fluid_batch_launch(lbm_fluid_batch* b, int parity, long step, hipStream_t st, lbm::FluidArgs* d_override);
lbm::FluidArgs* lbm_fluid_batch_scratch_args(lbm_fluid_batch* b, int parity, void* const* scratch);
int lbm_fluid_batch_items(const lbm_fluid_batch* b);
lbm_part_batch* lbm_part_batch_create(lbm_ctx* const* boxes, int nbox, int* rc);
void lbm_part_batch_destroy(lbm_part_batch* b);
int lbm_part_batch_count(const lbm_part_batch* b);
int lbm_part_batch_pre(lbm_part_batch* b, int parity, long step, hipStream_t st);
int lbm_part_batch_post(lbm_part_batch* b, int parity, long step, hipStream_t st);

/* records the start/stop events around one timed launch (lbm_api.cpp) */
void lbm_timing_begin(lbm_ctx* c);
void lbm_timing_end(lbm_ctx* c);

/* launchers implemented in lbm_fluid.hip / lbm_particles.hip */
lbm::FluidArgs lbm_fluid_args(lbm_ctx* c);
int lbm_launch_density(lbm_ctx* c);
int lbm_launch_collide(lbm_ctx* c);
int lbm_launch_fused(lbm_ctx* c);
int lbm_launch_fused_with(lbm_ctx* c, int tile, int zc, void* dst);
int lbm_fused_num_tiles();
bool lbm_fused_tile_valid(const lbm_ctx* c, int tile, int zc);
void lbm_fused_tile_shape(int tile, int* tx, int* ty, int* wx);
int lbm_launch_rho_from_pops(lbm_ctx* c, float* d_rr, float* d_rb);
int lbm_launch_copy_probe(const void* a, void* b, size_t n_float4, hipStream_t st);
int lbm_launch_stream_probe(const void* a, void* b, uint32_t nodes, hipStream_t st);
int lbm_launch_init_eq(lbm_ctx* c, int type, const float u0[3], const float* d_vel);
int lbm_launch_moments(lbm_ctx* c, float* d_rr, float* d_rb, float* d_vel, float* d_phase);
int lbm_launch_sums(lbm_ctx* c, double* d_out);
int lbm_launch_unpack_pops(lbm_ctx* c, int color, float* d_out);
int lbm_launch_pack_pops(lbm_ctx* c, const float* d_pr, const float* d_pb);
int lbm_particles_step_pre(lbm_ctx* c);   /* flags, erase, (density), fill, wetting */
int lbm_particles_flags(lbm_ctx* c, int level);
int lbm_particles_erase(lbm_ctx* c);
int lbm_particles_fill(lbm_ctx* c);
int lbm_particles_wetting(lbm_ctx* c);
int lbm_particles_shell_wetting(lbm_ctx* c);
int lbm_particles_max_neighbours();
int lbm_particles_bounce(lbm_ctx* c);
int lbm_particles_integrate(lbm_ctx* c);
