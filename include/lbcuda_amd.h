/*
 * lbcuda_amd.h -- C ABI of the MI355X (gfx950) lattice-Boltzmann time step.
 *
 * The reference (copmat/LBcuda) has no plugin or FFI layer: its host program
 * main.CUF launches CUDA-Fortran kernels with chevrons inside the step loop
 * (main.CUF:125-308) and keeps all state in module globals
 * (kernels_fluid.CUF:9-48, dimensions_m.CUF:318-347).  This header is the
 * drop-in boundary that replaces those launches and uploads.  Every entry
 * point below names the reference code it stands in for.  A Fortran host binds
 * it with ISO_C_BINDING (host/lbcuda_amd_mod.f90); Python binds it with ctypes
 * (lbcuda_amd/lib.py).
 *
 * Conventions at the boundary (reference: main.CUF:64,278, kernels_fluid.CUF):
 *   - host arrays are Fortran-ordered (x fastest), interior nodes only
 *     (1..nx, 1..ny, 1..nz), no halo;
 *   - population arrays are (nx, ny, nz, 0:18);
 *   - particle arrays follow pos(13, N) and forceAtoms(12, N) of
 *     dimensions_m.CUF:355-362, stored as double;
 *   - particle indices are 0-based in C, 1-based in Fortran;
 *   - every function returns 0 on success or a positive LBM_ERR_* code.
 */
#ifndef LBCUDA_AMD_H
#define LBCUDA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Population storage formats (the arithmetic is always FP32). */
#define LBM_STORE_F32  0
#define LBM_STORE_F16  1   /* f_i - w_i*rho0 stored as IEEE half  */
#define LBM_STORE_BF16 2   /* f_i - w_i*rho0 stored as bfloat16   */
/* With either 16-bit storage the rest pair (i = 0) is a 16-bit float with 3
 * exponent and 12 mantissa bits, |h| up to 2^(s-1) (1 - 2^-13), where the
 * offset s is the smallest that holds twice the nominal colour density
 * (s = 0, colour densities up to store_rho0 + 1.5, for the plain model);
 * a rest value beyond it raises LBM_ERR_STORE_RANGE. */

/* Wetting of particle surface nodes. */
#define LBM_WET_NONE    0  /* surface rho stays MINDENS (init_rho*_CG)              */
#define LBM_WET_JANUS   1  /* compute_virtual_rho, kernels_fluid_PART.CUF:299-352   */
#define LBM_WET_NEUTRAL 2  /* compute_densities_wall, kernels_fluid_PART.CUF:267-297 */

/* Initial-state types: input.dat line 4 (Tests/01/input.dat:4, main.CUF:2531-2608). */
#define LBM_INIT_DROPLET   1   /* setup2          kernels_fluid_CG.CUF:140-201 */
#define LBM_INIT_ZPLANES   2   /* setup2_zplanes  kernels_fluid_CG.CUF:203-256 */
#define LBM_INIT_RHO       3   /* rho from host + setupPops  main.CUF:2561-2568 */
#define LBM_INIT_POPS      4   /* rho and pops from host     main.CUF:2570-2577 */
#define LBM_INIT_CYLINDER  5   /* setupCyl        kernels_fluid_CG.CUF:258-311 */
#define LBM_INIT_RANDOM    6   /* host random rho + setupPops main.CUF:2582-2595 */
#define LBM_INIT_VELFIELD  7   /* rho and per-node velocity from host (no reference
                                  counterpart: needed for Taylor-Green starts)     */

/* Collide-stream kernel variants (lbm_set_kernel). */
#define LBM_KERNEL_FUSED   0   /* one pass: densities, phase field, collision and
                                  push streaming, z-marching tiles (default)    */
#define LBM_KERNEL_TWOPASS 1   /* density pre-pass + collide-push, the
                                  reference's init_rho*_CG / time_step_CG split */

/* Error codes (the reference stores __LINE__ in stop_d and aborts). */
#define LBM_OK               0
#define LBM_ERR_ARG          1   /* bad argument / config                      */
#define LBM_ERR_ALLOC        2   /* device or host allocation failed           */
#define LBM_ERR_HIP          3   /* HIP runtime error                          */
#define LBM_ERR_RHO_RANGE    4   /* kernels_fluid_CG.CUF:438-441 (init_rhoR_CG) */
#define LBM_ERR_BIG_FORCE    5   /* kernels_fluid_PART.CUF:1532-1537 (nve_lf)   */
#define LBM_ERR_EXTREME_VEL  6   /* kernels_fluid_PART.CUF:1544-1548 (nve_lf)   */
#define LBM_ERR_OUT_OF_BOX   7   /* kernels_fluid_PART.CUF:1581-1598 (nve_lf)   */
#define LBM_ERR_PART_SETUP   8   /* main.CUF:1207-1221 (SetupParts checks)      */
#define LBM_ERR_STATE        9   /* call order (e.g. step before init)          */
#define LBM_ERR_BATCH       10   /* batched launch configuration                */
#define LBM_ERR_LIST_OVERFLOW 11 /* more than 32 particles near one particle
                                    (the reference stops on >1000 particles per
                                    tile list, kernels_fluid_PART.CUF:2436-2439) */
#define LBM_ERR_STORE_RANGE 12   /* a rest population beyond the range of
                                    16-bit storage (no reference counterpart:
                                    the reference stores float) */

/*
 * Run-time configuration.  The reference fixes most of these at compile time
 * (dimensions_m.CUF:7, 50, 59-84, 103-108; defines.h:10-24); here they are
 * run-time so one library serves every box, tile and precision.
 */
typedef struct lbm_config {
  int32_t nx, ny, nz;        /* box (glx, gly, glz), periodic on all axes        */
  float   tau_r, tau_b;      /* relaxation times tauR, tauB (dimensions_m:103-104)*/
  float   sigma;             /* surface tension sigma_CG (dimensions_m:50)        */
  float   beta;              /* recolouring beta_CG (dimensions_m:50)             */
  int32_t storage;           /* LBM_STORE_*                                       */
  float   store_rho0;        /* rho0 of the shifted low-precision storage, 0..1.2 */
  float   min_rho, max_rho;  /* density range guard (kernels_fluid.CUF:35)        */
  /* particles (dimensions_m.CUF:55-84) */
  int32_t with_particles;
  int32_t n_particles;
  int32_t lrotate;
  float   radius;            /* rdim                                              */
  float   rmass;             /* 1/mass; 0 = prescribed constant velocity           */
  float   max_vel;           /* MAX_PARTICLE_VEL                                  */
  float   ext_force[3];      /* ext_fxx..ext_fzz                                   */
  float   ext_torque[3];     /* ext_tqx..ext_tqz                                   */
  int32_t wetting;           /* LBM_WET_*                                          */
  int32_t hertz;             /* withHz                                             */
  float   hz_k, hz_rmin, hz_rcap;
  int32_t lubrication;       /* llubrication (applied to the FP64 sums)            */
  float   lub_k, lub_rmin, lub_rcap;   /* cutoff gap, minimum gap (lattice units) */
  int32_t closure;           /* learned near-contact closure (no reference part)   */
  /* density-ratio colour gradient (the reference's DIFFDENS build,
   * kernels_fluid_CG.CUF:713-800, main.CUF:1047-1085): colour densities
   * densR/densB (one of them must be 1), colour-dependent rest weights
   * alphaCG and the psi/xi correction terms.  Runs on the two-kernel step. */
  int32_t diffdens;
  float   dens_r, dens_b;
  int32_t reserved[5];
} lbm_config;

typedef struct lbm_ctx lbm_ctx;

/* Fill *cfg with the reference defaults (dimensions_m.CUF:50-84, 103-108,
 * kernels_fluid.CUF:35) for an n x n x n box. */
void lbm_config_default(lbm_config* cfg, int32_t n);

/* Allocate device state for one box.  Replaces setupGPUgridAndAlloc
 * (main.CUF:2133-2407) and SetConstants (main.CUF:1043-1162).  Returns
 * LBM_ERR_ARG for boxes beyond 2^27 nodes (512^3) with float storage or
 * 2^28 nodes with 16-bit storage (one direction plane up to 1 GiB). */
int lbm_create(const lbm_config* cfg, lbm_ctx** out);

/* Release everything lbm_create allocated. */
int lbm_destroy(lbm_ctx* ctx);

/* Initial fluid state.  Replaces setupRho_Pops (main.CUF:2531-2608):
 * init_type is an LBM_INIT_* value; u0 is (vx,vy,vz) of input.dat line 5;
 * rho_r/rho_b are (nx,ny,nz) host arrays for LBM_INIT_RHO/RANDOM/VELFIELD/POPS
 * (else NULL); vel is (3,nx,ny,nz) for LBM_INIT_VELFIELD (else NULL). */
int lbm_init_fields(lbm_ctx* ctx, int32_t init_type, const float u0[3],
                    const float* rho_r, const float* rho_b, const float* vel);

/* Populations for LBM_INIT_POPS (read_and_setPops, main.CUF:2481-2529):
 * pops_r/pops_b are (nx,ny,nz,19) host arrays. */
int lbm_set_pops(lbm_ctx* ctx, const float* pops_r, const float* pops_b);

/* Particle state.  Replaces read_input_atom + SetupParts + init_isfluid_CG at
 * step 0 (main.CUF:1187-1300, 2596-2605).  pos is (13,n): x(3), v(3), omega(3),
 * q(4); the same values seed both time levels as the reference does. */
int lbm_set_particles(lbm_ctx* ctx, int32_t n, const double* pos);

/* Advance nsteps time steps.  Replaces one pass of the do-step loop body
 * (main.CUF:127-278) per step, in the same order.  Returns the first error
 * raised by any step (LBM_ERR_RHO_RANGE ...), polled once at the end. */
int lbm_step(lbm_ctx* ctx, int32_t nsteps);

/* Post-streaming populations of colour 0 (red) or 1 (blue), (nx,ny,nz,19),
 * as float (the state buffer a reference OutputVTK would copy, main.CUF:1823). */
int lbm_get_pops(lbm_ctx* ctx, int32_t color, float* out);

/* Density arrays rhoR_d / rhoB_d as they stand (lagging: last computed by the
 * density pass and wetting of the last step), (nx,ny,nz). */
int lbm_get_rho(lbm_ctx* ctx, float* rho_r, float* rho_b);

/* Host moments of the current state (moments2Fl, write_output.f90:975-1041):
 * rho_r, rho_b, phase (nx,ny,nz) and vel (3,nx,ny,nz); any may be NULL. */
int lbm_get_moments(lbm_ctx* ctx, float* rho_r, float* rho_b, float* vel,
                    float* phase);

/* Node-type flags of the current step (myfluid_d(:,:,:,flip)), (nx,ny,nz):
 * 1 fluid, 2 surface (spherelist), 4 inside (spheredead), 5 centre. */
int lbm_get_flags(lbm_ctx* ctx, int8_t* out);

/* Particle state at the current time level: pos (13,n) as lbm_set_particles,
 * force (12,n): F(3), F_bb(3), T(3), T_bb(3) as forceAtoms_d after the last
 * time_step_force_particles (kernels_fluid_PART.CUF:1266-1382). */
int lbm_get_particles(lbm_ctx* ctx, double* pos, double* force);

/* Global sums in FP64 (no reference counterpart; mass and momentum audit):
 * out[0]=sum rhoR, out[1]=sum rhoB, out[2..4]=fluid momentum. */
int lbm_get_sums(lbm_ctx* ctx, double out[5]);

/* Checkpoint of the whole box state as one opaque byte image in the storage
 * format of the box (no reference counterpart: the reference reads initType
 * 3/4 restart files, main.CUF:2444-2529, but writes none, and keeps no
 * particle restart).  size gives the image's bytes; save fills a host buffer
 * of at least that size; load restores it into a context created with an
 * identical lbm_config (and, with particles, after lbm_set_particles with the
 * same count), after which lbm_step continues bit for bit as the saved box
 * would have.  The closure block (lbm_set_closure) is configuration and is
 * not part of the image.  LBM_ERR_ARG for a foreign or truncated image or a
 * different config, LBM_ERR_STATE for a different particle count (both
 * checked before anything is copied); after a HIP error during the copy the
 * box state is undefined until the next successful load or init. */
int lbm_checkpoint_size(lbm_ctx* ctx, int64_t* bytes);
int lbm_checkpoint_save(lbm_ctx* ctx, void* buf, int64_t bytes);
int lbm_checkpoint_load(lbm_ctx* ctx, const void* buf, int64_t bytes);

/* Error state: code (LBM_ERR_*), step at which it was raised, message. */
int lbm_status(lbm_ctx* ctx, int32_t* code, int32_t* step, char* msg,
               int32_t msg_len);

/* Kernel timing with HIP events on the library's stream.  on != 0 brackets
 * every launch of the fluid collide-stream kernel with events; get returns
 * the summed milliseconds and the number of launches since enabling. */
int lbm_kernel_timing(lbm_ctx* ctx, int32_t on);
int lbm_get_kernel_timing(lbm_ctx* ctx, double* total_ms, int64_t* launches);

/* Kernel choice.  The reference fixes the thread-block tile at compile time
 * (TILE1..TILE3, defines.h:14-24, checked in main.CUF:2137-2148); here the
 * variant (LBM_KERNEL_*), the fused kernel's tile (index into the built-in
 * tile table, see lbm_tile_shape) and its z-chunk length are run-time
 * choices.  lbm_set_kernel rejects a tile that does not fit the box, and the
 * fused variant for a density-ratio (diffdens) box, which runs the two-kernel
 * step only. */
int lbm_set_kernel(lbm_ctx* ctx, int32_t variant, int32_t tile, int32_t zchunk);
int lbm_get_kernel(lbm_ctx* ctx, int32_t* variant, int32_t* tile, int32_t* zchunk);

/* Tile table of the fused kernel: number of entries, and the shape of entry
 * tile (x width, y height, wrap = 1 if the tile spans the whole x extent and
 * is only valid when nx equals the width). */
int32_t lbm_num_tiles(void);
int lbm_tile_shape(int32_t tile, int32_t* tx, int32_t* ty, int32_t* wrap);

/* Autotuner (no reference counterpart): times every valid (tile, z chunk)
 * of the fused kernel on this box's current state, reps launches each, into
 * a scratch level (the state is not advanced or changed), keeps the fastest
 * and returns it with its mean launch time in milliseconds.  Call after
 * lbm_init_fields (and lbm_set_particles).  LBM_ERR_ARG for a diffdens box. */
int lbm_autotune(lbm_ctx* ctx, int32_t reps, int32_t* tile, int32_t* zchunk, double* best_ms);

/* Name of the collide-stream kernel variant in use and its algorithmic bytes
 * per lattice update (2 colours x 19 pops x read+write x bytes/value + flag). */
int lbm_kernel_info(lbm_ctx* ctx, char* name, int32_t name_len,
                    double* bytes_per_lu);

/* Host-only helper (no GPU work): the seeded random colour field of
 * initType 6 (set_random_dens_fluids, main.CUF:2610-2684) in reproducible
 * form: rho_r = mean + amp*(2u-1), u = SplitMix64(seed ^ node)/2^64 with
 * node = x + nx*(y + ny*z), rho_b = 1 - rho_r.  The reference draws u from
 * the compiler's RANDOM_NUMBER with its default seed, which no other host can
 * reproduce.  rho_r/rho_b are (nx,ny,nz) host arrays. */
int lbm_fill_random_rho(uint64_t seed, double mean, double amp, int32_t nx, int32_t ny,
                        int32_t nz, float* rho_r, float* rho_b);

/* Learned near-contact closure (no reference counterpart; the reference's
 * near-contact model is the lubrication term of addHertzian,
 * kernels_fluid_PART.CUF:1463-1501, which stays the fallback).
 * A 4 -> 64 -> 64 -> 3 MLP (tanh hidden layers, FP32) maps the pair features
 * (gap/R, v_n, v_t, R) to the corrections (F_n, F_t, T):
 *   gap = r - 2R (r = centre distance), n = unit vector from i to j,
 *   u = v_i - v_j + R (omega_i + omega_j) x n, v_n = u.n (> 0 approaching),
 *   t = unit vector of u - v_n n, v_t = |u - v_n n|;
 *   force on i += -F_n n - F_t t, torque on i += T (n x t).
 * Parameters (LBM_MLP_NPARAM floats): W1[64][4] b1[64] W2[64][64] b2[64]
 * W3[3][64] b3[3].  The network sees (features - in_mean) / in_scale and its
 * outputs are multiplied by out_scale. */
#define LBM_MLP_NPARAM 4675
/* closure block: params, in_mean[4], in_scale[4], out_scale[3], lo[4], hi[4]
 * (the feature box of the training data: outside it, or at gap >= gap_max,
 * the lubrication term applies instead) and gap_max */
#define LBM_CLOSURE_NFLOAT (LBM_MLP_NPARAM + 20)

/* Batch inference on the GPU: x (n,4) normalised features -> y (n,3). */
int lbm_mlp_forward(const float* params, const float* x, float* y, int32_t n);
/* Mean squared error over the n samples and its gradient w.r.t. params. */
int lbm_mlp_loss_grad(const float* params, const float* x, const float* t, int32_t n, double* loss, float* grad);
/* Adam training on the GPU (forward, backward and update kernels of this
 * library); params are read as the initial weights and overwritten.  Each
 * epoch visits the samples in a seeded random order, `batch` at a time. */
int lbm_mlp_train(float* params, const float* x, const float* t, int32_t n, int32_t epochs, int32_t batch, float lr,
                  uint64_t seed, double* final_loss);
/* Enable (block != NULL) or disable the closure in a box's pair forces;
 * requires cfg.closure != 0 at create. */
int lbm_set_closure(lbm_ctx* ctx, const float* block);

/* Ensemble of independent boxes (no reference counterpart: the reference
 * runs one box per process).  Boxes may differ in size, particles and
 * parameters but share the storage format (and are not diffdens boxes).  Each box is an ordinary lbm_ctx
 * (initialise, set particles and read results through the calls above);
 * lbm_ensemble_step advances all of them together: per step one persistent
 * launch runs the fluid of every box from a shared work list of
 * (box, tile, z chunk) items, and each particle kernel runs once for the
 * particles of all boxes.  All boxes must be initialised and at the same
 * step.  lbm_ensemble_autotune picks the tile and z chunk of the persistent
 * launch (tiles without x wrap); lbm_ensemble_set_tile sets them. */
typedef struct lbm_ensemble lbm_ensemble;
int lbm_ensemble_create(int32_t nbox, const lbm_config* cfgs, lbm_ensemble** out);
int lbm_ensemble_destroy(lbm_ensemble* ens);
int lbm_ensemble_box(lbm_ensemble* ens, int32_t i, lbm_ctx** box);
int lbm_ensemble_step(lbm_ensemble* ens, int32_t nsteps);
int lbm_ensemble_set_tile(lbm_ensemble* ens, int32_t tile, int32_t zchunk);
int lbm_ensemble_autotune(lbm_ensemble* ens, int32_t reps, int32_t* tile, int32_t* zchunk, double* best_ms);

/* Device-to-device copy bandwidth in GB/s (read + write bytes over time) of
 * a plain copy kernel over two buffers of the given size, reps launches.
 * The reference probes host transfer bandwidth at start-up
 * (main.CUF:2396-2404); this is the HBM ceiling the step is compared with. */
int lbm_copy_bandwidth(int64_t bytes, int32_t reps, double* gbs);

/* Bandwidth in GB/s (read + write bytes over time) of the fused step's access
 * pattern without its arithmetic, over `nodes` nodes (<= 2^27): 19 direction
 * planes of float pairs read and 19 written, the planes with e_x = +-1
 * shifted by one node (the push).  No reference counterpart; the ceiling the
 * fused kernel's measured traffic is compared with. */
int lbm_stream_bandwidth(int64_t nodes, int32_t reps, double* gbs);

/* Library build identification (gfx target, storage formats). */
const char* lbm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LBCUDA_AMD_H */
