/*
 * lbm_oracle.h -- CPU ORACLE, TEST INFRASTRUCTURE ONLY.
 *
 * C entry points of the CPU restatement of the reference LBcuda step
 * (lbm_oracle.cpp).  They mirror the lbm_* functions of lbcuda_amd.h so a
 * test can drive the oracle and the HIP library with the same arguments.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H
#include <stdint.h>
#include "../include/lbcuda_amd.h"
#ifdef __cplusplus
extern "C" {
#endif
typedef struct orc_ctx orc_ctx;
int orc_create(const lbm_config* cfg, int real_bits, orc_ctx** out);  /* 32 or 64 */
int orc_destroy(orc_ctx* c);
int orc_init_fields(orc_ctx* c, int type, const float u0[3], const float* rho_r,
                    const float* rho_b, const float* vel);
int orc_set_pops(orc_ctx* c, const float* pops_r, const float* pops_b);
int orc_set_particles(orc_ctx* c, int n, const double* pos13);
int orc_step(orc_ctx* c, int n);
int orc_get_pops(orc_ctx* c, int color, double* out);        /* (nx,ny,nz,19) */
int orc_get_rho(orc_ctx* c, double* rho_r, double* rho_b);   /* (nx,ny,nz)    */
int orc_get_flags(orc_ctx* c, int8_t* out);
int orc_get_particles(orc_ctx* c, double* pos13, double* force12);
int orc_status(orc_ctx* c, int* code, long* step, char* msg, int len);
int orc_template(orc_ctx* c, int* rmax, int8_t* out, int len);
int orc_num_threads(void);
#ifdef __cplusplus
}
#endif
#endif
