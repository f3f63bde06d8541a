/*
 * pdac_oracle.h -- CPU restatement of PDAC's semi-implicit per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP
 * implementation is compared with; only the tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it.  The product never links it.
 *
 * It exposes the same entry points as include/pdacgpu.h with a pdac_cpu_
 * prefix, so a test drives both with identical calls.
 */
#ifndef PDAC_ORACLE_H
#define PDAC_ORACLE_H

#include "../include/pdacgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdac_cpu_ctx pdac_cpu_ctx;

int pdac_cpu_create(pdac_cpu_ctx **ctx, const pdac_dims *dims);
int pdac_cpu_destroy(pdac_cpu_ctx *ctx);
int pdac_cpu_set_grid(pdac_cpu_ctx *ctx, const double *dx, const double *dy,
                      const double *dz, const int32_t *flag,
                      const double *p_atm, const double *t_atm);
int pdac_cpu_set_params(pdac_cpu_ctx *ctx, const pdac_params *prm);
int pdac_cpu_set_ib(pdac_cpu_ctx *ctx, const int32_t *nfp, const int32_t *fpi, const double *fpd,
                    const double *coords, const int32_t *bmask, const double *ivf);
int pdac_cpu_set_ib_off(pdac_cpu_ctx *ctx, int bits);
int pdac_cpu_put(pdac_cpu_ctx *ctx, int field, const double *host, int64_t n);
int pdac_cpu_get(pdac_cpu_ctx *ctx, int field, double *host, int64_t n);
int pdac_cpu_field_ncomp(const pdac_cpu_ctx *ctx, int field);
int pdac_cpu_cnvert(pdac_cpu_ctx *ctx);
int pdac_cpu_boundary(pdac_cpu_ctx *ctx);
int pdac_cpu_fieldn(pdac_cpu_ctx *ctx);
int pdac_cpu_sgsg(pdac_cpu_ctx *ctx);
int pdac_cpu_begin_step(pdac_cpu_ctx *ctx);
int pdac_cpu_tilde(pdac_cpu_ctx *ctx, double dt);
int pdac_cpu_iter(pdac_cpu_ctx *ctx, double dt, int32_t *nit, int32_t *converged);
int pdac_cpu_ygas(pdac_cpu_ctx *ctx, double dt);
int pdac_cpu_htilde(pdac_cpu_ctx *ctx, double dt);
int pdac_cpu_ftem(pdac_cpu_ctx *ctx, double dt, int32_t *eos_fail);
int pdac_cpu_initialize_fieldn(pdac_cpu_ctx *ctx, double dt);
int pdac_cpu_explicit_iter(pdac_cpu_ctx *ctx, double dt, int32_t *eos_fail, double *cfl);
/* how calc_all_fluxes reads the west/south/bottom second neighbours of the
 * mass-flux stencil, which nb leaves unset: 1 = as 0, 0 = the reflected
 * second neighbours */
int pdac_cpu_set_explicit_stale(pdac_cpu_ctx *ctx, int zero);
int pdac_cpu_step(pdac_cpu_ctx *ctx, double dt, int32_t rungekut, int32_t nsteps,
                  int32_t *nit_last, int64_t *nit_total);

/* closed-form pieces exported for unit pinning */
double pdac_cpu_limiter(int lim_type, double beta, double erre);
double pdac_cpu_kdrags(const pdac_params *prm, int is, double du, double dv,
                       double dw, double ep, double rgp, double rlk, double mug);
double pdac_cpu_sdrag(const pdac_params *prm, int is, double du, double dv,
                      double dw, double ep, double rgp, double rlk, double mug);
int pdac_cpu_caloric_eosg(const pdac_params *prm, int ngas, double sieg,
                          const double *ygc, double *tg, double *cg);
void pdac_cpu_gauss_jordan(int np, const double *plim, double *a, double *b);
double pdac_cpu_extrapolate(double f2, double f1, double d12, double dx, int order);
void pdac_cpu_hcapg(const pdac_params *prm, const double *gmw, double tg, double *cpc);
int pdac_cpu_get_momflux(pdac_cpu_ctx *ctx, int comp, int face, double *out);

#ifdef __cplusplus
}
#endif
/* threads the order-independent cell loops run on (OpenMP) */
int pdac_cpu_threads(void);

#endif
