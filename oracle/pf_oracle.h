/*
 * pf_oracle.h -- CPU oracle for the coupled phase-field time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * FP64 CPU path (ileaof/phase-field-project-3D, source/ *.f90 and the *_body.inc files),
 * written to follow its operation order so that it can be checked bit for bit
 * against the reference compiled from source (oracle/ref/).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker / the CPU baseline -- never as the product path.
 *
 * Arrays are column-major (1:nx,1:ny,1:nz), FP64, caller-owned.
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H

#include <stdint.h>
#include "../include/pf_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

double pfo_pseudo_random(int32_t i, int32_t j, int32_t k, int32_t it, int32_t salt,
                         uint64_t seed);
double pfo_ruido_suave(int32_t i, int32_t j, int32_t k, int32_t it, uint64_t seed);
int32_t pfo_w0_e0(double lambda, float ajuste, float sigma, double *W0, double *E0);
double pfo_auxiliar(double TP, double Cl_eq, double Cs_eq, double R, double Vm,
                    double E0, double W0, float Sigma, float Beta, float Ke,
                    double Me, double D_liq, float ajuste);
double pfo_delta_t(double dx, double D_liq);
double pfo_noise_scale(double AAA, double dt, double dx, double D_liq,
                       int32_t ruido_tipo);

/* entrada.f90:226-262 (modes 0..3) plus the seeded modes 5 and 6. */
int pfo_initial_fields(int32_t nx, int32_t ny, int32_t nz, int32_t tipo,
                       double R_semente, double T0, double C0, float Ke,
                       int32_t n_seeds, uint64_t seed_rng, double *phi_old,
                       double *phi_new, double *c_old, double *c_new,
                       double *T_old, double *T_new, double *dpp);

/* One call of each reference subroutine, with its copy / boundary semantics. */
void pfo_temperatura(const pf_params *p, const double *T_old, const double *phi_old,
                     const double *dpp, double *T_new);
void pfo_equacao_campo_fase(pf_params *p, const double *phi_old,
                            const double *c_old, double *phi_new, double *dpp,
                            double probe[3]);
void pfo_concentracao_soluto(const pf_params *p, const double *phi_old,
                             double *c_old, double *c_new);
void pfo_diagnosticos(const pf_params *p, const double *phi_new,
                      const double *c_new, double out[5]);

/* principal.f90:93-226 without the output event: nsteps iterations. */
int pfo_run_steps(pf_params *p, pf_loop *L, int32_t nsteps, double *phi_old,
                  double *phi_new, double *c_old, double *c_new, double *T_old,
                  double *T_new, double *dpp, double *probes);

/* Number of OpenMP threads the oracle's parallel loops use (1 = serial). */
void pfo_set_threads(int32_t n);
int32_t pfo_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
