/* C-ABI of the quantitative Karma model (karma/karma3d.f90) on the MI355X.
 *
 * The reference is a stand-alone program with no subroutine interface: its
 * time loop (karma/karma3d.f90:159-312) and seed initialisation
 * (karma/karma3d.f90:130-144) are what these entry points replace, driven
 * by the parameters of its namelist (karma/karma3d.f90:75-89).  The
 * program around them (namelist, tip, output files) is bin/karma3d.  Fields are
 * FP64 and stay resident in HBM between calls; status codes are those of
 * pf_capi.h (PF_OK, PF_E_ARG, PF_E_HIP, PF_E_NOMEM) with pf_last_error(). */
#ifndef PF_KARMA_H
#define PF_KARMA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pf_karma pf_karma;

/* karma.nml (karma/karma3d.f90:75-89); the derived D = a2 lambda and
 * Omega = (c_l0 - C0)/((1-k) c_l0), c_l0 = (Tm - TP)/mliq, are formed as
 * karma/karma3d.f90:110-112 forms them. */
typedef struct {
  int32_t n;      /* cells per direction (>= 2) */
  int32_t pad_;
  double lambda, eps4, dx, dt, R0, Fnoise;
  double k, Tm, mliq, C0, TP;
} pf_karma_params;

/* Allocates the fields on the current device and sets the 1/8 seed at the
 * corner (karma/karma3d.f90:130-144); step counter 0. */
int pf_karma_create(const pf_karma_params *p, pf_karma **out);
/* nsteps steps of karma/karma3d.f90:159-291: mirror ghosts, phase equation,
 * solute equation with the anti-trapping current, clamp and advance.
 * Asynchronous on the model's stream. */
int pf_karma_run(pf_karma *s, int32_t nsteps);
/* Kernel form of the following steps: on (the default) the z-march kernels
 * (a work-group keeps three planes of its tile in LDS and marches along z),
 * off one thread per cell reading its neighbours through the caches.  Both
 * give the same bits. */
int pf_karma_set_zmarch(pf_karma *s, int32_t on);
/* Synchronous: the cells 1..n of phi and U (n^3 doubles each, x fastest,
 * as the reference's arrays phi(1:N,1:N,1:N), U(1:N,1:N,1:N)); either may be
 * null.  *it (nullable) = steps taken. */
int pf_karma_download(pf_karma *s, double *phi, double *U, int32_t *it);
/* Waits for the model's stream. */
int pf_karma_synchronize(pf_karma *s);
/* Frees the fields and the stream; reports a pending device error. */
int pf_karma_destroy(pf_karma *s);

#ifdef __cplusplus
}
#endif

#endif
