/* CPU restatement of the Karma model step (karma/karma3d.f90): test
 * infrastructure, the checker of the HIP Karma step. */
#ifndef KARMA_ORACLE_H
#define KARMA_ORACLE_H

typedef struct {
  double lambda, eps4, dx, dt, Fnoise, k;
  double a2;  /* 0.6267: D = a2 lambda (karma3d.f90:86, 112) */
} pfo_karma_par;

void pfo_karma_init(int n, double dx, double R0, double Omega, double *phi, double *U);
void pfo_karma_mirror(int n, double *f);
double pfo_karma_noise(int i, int j, int k, int it);
void pfo_karma_step(int n, const pfo_karma_par *p, int it, double *phi, double *U);

#endif
