/*
 * ls9_oracle.h -- CPU oracle for the Limit-sticky9 force/energy/RK path.
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Arrays are the reference's AoS layout: x, v, a are [n][3] doubles,
 * m is [n]; index 0 is the black hole and n counts it.
 * c is the speed of light; c = 1 reproduces Limit-sticky9.c bit for bit.
 */
#ifndef LS9_ORACLE_H
#define LS9_ORACLE_H

/* force()  Limit-sticky9.c:21-424 (a is overwritten) */
void orc_force(const double *x, const double *v, const double *m,
               double eps, double c, double *a, int n);
/* rows[0..nrows) (stars, 1 <= rows[r] < n) of force(); out is [nrows][3]
 * and bitwise those rows of orc_force */
void orc_force_rows(const double *x, const double *v, const double *m,
                    double eps, double c, const int *rows, int nrows,
                    double *out, int n);
/* force() over `threads` OpenMP threads (a multi-core CPU baseline): star
 * rows bitwise those of orc_force, a[0] to the reordering of one sum;
 * returns the number of threads that ran */
int orc_force_omp(const double *x, const double *v, const double *m, double eps,
                  double c, double *a, int n, int threads);
void orc_force_only(const double *x, const double *v, const double *m,
                    double eps, double c, double *a, int n);

/* energy terms of Limit-sticky9.c:445-558:
 * out = {kin, pot, pot_pn/c^2, sum(pot_pn2)/c^2, total} */
void orc_energy_terms(const double *x, const double *v, const double *m,
                      int n, double eps, double c, double *out);

/* initial_energy()  Limit-sticky9.c:566-714 (prints like the reference) */
void orc_initial_energy(const double *x, const double *v, const double *m,
                        int n, double *init_ene, double eps, double c);

/* energy()  Limit-sticky9.c:427-564 (prints like the reference) */
void orc_energy(double t, const double *x, const double *v, const double *m,
                int n, double init_ene, double eps, double c);

/* `steps` steps of the stage driver of Limit-sticky9.c:841-893 */
void orc_q1_steps(double *x, double *v, const double *m, int n,
                  double eps, double c, double dt, int steps);

/* Complex-step versions (ls9_oracle_cs.c): the same code on double
 * complex arrays (real and imaginary parts interleaved); eps, c, dt real. */
void orc_force_cs(const double *x, const double *v, const double *m,
                  double eps, double c, double *a, int n);
void orc_q1_steps_cs(double *x, double *v, const double *m, int n,
                     double eps, double c, double dt, int steps);

#endif
