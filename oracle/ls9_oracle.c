/*
 * ls9_oracle.c -- CPU oracle for the Limit-sticky9 force / energy / RK path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The HIP library (pn_amd/libpn1.so) never links or calls it.
 *
 * It restates, in FP64 and in the same floating-point evaluation order, the
 * reference C program Limit-sticky9.c:
 *   force()           Limit-sticky9.c:21-424
 *   energy()          Limit-sticky9.c:427-564
 *   initial_energy()  Limit-sticky9.c:566-714
 *   RK stage driver   Limit-sticky9.c:841-893
 * with one addition: the speed of light c.  The reference has no 1/c^2
 * factor (it is the c = 1 case).  Every 1PN increment that the reference
 * adds to an acceleration or an energy is multiplied here by lam = 1/(c*c);
 * at c = 1 that multiplication is by exactly 1.0, so results are bitwise
 * those of the reference when both are built without FMA contraction
 * (gcc -O2 on x86-64 emits none; the Makefile also passes -ffp-contract=off).
 *
 * Every expression keeps the reference's parenthesisation: C evaluates
 * a*b*c as (a*b)*c and a+b+c as (a+b)+c, and the comments give the tree.
 * Working arrays are heap-allocated (the reference keeps NMAX-sized arrays
 * on the stack, Limit-sticky9.c:26, 435, 441).
 *
 * Pinning (what checks this file): bitwise equality with the compiled
 * reference at c = 1 (force, initial_energy total), agreement with the
 * reference on rescaled inputs at c != 1 (a_c = c^2 a_ref(x, v/c, m/c^2)),
 * the n = 1 and n = 2 closed forms, and the invariants in tests/.
 */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ls9_oracle.h"

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p) {
        fprintf(stderr, "ls9_oracle: out of memory\n");
        exit(1);
    }
    return p;
}

/* force() and the stage driver (also dot3 / ref_rsqrt for the energy
 * code below), in FP64 */
#define REAL double
#define SQRT(x) sqrt(x)
#define FN(name) name
#include "ls9_force_body.h"
#undef REAL
#undef SQRT
#undef FN

/* ------------------------------------------------------------------ */
/* energy() / initial_energy() body  Limit-sticky9.c:445-558, 583-699   */
/* ------------------------------------------------------------------ */
void orc_energy_terms(const double *xs, const double *vs, const double *m,
                      int n, double eps, double c, double *out)
{
    const double (*x)[3] = (const double (*)[3])xs;
    const double (*v)[3] = (const double (*)[3])vs;
    const double lam = 1.0 / (c * c);
    const double eps2 = eps * eps;
    double *pot = xcalloc((size_t)n, sizeof *pot);
    double *pn2 = xcalloc((size_t)n, sizeof *pn2);
    double d[3], da[3], db[3], dab[3];

    double kin = 0.0;
    for (int i = 0; i < n; i++) kin += m[i] * dot3(v[i], v[i]);
    kin *= 0.5;

    for (int i = 0; i < n; i++) {
        double p = 0.0;
        for (int j = 0; j < n; j++) {
            for (int k = 0; k < 3; k++) d[k] = x[j][k] - x[i][k];
            double rinv = ref_rsqrt(dot3(d, d) + eps2);
            p -= rinv * m[j];
        }
        p += m[i] / sqrt(eps2);
        pot[i] = p * m[i];
    }
    double potn = 0.0;
    for (int i = 0; i < n; i++) potn += pot[i];
    potn *= 0.5;

    double potpn = 0.0;
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) d[k] = x[i][k] - x[0][k];
        double rai = ref_rsqrt(dot3(d, d) + eps2);
        double mrai = m[0] * rai;
        double vi2 = dot3(v[i], v[i]);
        /* 0.375 m vi2 vi2 + 1.5 mrai m vi2 + 0.5 mrai mrai m
         *   + 0.5 mrai m ((3 vi2 - 7 v0.vi) - (d.v0)(d.vi) rai rai) */
        potpn += 0.375 * m[i] * vi2 * vi2 + 1.5 * mrai * m[i] * vi2
                 + 0.5 * mrai * mrai * m[i]
                 + 0.5 * mrai * m[i]
                   * (3.0 * vi2 - 7.0 * dot3(v[0], v[i])
                      - dot3(d, v[0]) * dot3(d, v[i]) * rai * rai);
    }

    for (int i = 1; i < n; i++) {
        double p = 0.0;
        for (int j = 1; j < n; j++) {
            for (int k = 0; k < 3; k++) {
                da[k] = x[i][k] - x[0][k];
                db[k] = x[j][k] - x[0][k];
                dab[k] = x[j][k] - x[i][k];
            }
            double rai = ref_rsqrt(dot3(da, da) + eps2);
            double rb2e = dot3(db, db) + eps2;
            double rab2e = dot3(dab, dab) + eps2;
            double rbi = ref_rsqrt(rb2e);
            double rabi = ref_rsqrt(rab2e);
            double vi2 = dot3(v[i], v[i]);
            /* dab.vi is written dab0*vi0 + ... (:531-532); same products */
            p += 0.25 * rabi * m[i] * m[j]
                     * (6.0 * vi2 - 7.0 * dot3(v[i], v[j])
                        - dot3(dab, v[i]) * dot3(dab, v[j]) / rab2e)
                 + m[0] * m[i] * m[j] * rai * rabi
                 + 0.5 * m[0] * m[i] * m[j] * rai * rbi;
        }
        pn2[i] = p;
    }
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) da[k] = x[i][k] - x[0][k];
        double rai = ref_rsqrt(dot3(da, da) + eps2);
        double vi2 = dot3(v[i], v[i]);
        pn2[i] -= 0.25 * m[i] * m[i] / eps * (-vi2)
                  + m[0] * m[i] * m[i] * (rai / eps + 0.5 * rai * rai);
    }

    double ene = kin + potn + lam * potpn;
    double pn2sum = 0.0;
    for (int i = 1; i < n; i++) {
        ene += lam * pn2[i];
        pn2sum += pn2[i];
    }
    out[0] = kin;
    out[1] = potn;
    out[2] = lam * potpn;
    out[3] = lam * pn2sum;
    out[4] = ene;
    free(pot);
    free(pn2);
}

/* Print block of Limit-sticky9.c:560-563 / 701-703. */
static void print_energy_block(double t, const double *e)
{
    printf("time = %g\n", t);
    printf("pot = %22.15e kin = %22.15e \n pot_pn = %22.15e \n  total= %22.15e ratio = %e\n",
           e[1], e[0], e[2], e[4], e[0] / e[1]);
}

void orc_initial_energy(const double *x, const double *v, const double *m,
                        int n, double *init_ene, double eps, double c)
{
    double e[5];
    orc_energy_terms(x, v, m, n, eps, c, e);
    print_energy_block(0.0, e);
    *init_ene = e[4];
}

void orc_energy(double t, const double *x, const double *v, const double *m,
                int n, double init_ene, double eps, double c)
{
    double e[5];
    orc_energy_terms(x, v, m, n, eps, c, e);
    double de = (init_ene - e[4]) / e[4];
    print_energy_block(t, e);
    printf(" DE = %e %g\n", de, t);
}

/* force() with its O(n^2) loops spread over `threads` OpenMP threads: a CPU
 * baseline on all of a host's cores (not the reference, which is serial).
 * Star rows are bitwise those of orc_force (each row is one thread's serial
 * loop, in the reference's order); the BH's three-body a1c sum is split
 * into per-thread partials added in thread order, so a[0] differs from the
 * serial sum in the last bits.  Returns the number of threads that ran. */
int orc_force_omp(const double *xs, const double *vs, const double *m, double eps, double c,
                  double *as, int n, int threads)
{
    const double (*x)[3] = (const double (*)[3])xs;
    const double (*v)[3] = (const double (*)[3])vs;
    double (*a)[3] = (double (*)[3])as;
    const double lam = 1.0 / (c * c);
    const double eps2 = eps * eps;
    double acc1[3], a1c[3];
    if (threads < 1) threads = 1;
    double (*part)[3] = xcalloc((size_t)threads, sizeof *part);
    int ran = 1;

#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int i = 0; i < n; i++) row_newton(x, m, eps2, i, n, a[i]);

    bh_a1(x, v, m, eps2, n, acc1);
    for (int k = 0; k < 3; k++) a[0][k] += lam * acc1[k];
    bh_a1c_prefix(x, m, eps2, n, a1c);
#pragma omp parallel num_threads(threads)
    {
        const int t = omp_get_thread_num();
#pragma omp single
        ran = omp_get_num_threads();
#pragma omp for schedule(dynamic, 16)
        for (int i = 1; i < n; i++) bh_a1c_cross_row(x, v, m, eps2, i, n, part[t]);
    }
    for (int t = 0; t < threads; t++)
        for (int k = 0; k < 3; k++) a1c[k] += part[t][k];
    bh_a1c_removal(x, v, m, eps, eps2, n, a1c);
    for (int k = 0; k < 3; k++) a[0][k] += lam * a1c[k];

#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int i = 1; i < n; i++) {
        double t[3], s[3];
        row_star_bh_term(x, v, m, eps2, i, t);
        for (int k = 0; k < 3; k++) a[i][k] += lam * t[k];
        row_aa_cross_term(x, v, m, eps2, i, t);
        for (int k = 0; k < 3; k++) a[i][k] += lam * t[k];
        row_ac(x, v, m, eps2, i, n, s);
        for (int k = 0; k < 3; k++) a[i][k] += lam * s[k];
        row_self_term(x, v, m, eps, eps2, i, t);
        for (int k = 0; k < 3; k++) a[i][k] -= lam * t[k];
    }
    free(part);
    return ran;
}

/* One force evaluation, timed by the caller: a plain wrapper so a CPU
 * baseline can time the oracle without the RK driver around it. */
void orc_force_only(const double *x, const double *v, const double *m,
                    double eps, double c, double *a, int n)
{
    orc_force(x, v, m, eps, c, a, n);
}
