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
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ls9_oracle.h"

/* ((a0*b0) + (a1*b1)) + (a2*b2): the tree of every 3-term dot product
 * written out in Limit-sticky9.c (e.g. line 43 without the +eps2). */
static inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* Limit-sticky9.c:15-19 */
static inline double ref_rsqrt(double r2) { return 1.0 / sqrt(r2); }

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p) {
        fprintf(stderr, "ls9_oracle: out of memory\n");
        exit(1);
    }
    return p;
}

/* ------------------------------------------------------------------ */
/* force()  Limit-sticky9.c:21-424                                     */
/* ------------------------------------------------------------------ */
void orc_force(const double *xs, const double *vs, const double *m,
               double eps, double c, double *as, int n)
{
    const double (*x)[3] = (const double (*)[3])xs;
    const double (*v)[3] = (const double (*)[3])vs;
    double (*a)[3] = (double (*)[3])as;
    const double lam = 1.0 / (c * c);
    const double eps2 = eps * eps;
    double (*ac)[3] = xcalloc((size_t)n, sizeof *ac);
    double d[3], db[3], dc[3], dbc[3], dvbc[3];
    double acc1[3], a1c[3];

    /* Newtonian, all i, j in [0,n) including self (:39-51). */
    for (int i = 0; i < n; i++) {
        double s[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < n; j++) {
            for (int k = 0; k < 3; k++) d[k] = x[j][k] - x[i][k];
            double r2 = dot3(d, d) + eps2;
            double rinv = ref_rsqrt(r2);
            double mr3 = (m[j] * rinv) * rinv * rinv;
            for (int k = 0; k < 3; k++) s[k] += mr3 * d[k];
        }
        for (int k = 0; k < 3; k++) a[i][k] = s[k];
    }

    /* a1: BH <- star 1PN, unsoftened /r2 (:54-97). */
    acc1[0] = acc1[1] = acc1[2] = 0.0;
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) d[k] = x[i][k] - x[0][k];
        double r2 = dot3(d, d);
        double re = ref_rsqrt(r2 + eps2);
        double mr3 = (m[i] * re) * re * re;
        double vv = dot3(v[i], v[i]);
        double vd = dot3(v[i], d);
        /* ((5 m0) re - 2 vv) + ((1.5 vd) vd)/r2 */
        double br = (5.0 * m[0] * re - 2.0 * vv) + 1.5 * vd * vd / r2;
        for (int k = 0; k < 3; k++)
            acc1[k] += mr3 * d[k] * br + 3.0 * mr3 * v[i][k] * vd;
    }
    for (int k = 0; k < 3; k++) a[0][k] += lam * acc1[k];

    /* a1c prefix 4 m_i^2 e^4 d (:101-115). */
    a1c[0] = a1c[1] = a1c[2] = 0.0;
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) db[k] = x[i][k] - x[0][k];
        double re = ref_rsqrt(dot3(db, db) + eps2);
        double mr3 = (m[i] * re) * re * re;
        for (int k = 0; k < 3; k++) a1c[k] += 4.0 * mr3 * re * m[i] * db[k];
    }

    /* a1c three-body cross double sum, j = i included (:118-174). */
    for (int i = 1; i < n; i++) {
        for (int j = 1; j < n; j++) {
            for (int k = 0; k < 3; k++) {
                db[k] = x[i][k] - x[0][k];
                dc[k] = x[j][k] - x[0][k];
                dbc[k] = x[j][k] - x[i][k];
            }
            double rb2 = dot3(db, db);
            double rbe = ref_rsqrt(rb2 + eps2);
            double rc2 = dot3(dc, dc);
            double rce = ref_rsqrt(rc2 + eps2);
            double rbc2e = dot3(dbc, dbc) + eps2;
            double rbce = ref_rsqrt(rbc2e);
            double mr3 = (m[i] * rbe) * rbe * rbe;
            double vivj = dot3(v[i], v[j]);
            double vidb = dot3(v[i], db);
            double vjdb = dot3(v[j], db);
            /* ((4 rce + 1.25 rbce) - ((0.25 rc2)/rbc2e) rbce) + ((0.25 rb2)/rbc2e) rbce */
            double br = 4.0 * rce + 1.25 * rbce - 0.25 * rc2 / rbc2e * rbce
                        + 0.25 * rb2 / rbc2e * rbce;
            /* ((((3.5 (rbce rbce rbce)) rbe) m_i) m_j) */
            double cn = 3.5 * (rbce * rbce * rbce) * rbe * m[i] * m[j];
            double cv = mr3 * m[j] / m[0];
            for (int k = 0; k < 3; k++)
                a1c[k] += mr3 * m[j] * db[k] * br - cn * dbc[k]
                          - cv * (4.0 * vivj * db[k] - 3.0 * vidb * v[j][k]
                                  - 4.0 * vjdb * v[i][k]);
        }
    }

    /* a1c "remove double-counting", as written (:179-207). */
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) db[k] = x[0][k] - x[i][k];
        double re = ref_rsqrt(dot3(db, db) + eps2);
        double mr3 = (m[i] * re) * re * re;
        double vv = dot3(v[i], v[i]);
        double vd = dot3(v[i], db);
        for (int k = 0; k < 3; k++)
            a1c[k] -= mr3 * m[i] * db[k] * (4.0 * re + 1.25 * eps)
                      - mr3 * m[i] / m[0] * (4.0 * vv * db[k] - 7.0 * vd * v[i][k]);
    }
    for (int k = 0; k < 3; k++) a[0][k] += lam * a1c[k];

    /* star <- BH 1PN, b = x0 - x_i (:215-238). */
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) db[k] = x[0][k] - x[i][k];
        double re = ref_rsqrt(dot3(db, db) + eps2);
        double mri = m[0] * re;
        double mr3 = mri * re * re;
        double vv = dot3(v[i], v[i]);
        double vd = dot3(v[i], db);
        for (int k = 0; k < 3; k++)
            a[i][k] += lam * (mr3 * db[k] * (4.0 * mri - vv) + 4.0 * mr3 * vd * v[i][k]);
    }

    /* "aa_cross" single-star term with d = x_i - x0 (:249-274). */
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) db[k] = x[i][k] - x[0][k];
        double re = ref_rsqrt(dot3(db, db) + eps2);
        double mr3 = m[0] * re * re * re;
        double mie3 = m[i] * re * re * re;
        double vv = dot3(v[i], v[i]);
        double vd = dot3(v[i], db);
        for (int k = 0; k < 3; k++)
            a[i][k] += lam * (5.0 * mr3 * m[i] * re * db[k]
                              - mie3 * (4.0 * vv * db[k] - 7.0 * vd * v[i][k]));
    }

    /* ac three-body cross double sum, j = i included (:278-382). */
    for (int i = 1; i < n; i++) {
        double s[3] = {0.0, 0.0, 0.0};
        for (int j = 1; j < n; j++) {
            for (int k = 0; k < 3; k++) {
                db[k] = x[0][k] - x[i][k];
                dc[k] = x[0][k] - x[j][k];
                dbc[k] = x[j][k] - x[i][k];
                dvbc[k] = v[j][k] - v[i][k];
            }
            double rb2 = dot3(db, db);
            double rbe = ref_rsqrt(rb2 + eps2);
            double rc2 = dot3(dc, dc);
            double rce = ref_rsqrt(rc2 + eps2);
            double rbc2 = dot3(dbc, dbc);
            double rbc2e = rbc2 + eps2;
            double rbce = ref_rsqrt(rbc2e);
            double mr3 = (m[j] * rbce) * rbce * rbce;
            /* (((((m_j rbe) rbe) rbe) m0) */
            double c1 = m[j] * rbe * rbe * rbe * m[0];
            double b1 = 4.0 * rbce + 1.25 * rce + 0.25 * (rb2 - rbc2) * rce * rce * rce;
            double c2 = mr3 * m[0];
            double b2 = 4.0 * rbe + 1.25 * rce + 0.25 * (-rb2 + rbc2) * rce * rce * rce;
            double c3 = 3.5 * m[j] * m[0] * rce * rce * rce * (rbce - rbe);
            double c4 = m[j] * rbe * rbe * rbe;
            double vivj = dot3(v[i], v[j]);
            double vjdb = dot3(v[j], db);
            double vidb = dot3(v[i], db);
            double vvi = dot3(v[i], v[i]);
            double dvdv = dot3(dvbc, dvbc);
            double vjr = dot3(v[j], dbc);
            double b5 = vvi - 2.0 * dvdv + 1.5 * vjr * vjr / rbc2e;
            /* ((r0 (4vi0 - 3vj0) + r1 (4vi1 - 3vj1)) + r2 (4vi2 - 3vj2)) */
            double b6 = dbc[0] * (4.0 * v[i][0] - 3.0 * v[j][0])
                        + dbc[1] * (4.0 * v[i][1] - 3.0 * v[j][1])
                        + dbc[2] * (4.0 * v[i][2] - 3.0 * v[j][2]);
            for (int k = 0; k < 3; k++)
                s[k] += c1 * b1 * db[k] + c2 * b2 * dbc[k] - c3 * dc[k]
                        - c4 * (4.0 * vivj * db[k] - 3.0 * vjdb * v[i][k]
                                - 4.0 * vidb * v[j][k])
                        + mr3 * b5 * dbc[k] + mr3 * b6 * dvbc[k];
        }
        for (int k = 0; k < 3; k++) ac[i][k] = s[k];
    }

    /* add ac, then the exact self-term removal (:386-422). */
    for (int i = 1; i < n; i++)
        for (int k = 0; k < 3; k++) a[i][k] += lam * ac[i][k];
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) db[k] = x[0][k] - x[i][k];
        double rb2 = dot3(db, db);
        double rbe = ref_rsqrt(rb2 + eps2);
        double mr3 = (m[i] * rbe) * rbe * rbe;
        double vv = dot3(v[i], v[i]);
        double vd = dot3(v[i], db);
        double br = 4.0 / eps + 1.25 * rbe + 0.25 * rb2 * rbe * rbe * rbe;
        for (int k = 0; k < 3; k++)
            a[i][k] -= lam * (mr3 * m[0] * db[k] * br
                              - 3.5 * mr3 * m[0] * (1.0 / eps - rbe) * db[k]
                              - mr3 * (4.0 * vv * db[k] - 7.0 * vd * v[i][k]));
    }
    free(ac);
}

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

/* ------------------------------------------------------------------ */
/* RK stage driver  Limit-sticky9.c:841-893                            */
/* Three force evaluations; the state combination                     */
/* x <- ((Y0 + Y3) + 2 (Y1 + Y2)) / 6 of the stage states.             */
/* ------------------------------------------------------------------ */
void orc_q1_steps(double *x, double *v, const double *m, int n,
                  double eps, double c, double dt, int steps)
{
    size_t n3 = (size_t)n * 3;
    double *a = xcalloc(n3, sizeof *a);
    double *x1 = xcalloc(n3, sizeof *x1), *v1 = xcalloc(n3, sizeof *v1);
    double *hx = xcalloc(4 * n3, sizeof *hx), *hv = xcalloc(4 * n3, sizeof *hv);

    for (int s = 0; s < steps; s++) {
        memcpy(hx, x, n3 * sizeof *x);
        memcpy(hv, v, n3 * sizeof *v);

        orc_force(x, v, m, eps, c, a, n);
        for (size_t q = 0; q < n3; q++) {
            x1[q] = x[q] + 0.5 * v[q] * dt;
            v1[q] = v[q] + 0.5 * a[q] * dt;
            hx[n3 + q] = x1[q];
            hv[n3 + q] = v1[q];
        }
        orc_force(x1, v1, m, eps, c, a, n);
        for (size_t q = 0; q < n3; q++) {
            x1[q] = x[q] + 0.5 * v1[q] * dt; /* v1 of the previous stage */
            v1[q] = v[q] + 0.5 * a[q] * dt;
            hx[2 * n3 + q] = x1[q];
            hv[2 * n3 + q] = v1[q];
        }
        orc_force(x1, v1, m, eps, c, a, n);
        for (size_t q = 0; q < n3; q++) {
            x1[q] = x[q] + v1[q] * dt;
            v1[q] = v[q] + a[q] * dt;
            hx[3 * n3 + q] = x1[q];
            hv[3 * n3 + q] = v1[q];
        }
        for (size_t q = 0; q < n3; q++) {
            x[q] = (hx[q] + hx[3 * n3 + q] + 2.0 * (hx[n3 + q] + hx[2 * n3 + q])) / 6.0;
            v[q] = (hv[q] + hv[3 * n3 + q] + 2.0 * (hv[n3 + q] + hv[2 * n3 + q])) / 6.0;
        }
    }
    free(a); free(x1); free(v1); free(hx); free(hv);
}

/* One force evaluation, timed by the caller: a plain wrapper so a CPU
 * baseline can time the oracle without the RK driver around it. */
void orc_force_only(const double *x, const double *v, const double *m,
                    double eps, double c, double *a, int n)
{
    orc_force(x, v, m, eps, c, a, n);
}
