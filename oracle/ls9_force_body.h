/*
 * ls9_force_body.h -- the oracle's force() and stage driver, written once
 * over a number type.  TEST INFRASTRUCTURE ONLY.
 *
 * Included by ls9_oracle.c with REAL = double (the oracle proper, pinned
 * bit for bit to the reference) and by ls9_oracle_cs.c with REAL = double
 * complex (complex-step derivatives: f(x + ih) = f(x) + ih f'(x) + O(h^2),
 * so Im f(x + ih) / h is the derivative of this same code to rounding for
 * h ~ 1e-30; every operation here is analytic).  The includer defines
 *   REAL       the number type
 *   SQRT(x)    its square root
 *   FN(name)   the exported name of each function
 * and provides xcalloc().
 */
/* ((a0*b0) + (a1*b1)) + (a2*b2): the tree of every 3-term dot product
 * written out in Limit-sticky9.c (e.g. line 43 without the +eps2). */
static inline REAL FN(dot3)(const REAL *a, const REAL *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* Limit-sticky9.c:15-19 */
static inline REAL FN(ref_rsqrt)(REAL r2) { return 1.0 / SQRT(r2); }

/* ------------------------------------------------------------------ */
/* force()  Limit-sticky9.c:21-424                                     */
/* ------------------------------------------------------------------ */
/* Per-star pieces of force(), shared by orc_force and orc_force_rows so
 * that a row evaluated alone is bitwise that row of the full call.  Each
 * *_term writes the bracket that the reference scales by lam and adds to
 * (or subtracts from) a[i]; the caller applies lam, as the reference's
 * a[i][k] += lam * (...) does. */

/* Newtonian row i, j in [0,n) including self (:39-51). */
static inline void FN(row_newton)(const REAL (*x)[3], const REAL *m, double eps2,
                                  int i, int n, REAL s[3])
{
    REAL d[3];
    s[0] = s[1] = s[2] = 0.0;
    for (int j = 0; j < n; j++) {
        for (int k = 0; k < 3; k++) d[k] = x[j][k] - x[i][k];
        REAL r2 = FN(dot3)(d, d) + eps2;
        REAL rinv = FN(ref_rsqrt)(r2);
        REAL mr3 = (m[j] * rinv) * rinv * rinv;
        for (int k = 0; k < 3; k++) s[k] += mr3 * d[k];
    }
}

/* star <- BH 1PN, b = x0 - x_i (:215-238). */
static inline void FN(row_star_bh_term)(const REAL (*x)[3], const REAL (*v)[3],
                                        const REAL *m, double eps2, int i, REAL t[3])
{
    REAL db[3];
    for (int k = 0; k < 3; k++) db[k] = x[0][k] - x[i][k];
    REAL re = FN(ref_rsqrt)(FN(dot3)(db, db) + eps2);
    REAL mri = m[0] * re;
    REAL mr3 = mri * re * re;
    REAL vv = FN(dot3)(v[i], v[i]);
    REAL vd = FN(dot3)(v[i], db);
    for (int k = 0; k < 3; k++)
        t[k] = mr3 * db[k] * (4.0 * mri - vv) + 4.0 * mr3 * vd * v[i][k];
}

/* "aa_cross" single-star term with d = x_i - x0 (:249-274). */
static inline void FN(row_aa_cross_term)(const REAL (*x)[3], const REAL (*v)[3],
                                         const REAL *m, double eps2, int i, REAL t[3])
{
    REAL db[3];
    for (int k = 0; k < 3; k++) db[k] = x[i][k] - x[0][k];
    REAL re = FN(ref_rsqrt)(FN(dot3)(db, db) + eps2);
    REAL mr3 = m[0] * re * re * re;
    REAL mie3 = m[i] * re * re * re;
    REAL vv = FN(dot3)(v[i], v[i]);
    REAL vd = FN(dot3)(v[i], db);
    for (int k = 0; k < 3; k++)
        t[k] = 5.0 * mr3 * m[i] * re * db[k]
               - mie3 * (4.0 * vv * db[k] - 7.0 * vd * v[i][k]);
}

/* ac three-body cross sum of star i, j = i included (:278-382). */
static inline void FN(row_ac)(const REAL (*x)[3], const REAL (*v)[3], const REAL *m,
                              double eps2, int i, int n, REAL s[3])
{
    REAL db[3], dc[3], dbc[3], dvbc[3];
    s[0] = s[1] = s[2] = 0.0;
    for (int j = 1; j < n; j++) {
        for (int k = 0; k < 3; k++) {
            db[k] = x[0][k] - x[i][k];
            dc[k] = x[0][k] - x[j][k];
            dbc[k] = x[j][k] - x[i][k];
            dvbc[k] = v[j][k] - v[i][k];
        }
        REAL rb2 = FN(dot3)(db, db);
        REAL rbe = FN(ref_rsqrt)(rb2 + eps2);
        REAL rc2 = FN(dot3)(dc, dc);
        REAL rce = FN(ref_rsqrt)(rc2 + eps2);
        REAL rbc2 = FN(dot3)(dbc, dbc);
        REAL rbc2e = rbc2 + eps2;
        REAL rbce = FN(ref_rsqrt)(rbc2e);
        REAL mr3 = (m[j] * rbce) * rbce * rbce;
        /* (((((m_j rbe) rbe) rbe) m0) */
        REAL c1 = m[j] * rbe * rbe * rbe * m[0];
        REAL b1 = 4.0 * rbce + 1.25 * rce + 0.25 * (rb2 - rbc2) * rce * rce * rce;
        REAL c2 = mr3 * m[0];
        REAL b2 = 4.0 * rbe + 1.25 * rce + 0.25 * (-rb2 + rbc2) * rce * rce * rce;
        REAL c3 = 3.5 * m[j] * m[0] * rce * rce * rce * (rbce - rbe);
        REAL c4 = m[j] * rbe * rbe * rbe;
        REAL vivj = FN(dot3)(v[i], v[j]);
        REAL vjdb = FN(dot3)(v[j], db);
        REAL vidb = FN(dot3)(v[i], db);
        REAL vvi = FN(dot3)(v[i], v[i]);
        REAL dvdv = FN(dot3)(dvbc, dvbc);
        REAL vjr = FN(dot3)(v[j], dbc);
        REAL b5 = vvi - 2.0 * dvdv + 1.5 * vjr * vjr / rbc2e;
        /* ((r0 (4vi0 - 3vj0) + r1 (4vi1 - 3vj1)) + r2 (4vi2 - 3vj2)) */
        REAL b6 = dbc[0] * (4.0 * v[i][0] - 3.0 * v[j][0])
                    + dbc[1] * (4.0 * v[i][1] - 3.0 * v[j][1])
                    + dbc[2] * (4.0 * v[i][2] - 3.0 * v[j][2]);
        for (int k = 0; k < 3; k++)
            s[k] += c1 * b1 * db[k] + c2 * b2 * dbc[k] - c3 * dc[k]
                    - c4 * (4.0 * vivj * db[k] - 3.0 * vjdb * v[i][k]
                            - 4.0 * vidb * v[j][k])
                    + mr3 * b5 * dbc[k] + mr3 * b6 * dvbc[k];
    }
}

/* the exact self-term removal of star i (:396-422), subtracted. */
static inline void FN(row_self_term)(const REAL (*x)[3], const REAL (*v)[3],
                                     const REAL *m, double eps, double eps2, int i,
                                     REAL t[3])
{
    REAL db[3];
    for (int k = 0; k < 3; k++) db[k] = x[0][k] - x[i][k];
    REAL rb2 = FN(dot3)(db, db);
    REAL rbe = FN(ref_rsqrt)(rb2 + eps2);
    REAL mr3 = (m[i] * rbe) * rbe * rbe;
    REAL vv = FN(dot3)(v[i], v[i]);
    REAL vd = FN(dot3)(v[i], db);
    REAL br = 4.0 / eps + 1.25 * rbe + 0.25 * rb2 * rbe * rbe * rbe;
    for (int k = 0; k < 3; k++)
        t[k] = mr3 * m[0] * db[k] * br
               - 3.5 * mr3 * m[0] * (1.0 / eps - rbe) * db[k]
               - mr3 * (4.0 * vv * db[k] - 7.0 * vd * v[i][k]);
}

/* The BH's 1PN sums.  bh_a1c_cross_row adds the j loop of one i to the
 * caller's accumulator, so a serial caller looping i = 1 .. n-1 keeps the
 * reference's single-accumulator order (i outer, j inner). */

/* a1: BH <- star 1PN, unsoftened /r2 (:54-97). */
static inline void FN(bh_a1)(const REAL (*x)[3], const REAL (*v)[3], const REAL *m,
                             double eps2, int n, REAL acc1[3])
{
    REAL d[3];
    acc1[0] = acc1[1] = acc1[2] = 0.0;
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) d[k] = x[i][k] - x[0][k];
        REAL r2 = FN(dot3)(d, d);
        REAL re = FN(ref_rsqrt)(r2 + eps2);
        REAL mr3 = (m[i] * re) * re * re;
        REAL vv = FN(dot3)(v[i], v[i]);
        REAL vd = FN(dot3)(v[i], d);
        /* ((5 m0) re - 2 vv) + ((1.5 vd) vd)/r2 */
        REAL br = (5.0 * m[0] * re - 2.0 * vv) + 1.5 * vd * vd / r2;
        for (int k = 0; k < 3; k++)
            acc1[k] += mr3 * d[k] * br + 3.0 * mr3 * v[i][k] * vd;
    }
}

/* a1c prefix 4 m_i^2 e^4 d (:101-115); starts the a1c accumulator. */
static inline void FN(bh_a1c_prefix)(const REAL (*x)[3], const REAL *m, double eps2, int n,
                                     REAL a1c[3])
{
    REAL db[3];
    a1c[0] = a1c[1] = a1c[2] = 0.0;
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) db[k] = x[i][k] - x[0][k];
        REAL re = FN(ref_rsqrt)(FN(dot3)(db, db) + eps2);
        REAL mr3 = (m[i] * re) * re * re;
        for (int k = 0; k < 3; k++) a1c[k] += 4.0 * mr3 * re * m[i] * db[k];
    }
}

/* a1c three-body cross sum, row i, j = i included (:118-174). */
static inline void FN(bh_a1c_cross_row)(const REAL (*x)[3], const REAL (*v)[3],
                                        const REAL *m, double eps2, int i, int n,
                                        REAL a1c[3])
{
    REAL db[3], dc[3], dbc[3];
    for (int j = 1; j < n; j++) {
        for (int k = 0; k < 3; k++) {
            db[k] = x[i][k] - x[0][k];
            dc[k] = x[j][k] - x[0][k];
            dbc[k] = x[j][k] - x[i][k];
        }
        REAL rb2 = FN(dot3)(db, db);
        REAL rbe = FN(ref_rsqrt)(rb2 + eps2);
        REAL rc2 = FN(dot3)(dc, dc);
        REAL rce = FN(ref_rsqrt)(rc2 + eps2);
        REAL rbc2e = FN(dot3)(dbc, dbc) + eps2;
        REAL rbce = FN(ref_rsqrt)(rbc2e);
        REAL mr3 = (m[i] * rbe) * rbe * rbe;
        REAL vivj = FN(dot3)(v[i], v[j]);
        REAL vidb = FN(dot3)(v[i], db);
        REAL vjdb = FN(dot3)(v[j], db);
        /* ((4 rce + 1.25 rbce) - ((0.25 rc2)/rbc2e) rbce) + ((0.25 rb2)/rbc2e) rbce */
        REAL br = 4.0 * rce + 1.25 * rbce - 0.25 * rc2 / rbc2e * rbce
                    + 0.25 * rb2 / rbc2e * rbce;
        /* ((((3.5 (rbce rbce rbce)) rbe) m_i) m_j) */
        REAL cn = 3.5 * (rbce * rbce * rbce) * rbe * m[i] * m[j];
        REAL cv = mr3 * m[j] / m[0];
        for (int k = 0; k < 3; k++)
            a1c[k] += mr3 * m[j] * db[k] * br - cn * dbc[k]
                      - cv * (4.0 * vivj * db[k] - 3.0 * vidb * v[j][k]
                              - 4.0 * vjdb * v[i][k]);
    }
}

/* a1c "remove REAL-counting", as written (:179-207). */
static inline void FN(bh_a1c_removal)(const REAL (*x)[3], const REAL (*v)[3], const REAL *m,
                                      double eps, double eps2, int n, REAL a1c[3])
{
    REAL db[3];
    for (int i = 1; i < n; i++) {
        for (int k = 0; k < 3; k++) db[k] = x[0][k] - x[i][k];
        REAL re = FN(ref_rsqrt)(FN(dot3)(db, db) + eps2);
        REAL mr3 = (m[i] * re) * re * re;
        REAL vv = FN(dot3)(v[i], v[i]);
        REAL vd = FN(dot3)(v[i], db);
        for (int k = 0; k < 3; k++)
            a1c[k] -= mr3 * m[i] * db[k] * (4.0 * re + 1.25 * eps)
                      - mr3 * m[i] / m[0] * (4.0 * vv * db[k] - 7.0 * vd * v[i][k]);
    }
}

void FN(orc_force)(const REAL *xs, const REAL *vs, const REAL *m,
                   double eps, double c, REAL *as, int n)
{
    const REAL (*x)[3] = (const REAL (*)[3])xs;
    const REAL (*v)[3] = (const REAL (*)[3])vs;
    REAL (*a)[3] = (REAL (*)[3])as;
    const double lam = 1.0 / (c * c);
    const double eps2 = eps * eps;
    REAL (*ac)[3] = xcalloc((size_t)n, sizeof *ac);
    REAL t[3], acc1[3], a1c[3];

    /* Newtonian, all i, j in [0,n) including self (:39-51). */
    for (int i = 0; i < n; i++) FN(row_newton)(x, m, eps2, i, n, a[i]);

    FN(bh_a1)(x, v, m, eps2, n, acc1);
    for (int k = 0; k < 3; k++) a[0][k] += lam * acc1[k];
    FN(bh_a1c_prefix)(x, m, eps2, n, a1c);
    for (int i = 1; i < n; i++) FN(bh_a1c_cross_row)(x, v, m, eps2, i, n, a1c);
    FN(bh_a1c_removal)(x, v, m, eps, eps2, n, a1c);
    for (int k = 0; k < 3; k++) a[0][k] += lam * a1c[k];

    for (int i = 1; i < n; i++) {
        FN(row_star_bh_term)(x, v, m, eps2, i, t);
        for (int k = 0; k < 3; k++) a[i][k] += lam * t[k];
    }
    for (int i = 1; i < n; i++) {
        FN(row_aa_cross_term)(x, v, m, eps2, i, t);
        for (int k = 0; k < 3; k++) a[i][k] += lam * t[k];
    }
    for (int i = 1; i < n; i++) FN(row_ac)(x, v, m, eps2, i, n, ac[i]);

    /* add ac, then the exact self-term removal (:386-422). */
    for (int i = 1; i < n; i++)
        for (int k = 0; k < 3; k++) a[i][k] += lam * ac[i][k];
    for (int i = 1; i < n; i++) {
        FN(row_self_term)(x, v, m, eps, eps2, i, t);
        for (int k = 0; k < 3; k++) a[i][k] -= lam * t[k];
    }
    free(ac);
}

/* Rows rows[0..nrows) (stars, 1 <= rows[r] < n) of force(): out[r] is
 * bitwise a[rows[r]] of orc_force on the same input.  A star's row is
 * O(n); the BH's row needs every pair and is not offered here. */
void FN(orc_force_rows)(const REAL *xs, const REAL *vs, const REAL *m,
                        double eps, double c, const int *rows, int nrows,
                        REAL *outs, int n)
{
    const REAL (*x)[3] = (const REAL (*)[3])xs;
    const REAL (*v)[3] = (const REAL (*)[3])vs;
    REAL (*out)[3] = (REAL (*)[3])outs;
    const double lam = 1.0 / (c * c);
    const double eps2 = eps * eps;
    REAL a[3], t[3], s[3];

    for (int r = 0; r < nrows; r++) {
        int i = rows[r];
        if (i < 1 || i >= n) {
            fprintf(stderr, "ls9_oracle: row %d outside the stars [1, %d)\n", i, n);
            exit(1);
        }
        FN(row_newton)(x, m, eps2, i, n, a);
        FN(row_star_bh_term)(x, v, m, eps2, i, t);
        for (int k = 0; k < 3; k++) a[k] += lam * t[k];
        FN(row_aa_cross_term)(x, v, m, eps2, i, t);
        for (int k = 0; k < 3; k++) a[k] += lam * t[k];
        FN(row_ac)(x, v, m, eps2, i, n, s);
        for (int k = 0; k < 3; k++) a[k] += lam * s[k];
        FN(row_self_term)(x, v, m, eps, eps2, i, t);
        for (int k = 0; k < 3; k++) a[k] -= lam * t[k];
        for (int k = 0; k < 3; k++) out[r][k] = a[k];
    }
}

/* ------------------------------------------------------------------ */
/* RK stage driver  Limit-sticky9.c:841-893                            */
/* Three force evaluations; the state combination                     */
/* x <- ((Y0 + Y3) + 2 (Y1 + Y2)) / 6 of the stage states.             */
/* ------------------------------------------------------------------ */
void FN(orc_q1_steps)(REAL *x, REAL *v, const REAL *m, int n,
                      double eps, double c, double dt, int steps)
{
    size_t n3 = (size_t)n * 3;
    REAL *a = xcalloc(n3, sizeof *a);
    REAL *x1 = xcalloc(n3, sizeof *x1), *v1 = xcalloc(n3, sizeof *v1);
    REAL *hx = xcalloc(4 * n3, sizeof *hx), *hv = xcalloc(4 * n3, sizeof *hv);

    for (int s = 0; s < steps; s++) {
        memcpy(hx, x, n3 * sizeof *x);
        memcpy(hv, v, n3 * sizeof *v);

        FN(orc_force)(x, v, m, eps, c, a, n);
        for (size_t q = 0; q < n3; q++) {
            x1[q] = x[q] + 0.5 * v[q] * dt;
            v1[q] = v[q] + 0.5 * a[q] * dt;
            hx[n3 + q] = x1[q];
            hv[n3 + q] = v1[q];
        }
        FN(orc_force)(x1, v1, m, eps, c, a, n);
        for (size_t q = 0; q < n3; q++) {
            x1[q] = x[q] + 0.5 * v1[q] * dt; /* v1 of the previous stage */
            v1[q] = v[q] + 0.5 * a[q] * dt;
            hx[2 * n3 + q] = x1[q];
            hv[2 * n3 + q] = v1[q];
        }
        FN(orc_force)(x1, v1, m, eps, c, a, n);
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

