/*
 * pn1_host.c -- host-only parts of libpn1: input validation, seeded
 * initial conditions, and the file formats on either side of the path.
 *
 * Formats follow Limit-sticky9.c:
 *   input  (:795-809)  n-1 lines "x y z" then n-1 lines "vx vy vz", read
 *                      with "%f %f %f \n" into float; BH row absent;
 *   output (:820-826)  n lines "%lf %lf %lf \n" of x then n of v, BH first.
 * The reference ignores fscanf's result (a short file silently repeats the
 * last value read); here a short or malformed file is an error.
 */
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pn1.h"

/* ------------------------------------------------------------------ */
/* validation                                                          */
/* ------------------------------------------------------------------ */
static void setmsg(char *msg, int len, const char *fmt, ...)
{
    if (!msg || len <= 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, (size_t)len, fmt, ap);
    va_end(ap);
}

int pn1_check_input(const double *x, const double *v, const double *m,
                    int n, double eps, char *msg, int msglen)
{
    if (msg && msglen > 0) msg[0] = '\0';
    if (n < 1) {
        setmsg(msg, msglen, "n = %d: need at least the black hole (n >= 1)", n);
        return PN1_EINVAL;
    }
    if (!x || !v || !m) {
        setmsg(msg, msglen, "null array pointer (n = %d)", n);
        return PN1_EINVAL;
    }
    if (!(eps > 0.0) || !isfinite(eps)) {
        setmsg(msg, msglen, "eps must be finite and > 0 (eps = %g)", eps);
        return PN1_EINVAL;
    }
    /* the 1PN terms divide by m0 (Limit-sticky9.c:142, 192) */
    if (!(m[0] > 0.0) || !isfinite(m[0])) {
        setmsg(msg, msglen, "black hole mass m[0] must be finite and > 0 (m0 = %g)", m[0]);
        return PN1_EINVAL;
    }
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            if (!isfinite(x[3 * i + k]) || !isfinite(v[3 * i + k])) {
                setmsg(msg, msglen, "non-finite position or velocity at particle %d", i);
                return PN1_EINVAL;
            }
        }
        if (!isfinite(m[i])) {
            setmsg(msg, msglen, "non-finite mass at particle %d (%g)", i, m[i]);
            return PN1_EINVAL;
        }
    }
    /* a1 divides by the unsoftened star-BH distance (Limit-sticky9.c:74) */
    for (int i = 1; i < n; i++) {
        if (x[3 * i] == x[0] && x[3 * i + 1] == x[1] && x[3 * i + 2] == x[2]) {
            setmsg(msg, msglen, "star %d sits exactly on the black hole (unsoftened 1/r^2 term)", i);
            return PN1_EINVAL;
        }
    }
    return PN1_OK;
}

/* ------------------------------------------------------------------ */
/* SplitMix64 (Steele, Lea & Flood, OOPSLA 2014) -> uniforms, normals  */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t s;
    int have_spare;
    double spare;
} rng_t;

static uint64_t sm64_next(rng_t *r)
{
    uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform in [0, 1) with 53 random bits */
static double u01(rng_t *r) { return (double)(sm64_next(r) >> 11) * 0x1.0p-53; }

/* standard normal, Box-Muller (both values of each pair are used) */
static double gauss(rng_t *r)
{
    if (r->have_spare) {
        r->have_spare = 0;
        return r->spare;
    }
    double u1 = 1.0 - u01(r); /* (0, 1] */
    double u2 = u01(r);
    double rad = sqrt(-2.0 * log(u1));
    double ang = 2.0 * M_PI * u2;
    r->spare = rad * sin(ang);
    r->have_spare = 1;
    return rad * cos(ang);
}

static void iso_dir(rng_t *r, double len, double *out)
{
    double ct = 2.0 * u01(r) - 1.0;
    double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    double ph = 2.0 * M_PI * u01(r);
    out[0] = len * st * cos(ph);
    out[1] = len * st * sin(ph);
    out[2] = len * ct;
}

static double f32(double a) { return (double)(float)a; }

static void remove_mean(double *q, int n_stars)
{
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = 1; i <= n_stars; i++)
        for (int k = 0; k < 3; k++) s[k] += q[3 * i + k];
    for (int k = 0; k < 3; k++) s[k] /= (double)n_stars;
    for (int i = 1; i <= n_stars; i++)
        for (int k = 0; k < 3; k++) q[3 * i + k] -= s[k];
}

static void set_bh(double *x, double *v, double *m, int n_stars, double m_bh)
{
    for (int k = 0; k < 3; k++) x[k] = v[k] = 0.0;
    m[0] = m_bh;
    for (int i = 1; i <= n_stars; i++) m[i] = 1.0 / (double)n_stars;
}

int pn1_ic_uniform_sphere(int n_stars, unsigned long long seed, double R,
                          double Q, double m_bh, double eps,
                          double *x, double *v, double *m)
{
    if (n_stars < 1 || !(R > 0.0) || !(Q > 0.0) || !(m_bh > 0.0) || !(eps > 0.0) || !x || !v || !m)
        return PN1_EINVAL;
    rng_t r = {seed, 0, 0.0};
    set_bh(x, v, m, n_stars, m_bh);
    for (int i = 1; i <= n_stars; i++) iso_dir(&r, R * cbrt(u01(&r)), x + 3 * i);
    for (int i = 1; i <= n_stars; i++)
        for (int k = 0; k < 3; k++) v[3 * i + k] = gauss(&r);
    remove_mean(x, n_stars);
    remove_mean(v, n_stars);
    for (int i = 1; i <= n_stars; i++)
        for (int k = 0; k < 3; k++) x[3 * i + k] = f32(x[3 * i + k]);

    /* virial scaling: K / |W| = Q with W the softened potential energy of
     * all pairs, the BH included */
    const double mi = 1.0 / (double)n_stars, eps2 = eps * eps;
    double W = 0.0, K = 0.0;
    for (int i = 1; i <= n_stars; i++) {
        const double *xi = x + 3 * i;
        double wi = 0.0;
        for (int j = i + 1; j <= n_stars; j++) {
            const double *xj = x + 3 * j;
            double dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
            wi += 1.0 / sqrt(dx * dx + dy * dy + dz * dz + eps2);
        }
        double rb2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
        W -= mi * mi * wi + m_bh * mi / sqrt(rb2 + eps2);
        const double *vi = v + 3 * i;
        K += 0.5 * mi * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
    }
    /* K = 0 only for a single star, whose velocity is zero once the mean is
     * removed (and which then sits on the BH: pn1_check_input rejects it) */
    const double scale = K > 0.0 ? sqrt(Q * fabs(W) / K) : 0.0;
    for (int i = 1; i <= n_stars; i++)
        for (int k = 0; k < 3; k++) v[3 * i + k] = f32(v[3 * i + k] * scale);
    return PN1_OK;
}

int pn1_ic_plummer(int n_stars, unsigned long long seed, double a,
                   double m_bh, double *x, double *v, double *m)
{
    if (n_stars < 1 || !(a > 0.0) || !(m_bh > 0.0) || !x || !v || !m) return PN1_EINVAL;
    rng_t r = {seed, 0, 0.0};
    set_bh(x, v, m, n_stars, m_bh);
    const double vscale = sqrt(1.0 / a); /* G = M = 1 */
    for (int i = 1; i <= n_stars; i++) {
        /* radius from the cumulative mass fraction X (cut at 0.999) */
        double X;
        do { X = 0.999 * u01(&r); } while (X <= 0.0);
        double rad = a / sqrt(pow(X, -2.0 / 3.0) - 1.0);
        iso_dir(&r, rad, x + 3 * i);
        /* speed q v_esc with q from g(q) = q^2 (1 - q^2)^(7/2) by rejection */
        double q, y;
        do {
            q = u01(&r);
            y = 0.1 * u01(&r);
        } while (y >= q * q * pow(1.0 - q * q, 3.5));
        double vesc = M_SQRT2 * vscale * pow(1.0 + rad * rad / (a * a), -0.25);
        iso_dir(&r, q * vesc, v + 3 * i);
    }
    remove_mean(x, n_stars);
    remove_mean(v, n_stars);
    for (int i = 1; i <= n_stars; i++)
        for (int k = 0; k < 3; k++) {
            x[3 * i + k] = f32(x[3 * i + k]);
            v[3 * i + k] = f32(v[3 * i + k]);
        }
    return PN1_OK;
}

/* ------------------------------------------------------------------ */
/* .dat input / snapshot output                                        */
/* ------------------------------------------------------------------ */
int pn1_read_dat(const char *path, int n, double *x, double *v)
{
    if (!path || n < 1 || !x || !v) return PN1_EINVAL;
    FILE *fp = fopen(path, "r");
    if (!fp) return PN1_EIO;
    for (int k = 0; k < 3; k++) x[k] = v[k] = 0.0;
    int rc = PN1_OK;
    for (int pass = 0; pass < 2 && rc == PN1_OK; pass++) {
        double *q = pass == 0 ? x : v;
        for (int i = 1; i < n; i++) {
            float a, b, c;
            if (fscanf(fp, "%f %f %f \n", &a, &b, &c) != 3) {
                rc = PN1_EIO;
                break;
            }
            q[3 * i] = a;
            q[3 * i + 1] = b;
            q[3 * i + 2] = c;
        }
    }
    fclose(fp);
    return rc;
}

int pn1_write_snapshot(const char *path, int append, const double *x,
                       const double *v, int n)
{
    if (!path || n < 1 || !x || !v) return PN1_EINVAL;
    FILE *fp = fopen(path, append ? "a" : "w");
    if (!fp) return PN1_EIO;
    for (int i = 0; i < n; i++) fprintf(fp, "%lf %lf %lf \n", x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    for (int i = 0; i < n; i++) fprintf(fp, "%lf %lf %lf \n", v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    int bad = ferror(fp);
    if (fclose(fp) != 0) bad = 1;
    return bad ? PN1_EIO : PN1_OK;
}

/* ------------------------------------------------------------------ */
/* exact binary checkpoint                                             */
/* ------------------------------------------------------------------ */
static const char CKPT_MAGIC[8] = {'P', 'N', '1', 'C', 'K', 'P', 'T', '1'};

typedef struct {
    char magic[8];
    int32_t n, pad;
    double t;
    int64_t step;
    double eps, c, dt;
} ckpt_hdr_t;

int pn1_write_checkpoint(const char *path, int n, double t, long long step,
                         double eps, double c, double dt, const double *x,
                         const double *v, const double *m)
{
    if (!path || n < 1 || !x || !v || !m) return PN1_EINVAL;
    FILE *fp = fopen(path, "wb");
    if (!fp) return PN1_EIO;
    ckpt_hdr_t h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, CKPT_MAGIC, 8);
    h.n = n;
    h.t = t;
    h.step = step;
    h.eps = eps;
    h.c = c;
    h.dt = dt;
    size_t n3 = (size_t)n * 3;
    int ok = fwrite(&h, sizeof h, 1, fp) == 1 && fwrite(x, sizeof *x, n3, fp) == n3
             && fwrite(v, sizeof *v, n3, fp) == n3 && fwrite(m, sizeof *m, (size_t)n, fp) == (size_t)n;
    if (fclose(fp) != 0) ok = 0;
    return ok ? PN1_OK : PN1_EIO;
}

int pn1_read_checkpoint(const char *path, int *n, double *t, long long *step,
                        double *eps, double *c, double *dt, double *x,
                        double *v, double *m, int capacity)
{
    if (!path || !n) return PN1_EINVAL;
    FILE *fp = fopen(path, "rb");
    if (!fp) return PN1_EIO;
    ckpt_hdr_t h;
    int rc = PN1_OK;
    if (fread(&h, sizeof h, 1, fp) != 1 || memcmp(h.magic, CKPT_MAGIC, 8) != 0 || h.n < 1) {
        rc = PN1_EIO;
    } else {
        *n = h.n;
        if (t) *t = h.t;
        if (step) *step = h.step;
        if (eps) *eps = h.eps;
        if (c) *c = h.c;
        if (dt) *dt = h.dt;
        if (x || v || m) {
            size_t n3 = (size_t)h.n * 3;
            if (!x || !v || !m || h.n > capacity) rc = PN1_EINVAL;
            else if (fread(x, sizeof *x, n3, fp) != n3 || fread(v, sizeof *v, n3, fp) != n3
                     || fread(m, sizeof *m, (size_t)h.n, fp) != (size_t)h.n)
                rc = PN1_EIO;
        }
    }
    fclose(fp);
    return rc;
}
