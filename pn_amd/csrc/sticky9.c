/*
 * sticky9.c -- command-line host for libpn1: the driver of Limit-sticky9
 * (Limit-sticky9.c:740-920) with its constants as options.
 *
 *   sticky9 [--in FILE | --ic sphere|plummer --nstars N --seed S]
 *           [--n N] [--m0 M] [--eps E] [--dt DT] [--c C] [--endtime T]
 *           [--deouttime T] [--steps K] [--out FILE] [--mode fp64|mixed]
 *           [--legacy] [--autotune] [--checkpoint FILE] [--resume FILE]
 *
 * Defaults are the reference's: --in Init10k-sphere.dat, n = 10001,
 * m0 = 100/n and star masses 1/(n-1) (Limit-sticky9.c:813-816), eps =
 * 0.001, dt = 2^-17, endtime 10, deouttime 0.005, output
 * Sphere-r10-1PN.dat, c = 1 (the reference has no speed of light).
 * Output: the same stdout lines ("Init", "start", the energy blocks, the
 * final time) and the same snapshot file layout.
 *
 * --legacy runs the reference's own loop on host arrays, calling the
 * drop-in force() three times per step and energy() at each snapshot;
 * without it the state stays on the device (pn1_step between snapshots).
 * --autotune (resident mode) times the pair kernel's launch shapes on the
 * initial state first and keeps the fastest (pn1_autotune).
 * --steps K stops after K steps instead of at endtime.  --checkpoint
 * writes an exact binary state at the end; --resume starts from one.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pn1.h"

typedef struct {
    const char *in, *out, *ic, *ckpt, *resume;
    int n, nstars, mode, legacy, autotune;
    long long steps;
    unsigned long long seed;
    double m0, eps, dt, c, endtime, deouttime;
} opts_t;

static void usage(void)
{
    fprintf(stderr,
            "usage: sticky9 [--in FILE | --ic sphere|plummer --nstars N --seed S] [--n N]\n"
            "               [--m0 M] [--eps E] [--dt DT] [--c C] [--endtime T] [--deouttime T]\n"
            "               [--steps K] [--out FILE] [--mode fp64|mixed] [--legacy]\n"
            "               [--autotune] [--checkpoint FILE] [--resume FILE]\n");
    exit(2);
}

static double need_num(int argc, char **argv, int *i)
{
    if (*i + 1 >= argc) usage();
    char *end;
    double v = strtod(argv[++*i], &end);
    if (*end) usage();
    return v;
}

static void parse(int argc, char **argv, opts_t *o)
{
    memset(o, 0, sizeof *o);
    o->in = "Init10k-sphere.dat";
    o->out = "Sphere-r10-1PN.dat";
    o->n = 10001;
    o->m0 = -1.0;
    o->eps = 0.001;
    o->dt = 1.0 / 131072.0;
    o->c = 1.0;
    o->endtime = 10.0;
    o->deouttime = 0.005;
    o->steps = -1;
    o->seed = 1;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!strcmp(a, "--in") && i + 1 < argc) o->in = argv[++i];
        else if (!strcmp(a, "--out") && i + 1 < argc) o->out = argv[++i];
        else if (!strcmp(a, "--ic") && i + 1 < argc) o->ic = argv[++i];
        else if (!strcmp(a, "--checkpoint") && i + 1 < argc) o->ckpt = argv[++i];
        else if (!strcmp(a, "--resume") && i + 1 < argc) o->resume = argv[++i];
        else if (!strcmp(a, "--n")) o->n = (int)need_num(argc, argv, &i);
        else if (!strcmp(a, "--nstars")) o->nstars = (int)need_num(argc, argv, &i);
        else if (!strcmp(a, "--seed")) o->seed = (unsigned long long)need_num(argc, argv, &i);
        else if (!strcmp(a, "--m0")) o->m0 = need_num(argc, argv, &i);
        else if (!strcmp(a, "--eps")) o->eps = need_num(argc, argv, &i);
        else if (!strcmp(a, "--dt")) o->dt = need_num(argc, argv, &i);
        else if (!strcmp(a, "--c")) o->c = need_num(argc, argv, &i);
        else if (!strcmp(a, "--endtime")) o->endtime = need_num(argc, argv, &i);
        else if (!strcmp(a, "--deouttime")) o->deouttime = need_num(argc, argv, &i);
        else if (!strcmp(a, "--steps")) o->steps = (long long)need_num(argc, argv, &i);
        else if (!strcmp(a, "--legacy")) o->legacy = 1;
        else if (!strcmp(a, "--autotune")) o->autotune = 1;
        else if (!strcmp(a, "--mode") && i + 1 < argc) {
            const char *m = argv[++i];
            if (!strcmp(m, "fp64")) o->mode = PN1_FP64;
            else if (!strcmp(m, "mixed")) o->mode = PN1_MIXED;
            else usage();
        } else usage();
    }
}

static void *xmalloc(size_t n)
{
    void *p = calloc(n ? n : 1, 1);
    if (!p) {
        perror("sticky9");
        exit(1);
    }
    return p;
}

static void snapshot(const char *path, const double *x, const double *v, int n)
{
    if (pn1_write_snapshot(path, 1, x, v, n) != PN1_OK) {
        fprintf(stderr, "sticky9: cannot write %s\n", path);
        exit(1);
    }
}

static void die_ctx(const char *what, pn1_ctx *ctx, int rc)
{
    fprintf(stderr, "sticky9: %s failed (%d): %s\n", what, rc, ctx ? pn1_last_error(ctx) : "");
    exit(1);
}

/* Energy block of Limit-sticky9.c:560-563 for the resident path. */
static void print_energy(double t, const pn1_energy_t *e, double init_ene, int initial)
{
    printf("time = %g\n", t);
    printf("pot = %22.15e kin = %22.15e \n pot_pn = %22.15e \n  total= %22.15e ratio = %e\n",
           e->pot, e->kin, e->pot_pn, e->total, e->kin / e->pot);
    if (!initial) printf(" DE = %e %g\n", (init_ene - e->total) / e->total, t);
}

int main(int argc, char **argv)
{
    opts_t o;
    parse(argc, argv, &o);
    if (o.ic) o.n = o.nstars + 1;
    if (o.resume) {
        int rn = 0;
        if (pn1_read_checkpoint(o.resume, &rn, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, 0)
            != PN1_OK) {
            fprintf(stderr, "sticky9: cannot read checkpoint %s\n", o.resume);
            return 1;
        }
        o.n = rn;
    }
    const int n = o.n;
    if (n < 1) usage();
    double *x = xmalloc((size_t)n * 3 * sizeof *x), *v = xmalloc((size_t)n * 3 * sizeof *v);
    double *m = xmalloc((size_t)n * sizeof *m);
    double time = 0.0;
    long long step0 = 0;

    FILE *fp = fopen(o.out, "w");   /* truncate, as fopen(.., "w") in the reference */
    if (!fp) {
        perror("data_output");
        return 1;
    }
    fclose(fp);

    if (o.resume) {
        int rn;
        double eps, c, dt;
        if (pn1_read_checkpoint(o.resume, &rn, &time, &step0, &eps, &c, &dt, x, v, m, n)
            != PN1_OK) {
            fprintf(stderr, "sticky9: cannot read checkpoint %s\n", o.resume);
            return 1;
        }
        printf("Init \n");
    } else if (o.ic) {
        int rc;
        double m0 = o.m0 > 0.0 ? o.m0 : 0.01;
        if (!strcmp(o.ic, "sphere"))
            rc = pn1_ic_uniform_sphere(o.nstars, o.seed, 1.0, 0.5, m0, o.eps, x, v, m);
        else if (!strcmp(o.ic, "plummer"))
            rc = pn1_ic_plummer(o.nstars, o.seed, 1.0, m0, x, v, m);
        else
            usage();
        if (rc != PN1_OK) {
            fprintf(stderr, "sticky9: initial conditions failed (%d)\n", rc);
            return 1;
        }
        printf("Init \n");
    } else {
        FILE *fin = fopen(o.in, "r");
        if (!fin) {
            perror("data_input");   /* Limit-sticky9.c:781-784 */
            exit(1);
        }
        fclose(fin);
        printf("Init \n");
        if (pn1_read_dat(o.in, n, x, v) != PN1_OK) {
            fprintf(stderr, "sticky9: %s is short or malformed for n = %d\n", o.in, n);
            return 1;
        }
        m[0] = o.m0 > 0.0 ? o.m0 : 100.0 / (double)n;
        for (int i = 1; i < n; i++) m[i] = 1.0 / (double)(n - 1);
    }
    printf("start \n");
    snapshot(o.out, x, v, n);

    double eouttime = time + o.deouttime, init_ene = 0.0;
    long long step = step0;
#define DONE() (o.steps >= 0 ? step - step0 >= o.steps : !(time < o.endtime))

    if (o.legacy) {
        /* the reference's loop on host arrays, through the drop-in shapes */
        if (pn1_set_c(o.c) != PN1_OK || pn1_set_mode(o.mode) != PN1_OK) usage();
        size_t n3 = (size_t)n * 3;
        double *a = xmalloc(n3 * sizeof *a), *pot = xmalloc((size_t)n * sizeof *pot);
        double *x1 = xmalloc(n3 * sizeof *x1), *v1 = xmalloc(n3 * sizeof *v1);
        double *hx = xmalloc(4 * n3 * sizeof *hx), *hv = xmalloc(4 * n3 * sizeof *hv);
        double (*X)[3] = (double (*)[3])x, (*V)[3] = (double (*)[3])v;
        double (*A)[3] = (double (*)[3])a;
        double (*X1)[3] = (double (*)[3])x1, (*V1)[3] = (double (*)[3])v1;
        initial_energy(X, V, m, n, &init_ene, o.eps);
        while (!DONE()) {
            memcpy(hx, x, n3 * sizeof *x);
            memcpy(hv, v, n3 * sizeof *v);
            force(X, V, m, o.eps, A, pot, n);
            for (size_t q = 0; q < n3; q++) {
                x1[q] = x[q] + 0.5 * v[q] * o.dt;
                v1[q] = v[q] + 0.5 * a[q] * o.dt;
                hx[n3 + q] = x1[q];
                hv[n3 + q] = v1[q];
            }
            force(X1, V1, m, o.eps, A, pot, n);
            for (size_t q = 0; q < n3; q++) {
                x1[q] = x[q] + 0.5 * v1[q] * o.dt;
                v1[q] = v[q] + 0.5 * a[q] * o.dt;
                hx[2 * n3 + q] = x1[q];
                hv[2 * n3 + q] = v1[q];
            }
            force(X1, V1, m, o.eps, A, pot, n);
            for (size_t q = 0; q < n3; q++) {
                x1[q] = x[q] + v1[q] * o.dt;
                v1[q] = v[q] + a[q] * o.dt;
                hx[3 * n3 + q] = x1[q];
                hv[3 * n3 + q] = v1[q];
            }
            for (size_t q = 0; q < n3; q++) {
                x[q] = (hx[q] + hx[3 * n3 + q] + 2.0 * (hx[n3 + q] + hx[2 * n3 + q])) / 6.0;
                v[q] = (hv[q] + hv[3 * n3 + q] + 2.0 * (hv[n3 + q] + hv[2 * n3 + q])) / 6.0;
            }
            time += o.dt;
            step++;
            if (time >= eouttime) {
                energy(time, X, V, m, n, init_ene, o.eps);
                eouttime += o.deouttime;
                snapshot(o.out, x, v, n);
            }
        }
        free(a); free(pot); free(x1); free(v1); free(hx); free(hv);
    } else {
        pn1_ctx *ctx = NULL;
        int rc = pn1_create(&ctx, n, 0);
        if (rc != PN1_OK) die_ctx("pn1_create", ctx, rc);
        if ((rc = pn1_set_params(ctx, o.eps, o.c, o.dt, o.mode)) != PN1_OK) die_ctx("pn1_set_params", ctx, rc);
        if ((rc = pn1_upload(ctx, x, v, m, n)) != PN1_OK) die_ctx("pn1_upload", ctx, rc);
        if (o.autotune) {
            int S = 0, var = 0;
            double ms = 0.0;
            if ((rc = pn1_autotune(ctx, 2, &S, &var, &ms)) != PN1_OK)
                die_ctx("pn1_autotune", ctx, rc);
            fprintf(stderr, "sticky9: autotuned j split %d, variant %d (%.3f ms per pair launch)\n",
                    S, var, ms);
        }
        pn1_energy_t e;
        if ((rc = pn1_energy(ctx, &e)) != PN1_OK) die_ctx("pn1_energy", ctx, rc);
        init_ene = e.total;
        print_energy(0.0, &e, init_ene, 1);
        while (!DONE()) {
            /* steps up to the next snapshot (time advances by repeated
             * addition, as in the reference, so its cadence is kept) */
            int k = 0;
            double t = time;
            long long left = o.steps >= 0 ? o.steps - (step - step0) : -1;
            do {
                t += o.dt;
                k++;
            } while (t < eouttime && (left < 0 ? t < o.endtime : k < left));
            if ((rc = pn1_step(ctx, k)) != PN1_OK) die_ctx("pn1_step", ctx, rc);
            time = t;
            step += k;
            if (time >= eouttime) {
                if ((rc = pn1_energy(ctx, &e)) != PN1_OK) die_ctx("pn1_energy", ctx, rc);
                print_energy(time, &e, init_ene, 0);
                eouttime += o.deouttime;
                if ((rc = pn1_download(ctx, x, v)) != PN1_OK) die_ctx("pn1_download", ctx, rc);
                snapshot(o.out, x, v, n);
            }
        }
        if ((rc = pn1_download(ctx, x, v)) != PN1_OK) die_ctx("pn1_download", ctx, rc);
        pn1_destroy(ctx);
    }
    snapshot(o.out, x, v, n);
    if (o.ckpt && pn1_write_checkpoint(o.ckpt, n, time, step, o.eps, o.c, o.dt, x, v, m) != PN1_OK) {
        fprintf(stderr, "sticky9: cannot write checkpoint %s\n", o.ckpt);
        return 1;
    }
    printf("%lf \n", time);
    free(x); free(v); free(m);
    return 0;
}
