/* Exercises the host-side helpers of libpn1 (pn_amd/csrc/pn1_host.c: input
 * validation, seeded initial conditions, .dat reader, snapshot writer,
 * binary checkpoints) on good and bad input.  Built with AddressSanitizer
 * and UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py; any
 * out-of-bounds access, leak or undefined operation aborts the run.  Prints
 * "ok <checks>" at the end. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pn1.h"

static int checks = 0;
#define CHECK(cond)                                                            \
    do {                                                                       \
        ++checks;                                                              \
        if (!(cond)) {                                                         \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
            exit(1);                                                           \
        }                                                                      \
    } while (0)

static double *dalloc(int n) { return (double *)calloc((size_t)n, sizeof(double)); }

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";
    char path[1024], msg[256];
    /* initial conditions at several sizes, including one star */
    const int sizes[] = {1, 2, 3, 63, 257};
    for (unsigned k = 0; k < sizeof sizes / sizeof sizes[0]; ++k) {
        const int ns = sizes[k], n = ns + 1;
        double *x = dalloc(3 * n), *v = dalloc(3 * n), *m = dalloc(n);
        CHECK(pn1_ic_plummer(ns, 7 + k, 1.0, 0.05, x, v, m) == PN1_OK);
        CHECK(pn1_ic_uniform_sphere(ns, 11 + k, 1.0, 0.5, 0.01, 1e-3, x, v, m) == PN1_OK);
        for (int i = 0; i < 3 * n; ++i) CHECK(isfinite(x[i]) && isfinite(v[i]));
        /* a star may land on the BH for ns = 1 (the sampler centres the stars) */
        if (ns > 1) CHECK(pn1_check_input(x, v, m, n, 1e-3, msg, (int)sizeof msg) == PN1_OK);
        free(x); free(v); free(m);
    }
    /* validation rejects what the reference would turn into Inf or NaN */
    {
        const int n = 4;
        double *x = dalloc(3 * n), *v = dalloc(3 * n), *m = dalloc(n);
        CHECK(pn1_ic_plummer(n - 1, 3, 1.0, 0.05, x, v, m) == PN1_OK);
        m[0] = 0.0;
        CHECK(pn1_check_input(x, v, m, n, 1e-3, msg, (int)sizeof msg) != PN1_OK);
        m[0] = 0.05;
        CHECK(pn1_check_input(x, v, m, n, 0.0, msg, (int)sizeof msg) != PN1_OK);
        x[3] = x[0]; x[4] = x[1]; x[5] = x[2];          /* star 1 on the BH */
        CHECK(pn1_check_input(x, v, m, n, 1e-3, msg, (int)sizeof msg) != PN1_OK);
        x[3] = NAN;
        CHECK(pn1_check_input(x, v, m, n, 1e-3, msg, 1) != PN1_OK);   /* 1-byte message */
        CHECK(pn1_check_input(x, v, m, n, 1e-3, NULL, 0) != PN1_OK);
        free(x); free(v); free(m);
    }
    /* .dat reader: good file, short file, missing file */
    {
        const int n = 5;
        snprintf(path, sizeof path, "%s/in.dat", dir);
        FILE *f = fopen(path, "w");
        CHECK(f != NULL);
        for (int i = 0; i < 2 * (n - 1); ++i) fprintf(f, "%f %f %f \n", 0.1 * i, -0.2 * i, 0.3);
        fclose(f);
        double *x = dalloc(3 * n), *v = dalloc(3 * n);
        CHECK(pn1_read_dat(path, n, x, v) == PN1_OK);
        CHECK(x[0] == 0.0 && v[0] == 0.0);
        snprintf(path, sizeof path, "%s/short.dat", dir);
        f = fopen(path, "w");
        fprintf(f, "1 2 3\n4 5");
        fclose(f);
        CHECK(pn1_read_dat(path, n, x, v) != PN1_OK);
        snprintf(path, sizeof path, "%s/missing.dat", dir);
        CHECK(pn1_read_dat(path, n, x, v) != PN1_OK);
        free(x); free(v);
    }
    /* snapshot and checkpoint round trips */
    {
        const int ns = 63, n = ns + 1;
        double *x = dalloc(3 * n), *v = dalloc(3 * n), *m = dalloc(n);
        double *x2 = dalloc(3 * n), *v2 = dalloc(3 * n), *m2 = dalloc(n);
        CHECK(pn1_ic_plummer(ns, 5, 1.0, 0.05, x, v, m) == PN1_OK);
        snprintf(path, sizeof path, "%s/snap.dat", dir);
        CHECK(pn1_write_snapshot(path, 0, x, v, n) == PN1_OK);
        CHECK(pn1_write_snapshot(path, 1, x, v, n) == PN1_OK);
        snprintf(path, sizeof path, "%s/ck.bin", dir);
        CHECK(pn1_write_checkpoint(path, n, 1.5, 42, 1e-3, 100.0, 0x1p-10, x, v, m) == PN1_OK);
        int n2 = 0;
        long long step = 0;
        double t = 0, eps = 0, c = 0, dt = 0;
        CHECK(pn1_read_checkpoint(path, &n2, &t, &step, &eps, &c, &dt, NULL, NULL, NULL, 0) == PN1_OK);
        CHECK(n2 == n && step == 42 && t == 1.5 && c == 100.0);
        CHECK(pn1_read_checkpoint(path, &n2, &t, &step, &eps, &c, &dt, x2, v2, m2, n - 1) != PN1_OK);
        CHECK(pn1_read_checkpoint(path, &n2, &t, &step, &eps, &c, &dt, x2, v2, m2, n) == PN1_OK);
        CHECK(memcmp(x, x2, 3 * n * sizeof(double)) == 0 && memcmp(m, m2, n * sizeof(double)) == 0);
        /* a truncated checkpoint */
        snprintf(path, sizeof path, "%s/ck_trunc.bin", dir);
        FILE *f = fopen(path, "wb");
        fwrite("PN1", 1, 3, f);
        fclose(f);
        CHECK(pn1_read_checkpoint(path, &n2, &t, &step, &eps, &c, &dt, x2, v2, m2, n) != PN1_OK);
        free(x); free(v); free(m); free(x2); free(v2); free(m2);
    }
    printf("ok %d\n", checks);
    return 0;
}
