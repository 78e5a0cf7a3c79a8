/*
 * ls9_oracle_cs.c -- complex-step derivatives of the oracle's force() and
 * stage driver.  TEST INFRASTRUCTURE ONLY (the checker of the adjoint).
 *
 * The same source as the FP64 oracle (ls9_force_body.h) on double complex:
 * perturbing one input by i h (h ~ 1e-30) makes the imaginary part of every
 * output h times its derivative with respect to that input, exact to
 * rounding (no subtractive cancellation, unlike finite differences).  The
 * real parts are the FP64 oracle's values to rounding.
 *
 * (ls9_oracle.h declares these on double *, the interleaved layout of a
 * double complex array; this file does not include it for that reason.)
 *
 * Pinning: the FP64 instantiation of this code is the oracle, pinned bit
 * for bit to the compiled reference; the complex instantiation is checked
 * against central finite differences of the FP64 oracle in tests/.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>



static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p) {
        fprintf(stderr, "ls9_oracle_cs: out of memory\n");
        exit(1);
    }
    return p;
}

#define REAL double complex
#define SQRT(x) csqrt(x)
#define FN(name) name##_cs
#include "ls9_force_body.h"
