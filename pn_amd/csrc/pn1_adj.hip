// pn1_adj.hip -- adjoint of libpn1's force and stage driver (the
// pn1_force_vjp / pn1_adjoint_* entry points).  A translation unit of its
// own: it is built with -ffinite-math-only -fno-signed-zeros so that the
// dual-number arithmetic of the VJP kernels drops the zero tangents of its
// seeds at compile time; the forward kernels keep strict IEEE semantics in
// pn1_api.hip.
//
// The stage driver (Limit-sticky9.c:841-893) maps a state s = (x, v) to
//   Y1 = s + dt/2 F(s),  Y2 = s + dt/2 F(Y1),  s' = s + dt/6 (F(s) + F(Y1) + F(Y2))
// with F(y) = (y_v, a(y)) (its weights sum to 1/2, as in the reference).
// Its discrete adjoint, for the cotangent L' of s':
//   LY2 = J_F(Y2)^T (dt/6 L')
//   LY1 = J_F(Y1)^T (dt/6 L' + dt/2 LY2)
//   L   = L' + LY2 + LY1 + J_F(s)^T (dt/6 L' + dt/2 LY1)
// where J_F(y)^T (mu_x, mu_v) = (J_a,x^T mu_v, mu_x + J_a,v^T mu_v), and
// d/dm0 collects (d a / d m0)^T mu_v of all three.
//
// pn1_adjoint_forward steps the state with the FP64 kernels and keeps a
// checkpoint every `interval` steps; pn1_adjoint_backward re-runs each
// segment from its checkpoint (keeping its states), then walks it backward,
// recomputing Y1 and Y2 of each step with two force evaluations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pn1_adj_kernels.hpp"
#include "pn1_ctx.hpp"

using namespace pn1dev;
using namespace pn1host;

struct pn1_adj {
    int cap = 0;
    double *mu = nullptr, *out = nullptr, *L = nullptr, *LY1 = nullptr, *LY2 = nullptr;
    double *T = nullptr, *blk = nullptr, *gm0 = nullptr;
    ARec* rec = nullptr;
    double* ckpt = nullptr;     // nck states of 6 x cap
    double* seg = nullptr;      // interval states of 6 x cap
    int nck_cap = 0, seg_cap = 0;
    // recorded by pn1_adjoint_forward
    bool have_fwd = false;
    int steps = 0, interval = 0;
    unsigned gen = 0;
};

namespace pn1host {

void adj_free(pn1_ctx* c)
{
    pn1_adj* a = c ? c->adj : nullptr;
    if (!a) return;
    void* bufs[] = {a->mu, a->out, a->L, a->LY1, a->LY2, a->T, a->blk, a->gm0, a->rec, a->ckpt,
                    a->seg};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete a;
    c->adj = nullptr;
}

}  // namespace pn1host

namespace {

template <typename T>
int aalloc(pn1_ctx* c, T** p, size_t count)
{
    if (hipMalloc((void**)p, (count ? count : 1) * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return seterr(c, PN1_ENOMEM, "hipMalloc of %zu bytes (adjoint) failed", count * sizeof(T));
    }
    return PN1_OK;
}

// Buffers sized for the context's capacity (allocated once).
int adj_get(pn1_ctx* c, pn1_adj** out)
{
    if (!c->adj) {
        pn1_adj* a = new pn1_adj;
        c->adj = a;
        a->cap = c->cap;
        const size_t s6 = 6 * (size_t)c->cap;
        const size_t nb = (size_t)((c->cap + WG - 1) / WG + 1);
        int rc = PN1_OK;
        if (!rc) rc = aalloc(c, &a->mu, s6);
        if (!rc) rc = aalloc(c, &a->out, s6);
        if (!rc) rc = aalloc(c, &a->L, s6);
        if (!rc) rc = aalloc(c, &a->LY1, s6);
        if (!rc) rc = aalloc(c, &a->LY2, s6);
        if (!rc) rc = aalloc(c, &a->T, s6);
        if (!rc) rc = aalloc(c, &a->blk, nb * 4);
        if (!rc) rc = aalloc(c, &a->gm0, 2);
        if (!rc) rc = aalloc(c, &a->rec, (size_t)c->cap + 2);
        if (rc) {
            adj_free(c);
            return rc;
        }
    }
    *out = c->adj;
    return PN1_OK;
}

int nblk_of(const pn1_ctx* c) { return c->n > 1 ? (c->n - 1 + WG - 1) / WG : 1; }

// out = J_F(y)^T mu for the state y (SoA 6 x cap); d/dm0 to gm0[1], added
// to gm0[0].  FP64.
int enqueue_vjp(pn1_ctx* c, pn1_adj* a, const double* y, const double* mu, double* out,
                const Phys& P)
{
    const int nblk = nblk_of(c);
    if (c->n > 1) {
        hipLaunchKernelGGL(k_adj_prep, dim3(nblk), dim3(WG), 0, c->st, y, mu, c->m, P, a->rec);
        const dim3 grid(c->pair_grid), blk(WG);
        double* part = (double*)c->part;
        hipLaunchKernelGGL(k_adj_pair<1>, grid, blk, 0, c->st, a->rec, y, mu, c->m, part, P,
                           c->S, c->nIblk);
        hipLaunchKernelGGL(k_adj_pair<2>, grid, blk, 0, c->st, a->rec, y, mu, c->m, part, P,
                           c->S, c->nIblk);
    }
    hipLaunchKernelGGL(k_adj_fin, dim3(nblk), dim3(WG), 0, c->st, a->rec, y, mu, c->m,
                       (const double*)c->part, c->S, out, a->blk, P);
    hipLaunchKernelGGL(k_adj_bh, dim3(1), dim3(WG), 0, c->st, a->blk, nblk, mu, out, a->gm0, P);
    HIPCHK(c, hipGetLastError());
    return PN1_OK;
}

void lin(pn1_ctx* c, double* out, double ca, const double* A, double cb = 0.0,
         const double* B = nullptr, double cc = 0.0, const double* C = nullptr, double cd = 0.0,
         const double* D = nullptr)
{
    hipLaunchKernelGGL(k_adj_lin, dim3((c->n + WG - 1) / WG), dim3(WG), 0, c->st, out, c->n,
                       c->cap, ca, A, cb, B, cc, C, cd, D);
}

// AoS host [n][3] pair (x part, v part) -> SoA device 6 x cap.
int put6(pn1_ctx* c, const double* hx, const double* hv, double* dst)
{
    const size_t cap = (size_t)c->cap;
    double* h = c->h_stage;
    for (int i = 0; i < c->n; ++i)
        for (int k = 0; k < 3; ++k) {
            h[k * cap + i] = hx ? hx[3 * i + k] : 0.0;
            h[(3 + k) * cap + i] = hv ? hv[3 * i + k] : 0.0;
        }
    for (int k = 0; k < 6; ++k)
        HIPCHK(c, hipMemcpyAsync(dst + k * cap, h + k * cap, c->n * sizeof(double),
                                 hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PN1_OK;
}

int copy6(pn1_ctx* c, double* dst, const double* src)
{
    const size_t cap = (size_t)c->cap;
    for (int k = 0; k < 6; ++k)
        HIPCHK(c, hipMemcpyAsync(dst + k * cap, src + k * cap, c->n * sizeof(double),
                                 hipMemcpyDeviceToDevice, c->st));
    return PN1_OK;
}

// The adjoint runs the FP64 kernels whatever the context's mode.
struct Fp64Scope {
    pn1_ctx* c;
    int mode;
    explicit Fp64Scope(pn1_ctx* cc) : c(cc), mode(cc->mode) { c->mode = PN1_FP64; }
    ~Fp64Scope() { c->mode = mode; }
};

int enqueue_stages(pn1_ctx* c, int last, const Phys& P)
{
    for (int stage = 1; stage <= last; ++stage) {
        int rc = enqueue_force(c, stage, P);
        if (rc) return rc;
    }
    return PN1_OK;
}

}  // namespace

extern "C" {

int pn1_force_vjp(pn1_ctx* c, const double* u, double* gx, double* gv, double* gm0)
{
    int rc;
    if (!ready(c, &rc)) return rc;
    if (!u || !gx || !gv) return seterr(c, PN1_EINVAL, "null argument");
    pn1_adj* a;
    if ((rc = adj_get(c, &a))) return rc;
    const Phys P = make_phys(c);
    if ((rc = put6(c, nullptr, u, a->mu))) return rc;
    HIPCHK(c, hipMemsetAsync(a->gm0, 0, 2 * sizeof(double), c->st));
    if ((rc = enqueue_vjp(c, a, c->X, a->mu, a->out, P))) return rc;
    const size_t cap = (size_t)c->cap;
    if ((rc = fetch3(c, a->out, gx))) return rc;
    if ((rc = fetch3(c, a->out + 3 * cap, gv))) return rc;
    if (gm0) {
        double h[2];
        HIPCHK(c, hipMemcpy(h, a->gm0, sizeof h, hipMemcpyDeviceToHost));
        *gm0 = h[1];
    }
    return PN1_OK;
}

int pn1_adjoint_forward(pn1_ctx* c, int steps, int interval)
{
    int rc;
    if (!ready(c, &rc)) return rc;
    if (steps < 1 || interval < 1)
        return seterr(c, PN1_EINVAL, "steps (%d) and interval (%d) must be >= 1", steps, interval);
    pn1_adj* a;
    if ((rc = adj_get(c, &a))) return rc;
    const int nck = (steps + interval - 1) / interval;
    const size_t s6 = 6 * (size_t)c->cap;
    if (nck > a->nck_cap) {
        if (a->ckpt) (void)hipFree(a->ckpt);
        a->ckpt = nullptr;
        a->nck_cap = 0;
        if ((rc = aalloc(c, &a->ckpt, (size_t)nck * s6))) return rc;
        a->nck_cap = nck;
    }
    const int seg = std::min(interval, steps);
    if (seg > a->seg_cap) {
        if (a->seg) (void)hipFree(a->seg);
        a->seg = nullptr;
        a->seg_cap = 0;
        if ((rc = aalloc(c, &a->seg, (size_t)seg * s6))) return rc;
        a->seg_cap = seg;
    }
    Fp64Scope fp64(c);
    drop_graph(c);
    const Phys P = make_phys(c);
    for (int t = 0; t < steps; ++t) {
        if (t % interval == 0 && (rc = copy6(c, a->ckpt + (size_t)(t / interval) * s6, c->X)))
            return rc;
        if ((rc = enqueue_stages(c, 3, P))) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    a->have_fwd = true;
    a->steps = steps;
    a->interval = interval;
    a->gen = c->gen;
    return PN1_OK;
}

int pn1_adjoint_backward(pn1_ctx* c, const double* gx, const double* gv, double* gx0,
                         double* gv0, double* gm0)
{
    int rc;
    if (!ready(c, &rc)) return rc;
    if (!gx || !gv || !gx0 || !gv0) return seterr(c, PN1_EINVAL, "null argument");
    pn1_adj* a = c->adj;
    if (!a || !a->have_fwd) return seterr(c, PN1_EINVAL, "no pn1_adjoint_forward run to reverse");
    if (a->gen != c->gen)
        return seterr(c, PN1_EINVAL, "state or parameters changed since pn1_adjoint_forward");
    Fp64Scope fp64(c);
    drop_graph(c);
    const Phys P = make_phys(c);
    const double dt = c->dt;
    const size_t s6 = 6 * (size_t)c->cap;
    if ((rc = put6(c, gx, gv, a->L))) return rc;
    HIPCHK(c, hipMemsetAsync(a->gm0, 0, 2 * sizeof(double), c->st));
    const int nck = (a->steps + a->interval - 1) / a->interval;
    for (int k = nck - 1; k >= 0; --k) {
        const int t0 = k * a->interval, t1 = std::min(a->steps, t0 + a->interval);
        // states of the segment
        if ((rc = copy6(c, c->X, a->ckpt + (size_t)k * s6))) return rc;
        for (int t = t0; t < t1; ++t) {
            if ((rc = copy6(c, a->seg + (size_t)(t - t0) * s6, c->X))) return rc;
            if (t + 1 < t1 && (rc = enqueue_stages(c, 3, P))) return rc;
        }
        for (int t = t1 - 1; t >= t0; --t) {
            if ((rc = copy6(c, c->X, a->seg + (size_t)(t - t0) * s6))) return rc;
            if ((rc = enqueue_stages(c, 2, P))) return rc;     // H1 = Y1, H2 = Y2
            lin(c, a->mu, dt / 6.0, a->L);
            if ((rc = enqueue_vjp(c, a, c->H2, a->mu, a->LY2, P))) return rc;
            lin(c, a->mu, dt / 6.0, a->L, dt / 2.0, a->LY2);
            if ((rc = enqueue_vjp(c, a, c->H1, a->mu, a->LY1, P))) return rc;
            lin(c, a->mu, dt / 6.0, a->L, dt / 2.0, a->LY1);
            if ((rc = enqueue_vjp(c, a, c->X, a->mu, a->T, P))) return rc;
            lin(c, a->L, 1.0, a->L, 1.0, a->LY2, 1.0, a->LY1, 1.0, a->T);
        }
    }
    HIPCHK(c, hipGetLastError());
    const size_t cap = (size_t)c->cap;
    if ((rc = fetch3(c, a->L, gx0))) return rc;
    if ((rc = fetch3(c, a->L + 3 * cap, gv0))) return rc;
    if (gm0) {
        double h[2];
        HIPCHK(c, hipMemcpy(h, a->gm0, sizeof h, hipMemcpyDeviceToHost));
        *gm0 = h[0];
    }
    return PN1_OK;
}

}  // extern "C"
