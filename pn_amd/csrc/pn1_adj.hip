// pn1_adj.hip -- adjoint of libpn1's force and stage driver (the
// pn1_force_vjp / pn1_adjoint_* entry points).  A translation unit of its
// own: it is built with -ffinite-math-only -fno-signed-zeros so that the
// dual-number arithmetic of the per-star VJP terms drops the zero tangents
// of its seeds at compile time; the forward kernels keep strict IEEE
// semantics in pn1_api.hip.  The O(n^2) pair sums of the VJP run in one
// pass (k_adj_rev, pn1_adj_rev.hpp): a hand-written reverse sweep per pair
// gives both of its sides, the j records rotating through the wave's lanes.
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
// segment from its checkpoint, keeping every step's s, Y1 and Y2 (the
// stage buffers H1, H2 of the step), then walks it backward with three
// VJPs per step and no further force evaluation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pn1_adj_rev.hpp"
#include "pn1_ctx.hpp"
#include "pn1_ens.hpp"

using namespace pn1dev;
using namespace pn1host;

struct pn1_adj {
    int cap = 0;
    double *mu = nullptr, *out = nullptr, *L = nullptr, *LY1 = nullptr, *LY2 = nullptr;
    double *T = nullptr, *blk = nullptr, *gm0 = nullptr;
    ARec* rec = nullptr;
    double* rpart = nullptr;    // pair kernel partials, rpart_n doubles
    size_t rpart_n = 0;
    double* ckpt = nullptr;     // nck states of 6 x cap
    double* seg = nullptr;      // per step of a segment: s, Y1, Y2 (6 x cap each)
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
    void* bufs[] = {a->mu, a->out, a->L, a->LY1, a->LY2, a->T, a->blk, a->gm0, a->rec, a->rpart,
                    a->ckpt, a->seg};
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
    if (dev_malloc((void**)p, (count ? count : 1) * sizeof(T)) != hipSuccess) {
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

// j chunk L of the single-pass pair kernel for nst stars in nIb blocks of
// 256: whole 64-record tiles, the longest that still gives about
// PN1_ADJ_BLOCKS workgroups (several rounds of the whole device), within
// [128, 2048].  At n = 16385, L = 128 ... 1024 time within 2% of each
// other.  The partials take ceil(nst / L) * 7 + nIb * 6 rows of cap doubles:
// the j-side rows (one per 256-star block) dominate, ~n^2 / 5.3 bytes.
#ifndef PN1_ADJ_BLOCKS
#define PN1_ADJ_BLOCKS 2048
#endif
int adj_chunk(int nst, int nIb)
{
    const int s0 = std::max(1, (PN1_ADJ_BLOCKS + nIb - 1) / std::max(nIb, 1));
    int L = (nst + s0 - 1) / s0;
    L = (L + 63) / 64 * 64;
    return std::min(std::max(L, 128), 2048);
}

// out = J_F(y)^T mu for the state y (SoA 6 x cap); d/dm0 to gm0[1], added
// to gm0[0].  FP64.
int enqueue_vjp(pn1_ctx* c, pn1_adj* a, const double* y, const double* mu, double* out,
                const Phys& P)
{
    const int nblk = nblk_of(c);
    int S = 0, nIb = 0;
    if (c->n > 1) {
        // an explicit j split (pn1_set_tuning, pn1_autotune) sets the chunk
        // count, rounded to whole tiles; otherwise the VJP picks its own
        const int nst = c->n - 1;
        const int L = c->jsplit_req > 0 ? ((nst + c->S - 1) / c->S + 63) / 64 * 64
                                        : adj_chunk(nst, c->nIblk);
        S = (nst + L - 1) / L;
        nIb = c->nIblk;
        const size_t need = ((size_t)S * NADJ1 + (size_t)nIb * NADJ2) * (size_t)c->cap;
        if (need > a->rpart_n) {
            if (a->rpart) (void)hipFree(a->rpart);
            a->rpart = nullptr;
            a->rpart_n = 0;
            int rc = aalloc(c, &a->rpart, need);
            if (rc) return rc;
            a->rpart_n = need;
        }
        hipLaunchKernelGGL(k_adj_prep, dim3(nblk), dim3(WG), 0, c->st, y, mu, c->m, P, a->rec);
        hipLaunchKernelGGL(k_adj_rev, dim3(S * nIb), dim3(WG), 0, c->st, a->rec, y, mu, c->m,
                           a->rpart, P, L, S);
    }
    hipLaunchKernelGGL(k_adj_fin, dim3(nblk), dim3(WG), 0, c->st, a->rec, y, mu, c->m,
                       (const double*)a->rpart, S, nIb, out, a->blk, P);
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
    // a run that stops below leaves nothing for pn1_adjoint_backward to
    // reverse (its checkpoint buffers may have been replaced)
    a->have_fwd = false;
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
        if ((rc = aalloc(c, &a->seg, 3 * (size_t)seg * s6))) return rc;
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
        // the segment's states: per step s, Y1 = H1 and Y2 = H2 (the last
        // step runs two stages only: its end state is not needed)
        if ((rc = copy6(c, c->X, a->ckpt + (size_t)k * s6))) return rc;
        for (int t = t0; t < t1; ++t) {
            double* st = a->seg + 3 * (size_t)(t - t0) * s6;
            if ((rc = copy6(c, st, c->X))) return rc;
            if ((rc = enqueue_stages(c, t + 1 < t1 ? 3 : 2, P))) return rc;
            if ((rc = copy6(c, st + s6, c->H1))) return rc;
            if ((rc = copy6(c, st + 2 * s6, c->H2))) return rc;
        }
        for (int t = t1 - 1; t >= t0; --t) {
            const double* st = a->seg + 3 * (size_t)(t - t0) * s6;
            lin(c, a->mu, dt / 6.0, a->L);
            if ((rc = enqueue_vjp(c, a, st + 2 * s6, a->mu, a->LY2, P))) return rc;
            lin(c, a->mu, dt / 6.0, a->L, dt / 2.0, a->LY2);
            if ((rc = enqueue_vjp(c, a, st + s6, a->mu, a->LY1, P))) return rc;
            lin(c, a->mu, dt / 6.0, a->L, dt / 2.0, a->LY1);
            if ((rc = enqueue_vjp(c, a, st, a->mu, a->T, P))) return rc;
            lin(c, a->L, 1.0, a->L, 1.0, a->LY2, 1.0, a->LY1, 1.0, a->T);
        }
    }
    if ((rc = copy6(c, c->X, a->ckpt))) return rc;          // leave the initial state
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

// =========================================================================
// Ensembles (pn1_ens_force_vjp, pn1_ens_adjoint_*): the same VJP and stage
// adjoint for every cluster at once.  The single-pass pair kernel is a
// persistent launch over the ensemble's own work items (cluster, 256-star
// block, j chunk) and per-XCD queues; each item writes its own partial rows.
// =========================================================================
namespace pn1dev {

__global__ __launch_bounds__(WG) void k_ens_adj_prep(const double* __restrict__ y,
                                                    const double* __restrict__ mu,
                                                    const double* __restrict__ m,
                                                    const int* __restrict__ coff,
                                                    const EBlk* __restrict__ blk, Phys P,
                                                    ARec* __restrict__ rec)
{
    const EBlk b = blk[blockIdx.x];
    const int off = coff[b.c], nst = coff[b.c + 1] - off - 1;
    const int li = b.first + (int)threadIdx.x;
    if (li < nst) adj_prep_star(y, mu, m, (size_t)P.cap, P.eps2, off + 1 + li, off, rec);
}

__global__ void k_ens_adj_reset(unsigned int* __restrict__ counter)
{
    if (threadIdx.x < NXCD) counter[threadIdx.x] = 0u;
}

// Persistent single-pass pair kernel over the ensemble's work items.  Item
// (cluster c, block ib, chunk s): stars off + 1 + ib*WG + lane, records
// [j0, j1).  The partials of cluster c start at apart + aoff[c], rows of
// stride nst: i side (s*NADJ1 + k), s < S_c; then j side
// (S_c*NADJ1 + ib*NADJ2 + k).  Every item writes its own rows.
__global__ __launch_bounds__(WG) void k_ens_adj_rev(const ARec* __restrict__ rec,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ mu,
                                                   const double* __restrict__ m,
                                                   double* __restrict__ apart,
                                                   const long long* __restrict__ aoff,
                                                   const EItem* __restrict__ items,
                                                   unsigned int* __restrict__ counter,
                                                   const int* __restrict__ qoff, Phys P, int L)
{
    __shared__ double sh[NADJ2][4][64];
    __shared__ double tls[4][REVJ_F][64];
    const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    int q = 0;
    for (;;) {
        const int it = ens_take(counter, qoff, q);
        if (it < 0) break;
        const EItem w = items[it];
        const int nst = w.nst;
        const int li = w.ib * WG + (int)threadIdx.x;
        const bool iok = li < nst;
        const bool live = w.ib * WG + (__builtin_amdgcn_readfirstlane((int)threadIdx.x) & ~63) < nst;
        const int i = w.off + 1 + (iok ? li : nst - 1);
        const ABH B = load_abh(y, mu, m, P.cap, w.off);
        const RevU U = rev_uniform(B, P.eps2, P.lam);
        const ARec me = rec[i];
        const RevI I = rev_setup(me, U);
        RevIAcc ai;
        rev_zero(ai);
        const int S = ens_nchunks(nst, L), s = (w.j0 - w.off - 1) / L;
        double* base = apart + aoff[w.c];
        double* pj = base + (size_t)S * NADJ1 * nst;
        for (int jt = w.j0; jt < w.j1; jt += 64) {
            double g[NADJ2];
            rev_tile(rec, I, U, i, iok, live, jt, w.j1, ai, g, tls[wv]);
#pragma unroll
            for (int k = 0; k < NADJ2; ++k) sh[k][wv][lane] = g[k];
            __syncthreads();
            const int jl = jt + lane;
            if (wv == 0 && jl < w.j1) {
#pragma unroll
                for (int k = 0; k < NADJ2; ++k)
                    pj[((size_t)w.ib * NADJ2 + k) * nst + (jl - w.off - 1)] =
                        ((sh[k][0][lane] + sh[k][1][lane]) + sh[k][2][lane]) + sh[k][3][lane];
            }
            __syncthreads();
        }
        double g[NADJ1];
        rev_ifin(me, U, ai, g);
        if (iok) {
#pragma unroll
            for (int k = 0; k < NADJ1; ++k) base[((size_t)s * NADJ1 + k) * nst + li] = g[k];
        }
    }
}

__global__ __launch_bounds__(WG) void k_ens_adj_fin(const ARec* __restrict__ rec,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ mu,
                                                   const double* __restrict__ m,
                                                   const int* __restrict__ coff,
                                                   const EBlk* __restrict__ blk,
                                                   const double* __restrict__ apart,
                                                   const long long* __restrict__ aoff,
                                                   double* __restrict__ out,
                                                   double* __restrict__ ablk, Phys P, int L)
{
    const EBlk b = blk[blockIdx.x];
    const int off = coff[b.c], nst = coff[b.c + 1] - off - 1;
    const int li = b.first + (int)threadIdx.x;
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    if (li < nst) {
        const int p = off + 1 + li;
        const ABH B = load_abh(y, mu, m, P.cap, off);
        double g[NADJ1];
        adj_star_grad(rec[p], B, P.eps, P.eps2, P.lam, g);
        const double* base = apart + aoff[b.c] + li;
        const int S = ens_nchunks(nst, L), nIb = (nst + WG - 1) / WG;
        for (int s = 0; s < S; ++s) {
#pragma unroll
            for (int k = 0; k < NADJ1; ++k) g[k] += base[(size_t)(s * NADJ1 + k) * nst];
        }
        const double* pj = base + (size_t)S * NADJ1 * nst;
        for (int ib = 0; ib < nIb; ++ib) {
#pragma unroll
            for (int k = 0; k < NADJ2; ++k) g[k] += pj[(size_t)(ib * NADJ2 + k) * nst];
        }
        adj_store(g, mu, out, (size_t)P.cap, p, t);
    }
    block_sum<4>(t, ablk + (size_t)blockIdx.x * 4);
}

// One workgroup per cluster: its BH's rows and its m0 derivative
// (gm0[2c + 1] this evaluation, added to gm0[2c]).
__global__ __launch_bounds__(WG) void k_ens_adj_bh(const double* __restrict__ ablk,
                                                  const int* __restrict__ cblk,
                                                  const int* __restrict__ coff,
                                                  const double* __restrict__ mu,
                                                  double* __restrict__ out,
                                                  double* __restrict__ gm0, Phys P)
{
    const int c = blockIdx.x;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = cblk[c] + (int)threadIdx.x; b < cblk[c + 1]; b += WG)
#pragma unroll
        for (int k = 0; k < 4; ++k) s[k] += ablk[(size_t)b * 4 + k];
    __shared__ double tot[4];
    block_sum<4>(s, tot);
    if (threadIdx.x == 0) {
        const size_t cap = (size_t)P.cap;
        const int b0 = coff[c];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            out[(size_t)k * cap + b0] = -tot[k];
            out[(size_t)(3 + k) * cap + b0] = mu[(size_t)k * cap + b0];
        }
        gm0[2 * c + 1] = tot[3];
        gm0[2 * c] += tot[3];
    }
}

}  // namespace pn1dev

struct pn1_ens_adj {
    double *mu = nullptr, *out = nullptr, *L = nullptr, *LY1 = nullptr, *LY2 = nullptr;
    double *T = nullptr, *ablk = nullptr, *gm0 = nullptr;
    ARec* rec = nullptr;
    double* ckpt = nullptr;
    double* seg = nullptr;        // per step of a segment: s, Y1, Y2 (6 x cap each)
    int nck_cap = 0, seg_cap = 0, ablk_cap = 0;
    int grid = 0;                 // resident workgroups of k_ens_adj_rev
    double* apart = nullptr;      // pair kernel partials, apart_n doubles
    size_t apart_n = 0;
    long long* aoff = nullptr;    // per-cluster offsets into apart (capc)
    unsigned tab_gen = 0;         // (gen, L) the offsets were made for
    int tab_L = -1;
    bool have_fwd = false;
    int steps = 0, interval = 0;
    unsigned gen = 0;
};

namespace pn1host {

void ens_adj_free(pn1_ens* e)
{
    pn1_ens_adj* a = e ? e->adj : nullptr;
    if (!a) return;
    void* bufs[] = {a->mu, a->out, a->L, a->LY1, a->LY2, a->T, a->ablk, a->gm0, a->rec,
                    a->ckpt, a->seg, a->apart, a->aoff};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete a;
    e->adj = nullptr;
}

}  // namespace pn1host

namespace {

template <typename T>
int ealloc_adj(pn1_ens* e, T** p, size_t count)
{
    if (dev_malloc((void**)p, (count ? count : 1) * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return eerr(e, PN1_ENOMEM, "hipMalloc of %zu bytes (adjoint) failed", count * sizeof(T));
    }
    return PN1_OK;
}

template <typename K>
int adj_resident_grid(pn1_ens* e, K kernel, int* out)
{
    int per_cu = 0, dev = 0;
    hipDeviceProp_t prop;
    EHIPCHK(e, hipGetDevice(&dev));
    EHIPCHK(e, hipGetDeviceProperties(&prop, dev));
    EHIPCHK(e, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, WG, 0));
    *out = (per_cu > 0 ? per_cu : 1) * prop.multiProcessorCount;
    return PN1_OK;
}

int ens_adj_get(pn1_ens* e, pn1_ens_adj** out)
{
    if (!e->adj) {
        pn1_ens_adj* a = new pn1_ens_adj;
        e->adj = a;
        const size_t s6 = 6 * (size_t)e->cap;
        int rc = PN1_OK;
        if (!rc) rc = ealloc_adj(e, &a->mu, s6);
        if (!rc) rc = ealloc_adj(e, &a->out, s6);
        if (!rc) rc = ealloc_adj(e, &a->L, s6);
        if (!rc) rc = ealloc_adj(e, &a->LY1, s6);
        if (!rc) rc = ealloc_adj(e, &a->LY2, s6);
        if (!rc) rc = ealloc_adj(e, &a->T, s6);
        if (!rc) rc = ealloc_adj(e, &a->gm0, 2 * (size_t)e->capc);
        if (!rc) rc = ealloc_adj(e, &a->rec, (size_t)e->cap + 2);
        if (!rc) rc = ealloc_adj(e, &a->aoff, (size_t)e->capc);
        if (!rc) rc = adj_resident_grid(e, k_ens_adj_rev, &a->grid);
        if (rc) {
            ens_adj_free(e);
            return rc;
        }
    }
    pn1_ens_adj* a = e->adj;
    if (e->nb > a->ablk_cap) {          // per-block partials follow the block table
        if (a->ablk) (void)hipFree(a->ablk);
        a->ablk = nullptr;
        a->ablk_cap = 0;
        int rc = ealloc_adj(e, &a->ablk, (size_t)e->nb * 4);
        if (rc) return rc;
        a->ablk_cap = e->nb;
    }
    if (a->tab_L != e->L || a->tab_gen != e->gen) {   // partial layout of the clusters
        std::vector<long long> aoff(e->nc > 0 ? e->nc : 1, 0);
        size_t tot = 0;
        for (int c = 0; c < e->nc; ++c) {
            const int nst = e->off_h[c + 1] - e->off_h[c] - 1;
            aoff[c] = (long long)tot;
            if (nst > 0)
                tot += ((size_t)((nst + e->L - 1) / e->L) * NADJ1 +
                        (size_t)((nst + WG - 1) / WG) * NADJ2) * (size_t)nst;
        }
        if (tot > a->apart_n) {
            if (a->apart) (void)hipFree(a->apart);
            a->apart = nullptr;
            a->apart_n = 0;
            int rc = ealloc_adj(e, &a->apart, tot);
            if (rc) return rc;
            a->apart_n = tot;
        }
        if (e->nc > 0)
            EHIPCHK(e, hipMemcpyAsync(a->aoff, aoff.data(), e->nc * sizeof(long long),
                                      hipMemcpyHostToDevice, e->st));
        EHIPCHK(e, hipStreamSynchronize(e->st));
        a->tab_L = e->L;
        a->tab_gen = e->gen;
    }
    *out = a;
    return PN1_OK;
}

int ens_enqueue_vjp(pn1_ens* e, pn1_ens_adj* a, const double* y, const double* mu,
                    double* out, const Phys& P)
{
    if (e->nb)
        hipLaunchKernelGGL(k_ens_adj_prep, dim3(e->nb), dim3(WG), 0, e->st, y, mu, e->m, e->coff,
                           e->blk, P, a->rec);
    if (e->nitems) {
        hipLaunchKernelGGL(k_ens_adj_reset, dim3(1), dim3(64), 0, e->st, e->counter);
        hipLaunchKernelGGL(k_ens_adj_rev, dim3(std::min(a->grid, e->nitems)), dim3(WG), 0, e->st,
                           a->rec, y, mu, e->m, a->apart, a->aoff, e->items, e->counter, e->qoff,
                           P, e->L);
    }
    if (e->nb)
        hipLaunchKernelGGL(k_ens_adj_fin, dim3(e->nb), dim3(WG), 0, e->st, a->rec, y, mu, e->m,
                           e->coff, e->blk, (const double*)a->apart, a->aoff, out, a->ablk, P,
                           e->L);
    hipLaunchKernelGGL(k_ens_adj_bh, dim3(e->nc), dim3(WG), 0, e->st, a->ablk, e->cblk, e->coff,
                       mu, out, a->gm0, P);
    EHIPCHK(e, hipGetLastError());
    return PN1_OK;
}

void ens_lin(pn1_ens* e, double* out, double ca, const double* A, double cb = 0.0,
             const double* B = nullptr, double cc = 0.0, const double* C = nullptr,
             double cd = 0.0, const double* D = nullptr)
{
    hipLaunchKernelGGL(k_adj_lin, dim3((e->n + WG - 1) / WG), dim3(WG), 0, e->st, out, e->n,
                       e->cap, ca, A, cb, B, cc, C, cd, D);
}

int ens_put6(pn1_ens* e, const double* hx, const double* hv, double* dst)
{
    const size_t cap = (size_t)e->cap;
    double* h = e->h_stage;
    for (int i = 0; i < e->n; ++i)
        for (int k = 0; k < 3; ++k) {
            h[k * cap + i] = hx ? hx[3 * i + k] : 0.0;
            h[(3 + k) * cap + i] = hv ? hv[3 * i + k] : 0.0;
        }
    for (int k = 0; k < 6; ++k)
        EHIPCHK(e, hipMemcpyAsync(dst + k * cap, h + k * cap, e->n * sizeof(double),
                                  hipMemcpyHostToDevice, e->st));
    EHIPCHK(e, hipStreamSynchronize(e->st));
    return PN1_OK;
}

int ens_copy6(pn1_ens* e, double* dst, const double* src)
{
    const size_t cap = (size_t)e->cap;
    for (int k = 0; k < 6; ++k)
        EHIPCHK(e, hipMemcpyAsync(dst + k * cap, src + k * cap, e->n * sizeof(double),
                                  hipMemcpyDeviceToDevice, e->st));
    return PN1_OK;
}

int ens_get_gm0(pn1_ens* e, pn1_ens_adj* a, int which, double* gm0)
{
    std::vector<double> h(2 * (size_t)e->nc);
    EHIPCHK(e, hipMemcpyAsync(h.data(), a->gm0, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                              e->st));
    EHIPCHK(e, hipStreamSynchronize(e->st));
    for (int c = 0; c < e->nc; ++c) gm0[c] = h[2 * (size_t)c + which];
    return PN1_OK;
}

struct EnsFp64Scope {
    pn1_ens* e;
    int mode;
    explicit EnsFp64Scope(pn1_ens* ee) : e(ee), mode(ee->mode) { e->mode = PN1_FP64; }
    ~EnsFp64Scope() { e->mode = mode; }
};

int ens_stages(pn1_ens* e, int last, const Phys& P)
{
    for (int stage = 1; stage <= last; ++stage) {
        int rc = ens_enqueue_force(e, stage, P);
        if (rc) return rc;
    }
    return PN1_OK;
}

}  // namespace

extern "C" {

int pn1_ens_force_vjp(pn1_ens* e, const double* u, double* gx, double* gv, double* gm0)
{
    int rc;
    if (!ens_ready(e, &rc)) return rc;
    if (!u || !gx || !gv) return eerr(e, PN1_EINVAL, "null argument");
    pn1_ens_adj* a;
    if ((rc = ens_adj_get(e, &a))) return rc;
    const Phys P = ens_phys(e);
    if ((rc = ens_put6(e, nullptr, u, a->mu))) return rc;
    EHIPCHK(e, hipMemsetAsync(a->gm0, 0, 2 * (size_t)e->nc * sizeof(double), e->st));
    if ((rc = ens_enqueue_vjp(e, a, e->X, a->mu, a->out, P))) return rc;
    const size_t cap = (size_t)e->cap;
    if ((rc = ens_fetch3(e, a->out, gx))) return rc;
    if ((rc = ens_fetch3(e, a->out + 3 * cap, gv))) return rc;
    return gm0 ? ens_get_gm0(e, a, 1, gm0) : PN1_OK;
}

int pn1_ens_adjoint_forward(pn1_ens* e, int steps, int interval)
{
    int rc;
    if (!ens_ready(e, &rc)) return rc;
    if (steps < 1 || interval < 1)
        return eerr(e, PN1_EINVAL, "steps (%d) and interval (%d) must be >= 1", steps, interval);
    pn1_ens_adj* a;
    if ((rc = ens_adj_get(e, &a))) return rc;
    a->have_fwd = false;                 // as in pn1_adjoint_forward
    const int nck = (steps + interval - 1) / interval;
    const size_t s6 = 6 * (size_t)e->cap;
    if (nck > a->nck_cap) {
        if (a->ckpt) (void)hipFree(a->ckpt);
        a->ckpt = nullptr;
        a->nck_cap = 0;
        if ((rc = ealloc_adj(e, &a->ckpt, (size_t)nck * s6))) return rc;
        a->nck_cap = nck;
    }
    const int seg = std::min(interval, steps);
    if (seg > a->seg_cap) {
        if (a->seg) (void)hipFree(a->seg);
        a->seg = nullptr;
        a->seg_cap = 0;
        if ((rc = ealloc_adj(e, &a->seg, 3 * (size_t)seg * s6))) return rc;
        a->seg_cap = seg;
    }
    EnsFp64Scope fp64(e);
    edrop_graph(e);
    const Phys P = ens_phys(e);
    for (int t = 0; t < steps; ++t) {
        if (t % interval == 0 && (rc = ens_copy6(e, a->ckpt + (size_t)(t / interval) * s6, e->X)))
            return rc;
        if ((rc = ens_stages(e, 3, P))) return rc;
    }
    EHIPCHK(e, hipStreamSynchronize(e->st));
    a->have_fwd = true;
    a->steps = steps;
    a->interval = interval;
    a->gen = e->gen;
    return PN1_OK;
}

int pn1_ens_adjoint_backward(pn1_ens* e, const double* gx, const double* gv, double* gx0,
                             double* gv0, double* gm0)
{
    int rc;
    if (!ens_ready(e, &rc)) return rc;
    if (!gx || !gv || !gx0 || !gv0) return eerr(e, PN1_EINVAL, "null argument");
    pn1_ens_adj* a = e->adj;
    if (!a || !a->have_fwd)
        return eerr(e, PN1_EINVAL, "no pn1_ens_adjoint_forward run to reverse");
    if (a->gen != e->gen)
        return eerr(e, PN1_EINVAL, "state or parameters changed since pn1_ens_adjoint_forward");
    if ((rc = ens_adj_get(e, &a))) return rc;
    EnsFp64Scope fp64(e);
    edrop_graph(e);
    const Phys P = ens_phys(e);
    const double dt = e->dt;
    const size_t s6 = 6 * (size_t)e->cap;
    if ((rc = ens_put6(e, gx, gv, a->L))) return rc;
    EHIPCHK(e, hipMemsetAsync(a->gm0, 0, 2 * (size_t)e->nc * sizeof(double), e->st));
    const int nck = (a->steps + a->interval - 1) / a->interval;
    for (int k = nck - 1; k >= 0; --k) {
        const int t0 = k * a->interval, t1 = std::min(a->steps, t0 + a->interval);
        if ((rc = ens_copy6(e, e->X, a->ckpt + (size_t)k * s6))) return rc;
        for (int t = t0; t < t1; ++t) {
            double* st = a->seg + 3 * (size_t)(t - t0) * s6;
            if ((rc = ens_copy6(e, st, e->X))) return rc;
            if ((rc = ens_stages(e, t + 1 < t1 ? 3 : 2, P))) return rc;
            if ((rc = ens_copy6(e, st + s6, e->H1))) return rc;
            if ((rc = ens_copy6(e, st + 2 * s6, e->H2))) return rc;
        }
        for (int t = t1 - 1; t >= t0; --t) {
            const double* st = a->seg + 3 * (size_t)(t - t0) * s6;
            ens_lin(e, a->mu, dt / 6.0, a->L);
            if ((rc = ens_enqueue_vjp(e, a, st + 2 * s6, a->mu, a->LY2, P))) return rc;
            ens_lin(e, a->mu, dt / 6.0, a->L, dt / 2.0, a->LY2);
            if ((rc = ens_enqueue_vjp(e, a, st + s6, a->mu, a->LY1, P))) return rc;
            ens_lin(e, a->mu, dt / 6.0, a->L, dt / 2.0, a->LY1);
            if ((rc = ens_enqueue_vjp(e, a, st, a->mu, a->T, P))) return rc;
            ens_lin(e, a->L, 1.0, a->L, 1.0, a->LY2, 1.0, a->LY1, 1.0, a->T);
        }
    }
    if ((rc = ens_copy6(e, e->X, a->ckpt))) return rc;      // leave the initial state
    EHIPCHK(e, hipGetLastError());
    const size_t cap = (size_t)e->cap;
    if ((rc = ens_fetch3(e, a->L, gx0))) return rc;
    if ((rc = ens_fetch3(e, a->L + 3 * cap, gv0))) return rc;
    return gm0 ? ens_get_gm0(e, a, 0, gm0) : PN1_OK;
}

}  // extern "C"
