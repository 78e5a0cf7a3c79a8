// pn1_api.hip -- host side of libpn1: the context API, the three legacy
// shapes of Limit-sticky9 (force, energy, initial_energy), the autotuner and,
// through pn1_ens.inc, the ensemble API.  The adjoint is in pn1_adj.hip, the
// launches of the mixed-form pair kernels in pn1_pair32.hip.
//
// Device layout (FP64 SoA, stride cap, particle 0 = BH):
//   X[6][cap]   state at the start of a step (x, y, z, vx, vy, vz)
//   Y[6][cap]   stage input (the reference's x1, v1)
//   H1, H2      stage states 1 and 2 (the reference's hx[1..2], hv[1..2])
//   m[cap], acc[3][cap]
//   j64[cap] / j32[cap]   per-evaluation j-records (AoS)
//   part        pair partials: FP64 [S][16][cap]; mixed [S][16 + 3][cap] floats
//               (the sums, then the Kahan compensations of the Newtonian sums)
// Device memory is allocated by pn1_create, and grown by pn1_upload,
// pn1_set_tuning and pn1_autotune (the partial sums follow the j split) and
// by the first adjoint call; pn1_force, pn1_energy and pn1_step only launch
// kernels and copy results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "pn1.h"
#include "pn1_ctx.hpp"
#include "pn1_ens.hpp"
#include "pn1_ens_kernels.hpp"
#include "pn1_kernels.hpp"
#include "pn1_pair32.hpp"

using namespace pn1dev;

namespace pn1host {

int seterr(pn1_ctx* c, int code, const char* fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

template <typename T>
int dalloc(pn1_ctx* c, T** p, size_t count)
{
    if (dev_malloc((void**)p, count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return seterr(c, PN1_ENOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
    }
    return PN1_OK;
}

int nblocks(int nst) { return nst > 0 ? (nst + WG - 1) / WG : 1; }

Phys make_phys(const pn1_ctx* c)
{
    Phys P;
    P.eps = c->eps;
    P.eps2 = c->eps * c->eps;
    P.lam = 1.0 / (c->c * c->c);
    P.dt = c->dt;
    P.m0 = c->m0;
    P.ws = 1.0 / std::sqrt(P.eps2);
    P.mmax = c->mmax;
    P.eps2f = (float)P.eps2;
    P.m0f = (float)c->m0;
    P.n = c->n;
    P.cap = c->cap;
    return P;
}

// The mixed form's BH-close threshold for stars up to dmax from the BH (see
// PN1_FLAG_ETA): tauf = eta / (6 u D) with D = 2 dmax, room for the stars to
// move away before the next upload.
double flag_threshold(double dmax)
{
    if (!(dmax > 0.0)) return HUGE_VAL;
    return PN1_FLAG_ETA / (6.0 * 0x1p-24 * 2.0 * dmax);
}

// j split S and grid of the pair kernels for the current n.
void plan(pn1_ctx* c)
{
    const int nst = c->n - 1;
    c->nblk = nblocks(nst);
    c->nIblk = nst > 0 ? (nst + WG - 1) / WG : 0;
    int S = c->jsplit_req;
    if (S <= 0) {
        // Many short rounds of workgroups: the last, partly filled round of
        // a grid costs little when there are tens of rounds (measured on
        // MI355X at n = 131073: the mixed pair kernel takes 13.3 ms at split
        // 64, 14.4 ms at 16 and 22.9 ms at 4; at n = 10241 split 128 is 1.2x
        // faster than 32).  Chunks stay >= 64 records.
        S = 1;
        while (S < 128 && (long)S * c->nIblk < 32768 && (nst + 2 * S - 1) / (2 * S) >= 64) S *= 2;
    }
    if (nst > 0 && S > nst) S = nst;
    if (S < 1) S = 1;
    c->S = S;
    if (S <= 8 && 8 % S == 0) {
        const int per = 8 / S;
        c->pair_grid = 8 * ((c->nIblk + per - 1) / per);
    } else {
        c->pair_grid = S * c->nIblk;
    }
    c->tile = c->tile_req > 0 ? (c->tile_req < 2 ? 2 : c->tile_req) : 64;
}

// Bytes of pair partials a j split of S chunks needs (0: the plan's).
size_t part_need(const pn1_ctx* c, int split = 0)
{
    const size_t cap = (size_t)c->cap, S = (size_t)(split > 0 ? split : c->S);
    size_t need = S * NACC * cap * sizeof(double);
    const size_t needE = S * NACCE * cap * sizeof(double);
    const size_t need32 = S * MROWS * cap * sizeof(float);
    if (needE > need) need = needE;
    if (need32 > need) need = need32;
    return need;
}

// Partial-sum buffer for a j split of S chunks (default: the plan's).  The
// new buffer is allocated before the old one is freed: if the allocation
// is refused, the old buffer stays (ready() then refuses launches of a
// plan it does not cover).
int ensure_part(pn1_ctx* c, int split = 0)
{
    const size_t need = part_need(c, split);
    if (need <= c->part_bytes) return PN1_OK;
    void* p = nullptr;
    if (dev_malloc(&p, need) != hipSuccess)
        return seterr(c, PN1_ENOMEM, "hipMalloc of %zu bytes for pair partials failed", need);
    drop_graph(c);                      // a captured step holds the old buffer
    if (c->part) (void)hipFree(c->part);
    c->part = p;
    c->part_bytes = need;
    return PN1_OK;
}

// Kernel form in use: the requested one, or per mode the faster one
// measured at n = 131073: one record at a time in both modes (the mixed
// kernel then fits 8 waves per SIMD).
int kernel_variant(const pn1_ctx* c)
{
    if (c->variant_req >= 0) return c->mode == PN1_MIXED ? c->variant_req : (c->variant_req & 1);
    return 0;
}

// Whether the mixed pair kernel combines four chunks per workgroup
// (k_pair32q, variant bit 2, splits into a multiple of 4 chunks): a quarter
// of the partial-sum traffic, at ~1.6% more pair-kernel time than one chunk
// per 256-thread workgroup (k_pair32) at n = 131073 on MI355X.
bool quad_pairs(const pn1_ctx* c, int S)
{
    return c->mode == PN1_MIXED && S >= QSUB && S % QSUB == 0 && (kernel_variant(c) & 4);
}

// Partial-sum slots per star of a launch split in S chunks.
int part_slots(const pn1_ctx* c, int S) { return quad_pairs(c, S) ? S / QSUB : S; }

// Workgroups of a pair launch over nib i blocks with `slots` chunks each
// (map_block's arrangement when slots divides 8).
int pair_grid(int slots, int nib)
{
    if (slots <= 8 && 8 % slots == 0) return 8 * ((nib + 8 / slots - 1) / (8 / slots));
    return slots * nib;
}

void drop_graph(pn1_ctx* c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
}

bool ready(pn1_ctx* c, int* rc)
{
    if (!c) {
        *rc = PN1_EINVAL;
        return false;
    }
    if (!c->have_state) {
        *rc = seterr(c, PN1_EINVAL, "no state uploaded");
        return false;
    }
    if (c->part_bytes < part_need(c)) {
        *rc = seterr(c, PN1_ENOMEM, "the pair partials (%zu bytes) do not cover the launch plan "
                     "(%zu bytes): an earlier allocation was refused", c->part_bytes,
                     part_need(c));
        return false;
    }
    if (hipSetDevice(c->dev) != hipSuccess) {
        *rc = seterr(c, PN1_EDEVICE, "hipSetDevice(%d) failed", c->dev);
        return false;
    }
    *rc = PN1_OK;
    return true;
}

// ---- one force evaluation (+ stage update) ------------------------------
template <int MODE, int STAGE>
void launch_fin(pn1_ctx* c, const double* in, const Phys& P)
{
    // the cancellation ratios are refreshed by plain evaluations and the
    // first stage of each step (the pair kernel accumulated their sums)
    float* kout = (MODE == 1 && STAGE <= 1) ? c->kr : nullptr;
    hipLaunchKernelGGL((k_fin<MODE, STAGE>), dim3(c->nblk), dim3(WG), 0, c->st, in, c->X, c->Y,
                       c->H1, c->H2, c->m, (const void*)c->part, part_slots(c, c->S), c->gs,
                       c->acc, c->bhblk,
                       kout, (const int*)c->fn, (const double*)c->xp, (const int*)c->fl,
                       (const double*)c->xh, P, 1, c->n);
    hipLaunchKernelGGL((k_bh<STAGE>), dim3(1), dim3(WG), 0, c->st, c->bhblk, c->nblk, c->gs, c->X,
                       c->Y, c->H1, c->H2, c->acc, P);
}

// The pair kernel over i blocks ib0 .. ib0 + nib - 1 against every star,
// with the j range split in S chunks.  ts (mixed form): also accumulate the
// sums behind the stars' cancellation ratios.
void launch_pair(pn1_ctx* c, const double* in, const Phys& P, int ib0, int nib, int S, bool ts)
{
    const bool mixed = c->mode == PN1_MIXED;
    const dim3 g(pair_grid(S, nib)), blk(WG);
    const int variant = kernel_variant(c);
    if (mixed) {
        launch_pair32(c, in, P, ib0, nib, S, pair32_form(variant, ts, c->eqm));
    } else if (variant & 1) {
        hipLaunchKernelGGL(k_pair64<1>, g, blk, 0, c->st, c->j64, in, (double*)c->part, P, S, nib,
                           ib0);
    } else {
        hipLaunchKernelGGL(k_pair64<0>, g, blk, 0, c->st, c->j64, in, (double*)c->part, P, S, nib,
                           ib0);
    }
}

// k_prep's mixed-form inputs (prev: the thresholds from the previous
// evaluation; false: none, the first pass after an upload).
MixPrep mix_prep(const pn1_ctx* c, const Phys& P, bool prev)
{
    MixPrep M;
    M.j32 = c->j32;
    M.thr = c->thr;
    M.aprev = prev ? c->acc : nullptr;
    M.kprev = prev ? c->kr : nullptr;
    M.mmax = c->mmax;
    M.tauf = c->tauf;
    M.lam = P.lam;
    M.hard_rows = 1;
    return M;
}

MixPrep fp64_prep()
{
    MixPrep M;
    M.j32 = nullptr;
    M.thr = nullptr;
    M.aprev = nullptr;
    M.kprev = nullptr;
    M.mmax = M.tauf = M.lam = 0.0;
    M.hard_rows = 0;
    return M;
}

// The records and separable sums of one evaluation on `in`; mixed form:
// also the list of stars taken out of the FP32 pass and the FP64 sums of
// their pairs.
void launch_prep(pn1_ctx* c, const double* in, const Phys& P)
{
    const bool mixed = c->mode == PN1_MIXED;
    const MixPrep M = mixed ? mix_prep(c, P, c->acc_ok) : fp64_prep();
    hipLaunchKernelGGL(k_prep, dim3(c->nblk), dim3(WG), 0, c->st, in, c->m, P, c->j64, c->gblk, M,
                       c->fl, c->fblk);
    hipLaunchKernelGGL((k_rows_sum<NGS>), dim3(1), dim3(WG), 0, c->st, c->gblk, c->nblk, c->gs);
    if (mixed) {
        hipLaunchKernelGGL(k_flist, dim3(1), dim3(WG), 0, c->st, (const int*)c->fl,
                           (const int*)c->fblk, c->nblk, c->n, c->flist, c->hlist, c->fn);
        hipLaunchKernelGGL(k_pair_fl, dim3(c->nblk), dim3(WG), 0, c->st, (const JRec64*)c->j64,
                           in, (const int*)c->flist, (const int*)c->fn, c->xp, P);
        hipLaunchKernelGGL(k_hard_rows, dim3(c->hgrid), dim3(HWG), 0, c->st,
                           (const JRec64*)c->j64, in, (const int*)c->hlist, (const int*)c->fn,
                           c->xh, c->hseg, P);
        hipLaunchKernelGGL(k_hard_sum, dim3((HCAP * (NACC + 1) + WG - 1) / WG), dim3(WG), 0,
                           c->st, (const int*)c->hlist, (const int*)c->fn,
                           (const double*)c->hseg, c->xh, c->cap);
    }
}

// stage 0: plain evaluation on X; stages 1..3: the driver's force calls.
int enqueue_force(pn1_ctx* c, int stage, const Phys& P)
{
    const double* in = stage <= 1 ? c->X : c->Y;
    const bool mixed = c->mode == PN1_MIXED;
    launch_prep(c, in, P);
    if (c->n > 1) {
        hipEvent_t e1 = nullptr;
        if (c->prof) HIPCHK(c, prof_open(c->ev, c->st, &e1));
        launch_pair(c, in, P, 0, c->nIblk, c->S, stage <= 1);
        if (e1) HIPCHK(c, hipEventRecord(e1, c->st));
    }
    switch (stage * 2 + (mixed ? 1 : 0)) {
    case 0: launch_fin<0, 0>(c, in, P); break;
    case 1: launch_fin<1, 0>(c, in, P); break;
    case 2: launch_fin<0, 1>(c, in, P); break;
    case 3: launch_fin<1, 1>(c, in, P); break;
    case 4: launch_fin<0, 2>(c, in, P); break;
    case 5: launch_fin<1, 2>(c, in, P); break;
    case 6: launch_fin<0, 3>(c, in, P); break;
    default: launch_fin<1, 3>(c, in, P); break;
    }
    HIPCHK(c, hipGetLastError());
    c->acc_ok = mixed;
    return PN1_OK;
}

int ensure_prev_acc(pn1_ctx* c, const Phys& P)
{
    if (c->mode != PN1_MIXED || c->acc_ok) return PN1_OK;
    const bool prof = c->prof;      // not one of the caller's timed launches
    c->prof = false;
    const int rc = enqueue_force(c, 0, P);
    c->prof = prof;
    return rc;
}

int enqueue_step(pn1_ctx* c, const Phys& P)
{
    for (int stage = 1; stage <= 3; ++stage) {
        int rc = enqueue_force(c, stage, P);
        if (rc) return rc;
    }
    return PN1_OK;
}

// SoA device [k][cap] -> AoS host [n][3] for k in 0..2, through the stage.
int fetch3(pn1_ctx* c, const double* dsrc, double* out)
{
    const size_t cap = (size_t)c->cap;
    for (int k = 0; k < 3; ++k)
        HIPCHK(c, hipMemcpyAsync(c->h_stage + k * cap, dsrc + k * cap, c->n * sizeof(double),
                                 hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int i = 0; i < c->n; ++i)
        for (int k = 0; k < 3; ++k) out[3 * i + k] = c->h_stage[k * cap + i];
    return PN1_OK;
}

// ---- scratch of pn1_force_noalloc -----------------------------------------
// Sized for n <= PN1_NOALLOC_MAX_N and a j split of at most NA_MAXS chunks:
// j-records, per-block separable sums, their totals, pair partials and the
// per-block BH partials of one FP64 force evaluation, each part 256-byte
// aligned (offsets in doubles).
constexpr int NA_MAXS = 32;
constexpr int NA_MAXB = (PN1_NOALLOC_MAX_N + WG - 1) / WG;
constexpr size_t na_round(size_t d) { return (d + 31) & ~(size_t)31; }
constexpr size_t NA_J64 = 0;
constexpr size_t NA_PART =
    NA_J64 + na_round((size_t)PN1_NOALLOC_MAX_N * (sizeof(JRec64) / sizeof(double)));
constexpr size_t NA_GBLK = NA_PART + na_round((size_t)NA_MAXS * NACC * PN1_NOALLOC_MAX_N);
constexpr size_t NA_GS = NA_GBLK + na_round((size_t)NA_MAXB * NGS);
constexpr size_t NA_BH = NA_GS + na_round(NGS);
constexpr size_t NA_TOTAL = NA_BH + na_round((size_t)NA_MAXB * 3);
static_assert(sizeof(JRec64) % sizeof(double) == 0, "JRec64 in doubles");

// j split of pn1_force_noalloc: NA_MAXS chunks, halved while a chunk would
// hold fewer than 64 stars.
int na_split(int nst)
{
    int S = NA_MAXS;
    while (S > 1 && (nst + S - 1) / S < 64) S >>= 1;
    return S;
}

}  // namespace pn1host

using namespace pn1host;

// One per GPU: the code object's __device__ array, placed on each device the
// library runs on when the code object is loaded there, and held until the
// process ends.  Host code never passes its name to a kernel or a copy (on
// the host the name is only a handle): it uses the device address that
// hipGetSymbolAddress returns for the current device (na_scratch).
static __device__ __align__(256) double scratch[NA_TOTAL];

namespace pn1host {

constexpr int NA_MAXDEV = 64;
double* g_na_ptr[NA_MAXDEV];
std::mutex g_na_mu;

// Device address of the current GPU's scratch, looked up once per device.
int na_scratch(int dev, double** out)
{
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (dev < 0 || dev >= NA_MAXDEV) return PN1_ENODEV;
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PN1_EDEVICE;
        g_na_ptr[dev] = (double*)p;
    }
    *out = g_na_ptr[dev];
    return PN1_OK;
}

}  // namespace pn1host

// =========================================================================
// context API
// =========================================================================
extern "C" {

const char* pn1_last_error(const pn1_ctx* c) { return c ? c->err : "null context"; }

void pn1_destroy(pn1_ctx* c)
{
    if (!c) return;
    if (c->dev >= 0) (void)hipSetDevice(c->dev);
    if (c->st) (void)hipStreamSynchronize(c->st);
    drop_graph(c);
    adj_free(c);
    for (auto& e : c->ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    double* bufs[] = {c->X, c->Y, c->H1, c->H2, c->m, c->acc, c->gblk, c->gs, c->bhblk, c->eblk,
                      c->eout};
    for (double* b : bufs)
        if (b) (void)hipFree(b);
    if (c->j64) (void)hipFree(c->j64);
    if (c->j32) (void)hipFree(c->j32);
    if (c->thr) (void)hipFree(c->thr);
    if (c->kr) (void)hipFree(c->kr);
    void* fbufs[] = {c->fl, c->fblk, c->flist, c->fn, c->xp, c->hlist, c->xh, c->hseg};
    for (void* b : fbufs)
        if (b) (void)hipFree(b);
    if (c->part) (void)hipFree(c->part);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

int pn1_create(pn1_ctx** out, int capacity, int device)
{
    if (!out) return PN1_EINVAL;
    *out = nullptr;
    if (capacity < 1 || capacity > PN1_MAX_CAPACITY) return PN1_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PN1_ENODEV;
    if (device < 0 || device >= ndev) return PN1_ENODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return PN1_ENODEV;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        std::fprintf(stderr, "libpn1: device %d is %s; this library is built for gfx950 only\n",
                     device, prop.gcnArchName);
        return PN1_ENODEV;
    }
    pn1_ctx* c = new pn1_ctx;
    c->dev = device;
    c->cap = capacity;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        pn1_destroy(c);
        return PN1_EDEVICE;
    }
    const size_t cap = (size_t)capacity;
    const size_t nb = (size_t)nblocks(capacity) + 1;
    int rc = PN1_OK;
    if (!rc) rc = dalloc(c, &c->X, 6 * cap);
    if (!rc) rc = dalloc(c, &c->Y, 6 * cap);
    if (!rc) rc = dalloc(c, &c->H1, 6 * cap);
    if (!rc) rc = dalloc(c, &c->H2, 6 * cap);
    if (!rc) rc = dalloc(c, &c->m, cap);
    if (!rc) rc = dalloc(c, &c->acc, 3 * cap);
    if (!rc) rc = dalloc(c, &c->gblk, nb * NGS);
    if (!rc) rc = dalloc(c, &c->gs, 16);
    if (!rc) rc = dalloc(c, &c->bhblk, nb * 3);
    if (!rc) rc = dalloc(c, &c->eblk, nb * NEB);
    if (!rc) rc = dalloc(c, &c->eout, 8);
    // two records more than particles: the pair loops fetch up to jr[j + 2]
    if (!rc) rc = dalloc(c, &c->j64, cap + 2);
    if (!rc) rc = dalloc(c, &c->j32, cap + 2);
    if (!rc) rc = dalloc(c, &c->thr, cap);
    if (!rc) rc = dalloc(c, &c->kr, cap);
    if (!rc) rc = dalloc(c, &c->fl, cap);
    if (!rc) rc = dalloc(c, &c->fblk, 2 * nb);
    if (!rc) rc = dalloc(c, &c->flist, cap);
    if (!rc) rc = dalloc(c, &c->hlist, cap);
    if (!rc) rc = dalloc(c, &c->xh, (size_t)(NACC + 1) * cap);
    if (!rc) rc = dalloc(c, &c->hseg, (size_t)HCAP * HSEG * (NACC + 1));
    if (!rc) rc = dalloc(c, &c->fn, 4);
    if (!rc) rc = dalloc(c, &c->xp, (size_t)(NACC + 1) * cap);
    if (!rc && host_malloc((void**)&c->h_stage, 7 * cap * sizeof(double)) != hipSuccess) {
        c->h_stage = nullptr;
        rc = PN1_ENOMEM;
    }
    if (rc) {
        pn1_destroy(c);
        return rc;
    }
    *out = c;
    return PN1_OK;
}

int pn1_set_params(pn1_ctx* c, double eps, double cl, double dt, int mode)
{
    if (!c) return PN1_EINVAL;
    if (!(eps > 0.0) || !std::isfinite(eps)) return seterr(c, PN1_EINVAL, "eps must be > 0");
    if (!(cl > 0.0) || !std::isfinite(cl)) return seterr(c, PN1_EINVAL, "c must be > 0");
    if (!(dt > 0.0) || !std::isfinite(dt)) return seterr(c, PN1_EINVAL, "dt must be > 0");
    if (mode != PN1_FP64 && mode != PN1_MIXED) return seterr(c, PN1_EINVAL, "unknown mode %d", mode);
    if (eps != c->eps || cl != c->c || dt != c->dt || mode != c->mode) drop_graph(c);
    if (eps != c->eps || cl != c->c || dt != c->dt) ++c->gen;
    // the mixed form's pair thresholds come from the last evaluation's
    // accelerations, which depend on eps and c: take them afresh
    if (eps != c->eps || cl != c->c) c->acc_ok = false;
    c->eps = eps;
    c->c = cl;
    c->dt = dt;
    c->mode = mode;
    return PN1_OK;
}

int pn1_set_tuning(pn1_ctx* c, int jsplit, int tile, int variant)
{
    if (!c) return PN1_EINVAL;
    if (jsplit < 0 || jsplit > 4096 || tile < 0 || tile > 65536 || variant < -1 || variant > 7)
        return seterr(c, PN1_EINVAL, "bad tuning (jsplit %d, tile %d, variant %d)", jsplit, tile,
                      variant);
    const int keep_S = c->jsplit_req, keep_tile = c->tile_req, keep_var = c->variant_req;
    c->jsplit_req = jsplit;
    c->tile_req = tile;
    c->variant_req = variant;
    drop_graph(c);
    if (!c->have_state) return PN1_OK;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->st));
    plan(c);
    const int rc = ensure_part(c);
    if (rc) {                           // refused: the shape in use before stays
        c->jsplit_req = keep_S;
        c->tile_req = keep_tile;
        c->variant_req = keep_var;
        plan(c);
    }
    return rc;
}

int pn1_get_tuning(const pn1_ctx* c, int* jsplit, int* tile, int* variant)
{
    if (!c) return PN1_EINVAL;
    if (jsplit) *jsplit = c->S;
    if (tile) *tile = c->tile;
    if (variant) *variant = kernel_variant(c);
    return PN1_OK;
}

int pn1_upload(pn1_ctx* c, const double* x, const double* v, const double* m, int n)
{
    if (!c) return PN1_EINVAL;
    char msg[256];
    int rc = pn1_check_input(x, v, m, n, c->eps, msg, sizeof msg);
    if (rc) return seterr(c, rc, "%s", msg);
    if (n > c->cap) return seterr(c, PN1_EINVAL, "n = %d exceeds capacity %d", n, c->cap);
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const size_t cap = (size_t)c->cap;
    double* h = c->h_stage;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            h[k * cap + i] = x[3 * i + k];
            h[(3 + k) * cap + i] = v[3 * i + k];
        }
        h[6 * cap + i] = m[i];
    }
    for (int k = 0; k < 6; ++k)
        HIPCHK(c, hipMemcpyAsync(c->X + k * cap, h + k * cap, n * sizeof(double),
                                 hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->m, h + 6 * cap, n * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    // a captured step holds this state's launch constants (n, the BH mass,
    // the mixed form's mass scale, equal-mass kernels, BH-close threshold)
    drop_graph(c);
    ++c->gen;
    c->n = n;
    c->m0 = m[0];
    // mass scale of the mixed records: the largest |m_j| of the stars (0 if
    // every star is massless: their records then have mass 0, and no pair
    // needs a threshold); the equal-mass kernels take unit record masses,
    // so only for stars of one positive mass
    c->mmax = 0.0;
    c->eqm = n > 1 && m[1] > 0.0;
    double dmax2 = 0.0;
    for (int i = 1; i < n; ++i) {
        c->mmax = std::max(c->mmax, std::fabs(m[i]));
        c->eqm = c->eqm && m[i] == m[1];
        double d2 = 0.0;
        for (int k = 0; k < 3; ++k) d2 += (x[3 * i + k] - x[k]) * (x[3 * i + k] - x[k]);
        dmax2 = std::max(dmax2, d2);
    }
    c->tauf = flag_threshold(std::sqrt(dmax2));
    c->acc_ok = false;
    c->have_state = true;
    plan(c);
    return ensure_part(c);
}

int pn1_download(pn1_ctx* c, double* x, double* v)
{
    int rc;
    if (!ready(c, &rc)) return rc;
    if (!x || !v) return seterr(c, PN1_EINVAL, "null output");
    const size_t cap = (size_t)c->cap;
    if ((rc = fetch3(c, c->X, x))) return rc;
    return fetch3(c, c->X + 3 * cap, v);
}

int pn1_force(pn1_ctx* c, double* a)
{
    int rc;
    if (!ready(c, &rc)) return rc;
    if (!a) return seterr(c, PN1_EINVAL, "null output");
    const Phys P = make_phys(c);
    if ((rc = ensure_prev_acc(c, P))) return rc;
    if ((rc = enqueue_force(c, 0, P))) return rc;
    return fetch3(c, c->acc, a);
}

int pn1_force_range(pn1_ctx* c, int i0, int i1, double* a)
{
    int rc;
    if (!ready(c, &rc)) return rc;
    if (!a) return seterr(c, PN1_EINVAL, "null output");
    const int n = c->n;
    if (i0 < 1 || i1 > n || i0 >= i1 || (i0 - 1) % WG != 0 || ((i1 - 1) % WG != 0 && i1 != n))
        return seterr(c, PN1_EINVAL,
                      "star range [%d, %d) must lie in [1, %d) and start and end on %d-star "
                      "block boundaries (or end at n)", i0, i1, n, WG);
    const Phys P = make_phys(c);
    const bool mixed = c->mode == PN1_MIXED;
    const double* in = c->X;
    const int nib = (i1 - i0 + WG - 1) / WG;
    // The j split: an explicit one (pn1_set_tuning, pn1_autotune) as set,
    // which keeps the rows bitwise equal to pn1_force's; otherwise the
    // shard's own plan, the full launch's split doubled while the shard's
    // grid has fewer than 16384 workgroups (chunks of >= 64 records, at most
    // 512 chunks).
    int S = c->S;
    if (c->jsplit_req <= 0) {
        const int nst = n - 1;
        while (S < 512 && (long)S * nib < 16384 && (nst + 2 * S - 1) / (2 * S) >= 64) S *= 2;
    }
    if ((rc = ensure_prev_acc(c, P))) return rc;
    if ((rc = ensure_part(c, S))) return rc;
    // every star's j-record and the separable sums: what a sharded run would
    // all-gather and all-reduce before its pair kernel
    launch_prep(c, in, P);
    hipEvent_t e1 = nullptr;
    if (c->prof) HIPCHK(c, prof_open(c->ev, c->st, &e1));
    launch_pair(c, in, P, (i0 - 1) / WG, nib, S, true);
    if (e1) HIPCHK(c, hipEventRecord(e1, c->st));
    const dim3 gf(nib);
    if (mixed)
        hipLaunchKernelGGL((k_fin<1, 0>), gf, dim3(WG), 0, c->st, in, c->X, c->Y, c->H1, c->H2,
                           c->m, (const void*)c->part, part_slots(c, S), c->gs, c->acc, c->bhblk, c->kr,
                           (const int*)c->fn, (const double*)c->xp, (const int*)c->fl,
                           (const double*)c->xh, P, i0, i1);
    else
        hipLaunchKernelGGL((k_fin<0, 0>), gf, dim3(WG), 0, c->st, in, c->X, c->Y, c->H1, c->H2,
                           c->m, (const void*)c->part, part_slots(c, S), c->gs, c->acc, c->bhblk, c->kr,
                           (const int*)c->fn, (const double*)c->xp, (const int*)c->fl,
                           (const double*)c->xh, P, i0, i1);
    if (!mixed) c->acc_ok = false;
    HIPCHK(c, hipGetLastError());
    const size_t cap = (size_t)c->cap, cnt = (size_t)(i1 - i0);
    for (int k = 0; k < 3; ++k)
        HIPCHK(c, hipMemcpyAsync(c->h_stage + k * cap, c->acc + k * cap + i0,
                                 cnt * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (size_t i = 0; i < cnt; ++i)
        for (int k = 0; k < 3; ++k) a[3 * i + k] = c->h_stage[k * cap + i];
    return PN1_OK;
}

int pn1_energy(pn1_ctx* c, pn1_energy_t* out)
{
    int rc;
    if (!ready(c, &rc)) return rc;
    if (!out) return seterr(c, PN1_EINVAL, "null output");
    const Phys P = make_phys(c);
    hipLaunchKernelGGL(k_prep, dim3(c->nblk), dim3(WG), 0, c->st, c->X, c->m, P, c->j64,
                       c->gblk, fp64_prep(), c->fl, c->fblk);
    hipLaunchKernelGGL((k_rows_sum<NGS>), dim3(1), dim3(WG), 0, c->st, c->gblk, c->nblk, c->gs);
    if (c->n > 1)
        hipLaunchKernelGGL(k_epair64, dim3(c->pair_grid), dim3(WG), 0, c->st, c->j64, c->X,
                           (double*)c->part, P, c->S, c->nIblk);
    hipLaunchKernelGGL(k_efin, dim3(c->nblk), dim3(WG), 0, c->st, c->X, c->m,
                       (const double*)c->part, c->n > 1 ? c->S : 0, c->gs, c->eblk, P);
    hipLaunchKernelGGL(k_ered, dim3(1), dim3(WG), 0, c->st, c->eblk, c->nblk, c->X, c->m, c->gs,
                       c->eout, P);
    HIPCHK(c, hipGetLastError());
    double e[5];
    HIPCHK(c, hipMemcpyAsync(e, c->eout, sizeof e, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    out->kin = e[0];
    out->pot = e[1];
    out->pot_pn = e[2];
    out->pot_pn2 = e[3];
    out->total = e[4];
    return PN1_OK;
}

int pn1_step(pn1_ctx* c, int steps)
{
    int rc;
    if (!ready(c, &rc)) return rc;
    if (steps < 0) return seterr(c, PN1_EINVAL, "steps < 0");
    if (steps == 0) return PN1_OK;
    const Phys P = make_phys(c);
    if ((rc = ensure_prev_acc(c, P))) return rc;
    if (c->prof) {
        for (int s = 0; s < steps; ++s)
            if ((rc = enqueue_step(c, P))) return rc;
        return PN1_OK;
    }
    if (!c->gexec) {
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        rc = enqueue_step(c, P);
        hipError_t ec = hipStreamEndCapture(c->st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess)
            return seterr(c, PN1_EDEVICE, "hipStreamEndCapture failed: %s", hipGetErrorString(ec));
        hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            c->gexec = nullptr;
            return seterr(c, PN1_EDEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
        }
    }
    for (int s = 0; s < steps; ++s) HIPCHK(c, hipGraphLaunch(c->gexec, c->st));
    return PN1_OK;
}

int pn1_sync(pn1_ctx* c)
{
    if (!c) return PN1_EINVAL;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PN1_OK;
}

int pn1_profile_enable(pn1_ctx* c, int on)
{
    if (!c) return PN1_EINVAL;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (auto& e : c->ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    c->ev.clear();
    c->prof = on != 0;
    c->prof_launches = 0;
    c->prof_ms = 0.0;
    return PN1_OK;
}

int pn1_profile_read(pn1_ctx* c, int* launches, double* total_ms)
{
    if (!c) return PN1_EINVAL;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (auto& e : c->ev) {
        float ms = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&ms, e.first, e.second));
        c->prof_ms += ms;
        c->prof_launches += 1;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    c->ev.clear();
    if (launches) *launches = c->prof_launches;
    if (total_ms) *total_ms = c->prof_ms;
    return PN1_OK;
}

// Autotuner: times the pair kernel of `reps` force evaluations on the
// resident state per candidate shape (j split x kernel variant; the mixed
// form's compensation interval is left alone: it trades accuracy), after
// one warm-up evaluation, and keeps the fastest.
int pn1_autotune(pn1_ctx* c, int reps, int* jsplit, int* variant, double* best_ms)
{
    int rc;
    if (!ready(c, &rc)) return rc;
    if (reps < 1) return seterr(c, PN1_EINVAL, "reps must be >= 1");
    const int nst = c->n - 1;
    const bool was_prof = c->prof;
    const int keep_tile = c->tile_req, keep_S = c->jsplit_req, keep_var = c->variant_req;
    std::vector<int> splits = {0};                  // 0 = the automatic plan
    for (int S = 8; S <= 256; S *= 2)
        if (S <= std::max(1, nst / 64)) splits.push_back(S);
    const Phys P = make_phys(c);
    double best = 1e300;
    int bS = 0, bV = -1;
    // whole force evaluations (every kernel: the partial-sum reduction in
    // k_fin grows with the split), timed with events on the stream
    hipEvent_t ea = nullptr, eb = nullptr;
    if (hipEventCreate(&ea) != hipSuccess || hipEventCreate(&eb) != hipSuccess) {
        if (ea) (void)hipEventDestroy(ea);
        return seterr(c, PN1_EDEVICE, "hipEventCreate failed");
    }
    if ((rc = pn1_profile_enable(c, 0))) {
        (void)hipEventDestroy(ea);
        (void)hipEventDestroy(eb);
        return rc;
    }
    auto time_shape = [&](int S, int var, double* avg) {
        c->jsplit_req = S;
        c->variant_req = var;
        c->tile_req = keep_tile;
        plan(c);
        int r = ensure_part(c);
        if (!r) r = enqueue_force(c, 0, P);                      // warm-up
        if (!r && hipEventRecord(ea, c->st) != hipSuccess) r = seterr(c, PN1_EDEVICE, "hipEventRecord failed");
        for (int k = 0; k < reps && !r; ++k) r = enqueue_force(c, 0, P);
        if (!r && hipEventRecord(eb, c->st) != hipSuccess) r = seterr(c, PN1_EDEVICE, "hipEventRecord failed");
        float ms = 0.0f;
        if (!r && (hipEventSynchronize(eb) != hipSuccess || hipEventElapsedTime(&ms, ea, eb) != hipSuccess))
            r = seterr(c, PN1_EDEVICE, "timing the force evaluations failed");
        *avg = (double)ms / reps;
        return r;
    };
    for (int S : splits) {
        for (int var = 0; var < 2 && !rc; ++var) {
            double avg = 0.0;
            rc = time_shape(S, var, &avg);
            if (!rc && avg < best) {
                best = avg;
                bS = c->S;
                bV = var;
            }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    (void)pn1_profile_enable(c, was_prof ? 1 : 0);
    if (rc) {                              // leave the shape the caller had
        c->jsplit_req = keep_S;
        c->variant_req = keep_var;
        plan(c);
        (void)ensure_part(c);
        return rc;
    }
    c->jsplit_req = bS;
    c->variant_req = bV;
    drop_graph(c);
    plan(c);
    if ((rc = ensure_part(c))) return rc;
    if (jsplit) *jsplit = c->S;
    if (variant) *variant = kernel_variant(c);
    if (best_ms) *best_ms = best;
    return PN1_OK;
}

// =========================================================================
// no-allocation force
// =========================================================================
// The FP64 force of pn1_force on a state the caller holds in device memory,
// with every intermediate in this GPU's scratch: the same kernels, the j
// split na_split(n - 1).  Enqueues only (no hipMalloc, hipFree, memset or
// synchronisation); nothing in the scratch needs clearing, every value read
// is written earlier in the same call.
int pn1_force_noalloc(const double* st, const double* m, int n, double eps, double c,
                      double* acc, void* stream)
{
    if (!st || !m || !acc || n < 1 || n > PN1_NOALLOC_MAX_N) return PN1_EINVAL;
    if (!(eps > 0.0) || !std::isfinite(eps) || !(c > 0.0) || !std::isfinite(c))
        return PN1_EINVAL;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return PN1_EDEVICE;
    double* base = nullptr;
    const int rc = na_scratch(dev, &base);
    if (rc) return rc;
    const hipStream_t s = (hipStream_t)stream;
    Phys P;
    P.eps = eps;
    P.eps2 = eps * eps;
    P.lam = 1.0 / (c * c);
    P.dt = 0.0;
    P.m0 = 0.0;                 // the kernels read the BH mass from m[0]
    P.ws = 1.0 / std::sqrt(P.eps2);
    P.mmax = 0.0;
    P.eps2f = (float)P.eps2;
    P.m0f = 0.0f;
    P.n = n;
    P.cap = n;
    JRec64* j64 = (JRec64*)(base + NA_J64);
    double* part = base + NA_PART;
    double* gblk = base + NA_GBLK;
    double* gs = base + NA_GS;
    double* bhblk = base + NA_BH;
    const int nst = n - 1, nblk = nblocks(nst), nib = (nst + WG - 1) / WG;
    const int S = na_split(nst);
    hipLaunchKernelGGL(k_prep, dim3(nblk), dim3(WG), 0, s, st, m, P, j64, gblk, fp64_prep(),
                       (int*)nullptr, (int*)nullptr);
    hipLaunchKernelGGL((k_rows_sum<NGS>), dim3(1), dim3(WG), 0, s, (const double*)gblk, nblk, gs);
    if (nst > 0) {
        int grid = S * nib;
        if (S <= 8 && 8 % S == 0) grid = 8 * ((nib + 8 / S - 1) / (8 / S));
        hipLaunchKernelGGL(k_pair64<0>, dim3(grid), dim3(WG), 0, s, (const JRec64*)j64, st, part,
                           P, S, nib, 0);
    }
    hipLaunchKernelGGL((k_fin<0, 0>), dim3(nblk), dim3(WG), 0, s, st, (double*)nullptr,
                       (double*)nullptr, (double*)nullptr, (double*)nullptr, m,
                       (const void*)part, S, (const double*)gs, acc, bhblk, (float*)nullptr,
                       (const int*)nullptr, (const double*)nullptr, (const int*)nullptr,
                       (const double*)nullptr, P, 1, n);
    hipLaunchKernelGGL((k_bh<0>), dim3(1), dim3(WG), 0, s, (const double*)bhblk, nblk,
                       (const double*)gs, (double*)nullptr, (double*)nullptr, (double*)nullptr,
                       (double*)nullptr, acc, P);
    return hipGetLastError() == hipSuccess ? PN1_OK : PN1_EDEVICE;
}

// =========================================================================
// legacy shapes
// =========================================================================
static pn1_ctx* g_ctx = nullptr;
static double g_c = 1.0;
static int g_mode = PN1_FP64;

int pn1_set_c(double cl)
{
    if (!(cl > 0.0) || !std::isfinite(cl)) return PN1_EINVAL;
    g_c = cl;
    return PN1_OK;
}

int pn1_set_mode(int mode)
{
    if (mode != PN1_FP64 && mode != PN1_MIXED) return PN1_EINVAL;
    g_mode = mode;
    return PN1_OK;
}

static void die(const char* what, const char* msg)
{
    std::fprintf(stderr, "%s: %s\n", what, msg);
    std::exit(1);
}

static pn1_ctx* legacy_ctx(const char* what, const double* x, const double* v, const double* m,
                           int n, double eps, int mode)
{
    char msg[256];
    if (pn1_check_input(x, v, m, n, eps, msg, sizeof msg) != PN1_OK) die(what, msg);
    if (!g_ctx || g_ctx->cap < n) {
        pn1_destroy(g_ctx);
        g_ctx = nullptr;
        int cap = n < 1024 ? 1024 : n;
        int rc = pn1_create(&g_ctx, cap, 0);
        if (rc != PN1_OK) {
            std::snprintf(msg, sizeof msg, "cannot create a gfx950 device context (status %d)", rc);
            die(what, msg);
        }
    }
    if (pn1_set_params(g_ctx, eps, g_c, g_ctx->dt, mode) != PN1_OK ||
        pn1_upload(g_ctx, x, v, m, n) != PN1_OK)
        die(what, g_ctx->err);
    return g_ctx;
}

void force(double (*x)[3], double (*v)[3], double* m, double eps, double (*a)[3], double* pot, int n)
{
    (void)pot;  // ignored, as in Limit-sticky9.c:21-424
    pn1_ctx* c = legacy_ctx("force", &x[0][0], &v[0][0], m, n, eps, g_mode);
    if (pn1_force(c, &a[0][0]) != PN1_OK) die("force", c->err);
}

static void legacy_energy(const char* what, double (*x)[3], double (*v)[3], double* m, int n,
                          double eps, pn1_energy_t* e)
{
    pn1_ctx* c = legacy_ctx(what, &x[0][0], &v[0][0], m, n, eps, PN1_FP64);
    if (pn1_energy(c, e) != PN1_OK) die(what, c->err);
}

// The print block of Limit-sticky9.c:560-563 / 701-703.
static void print_block(double t, const pn1_energy_t* e)
{
    std::printf("time = %g\n", t);
    std::printf("pot = %22.15e kin = %22.15e \n pot_pn = %22.15e \n  total= %22.15e ratio = %e\n",
                e->pot, e->kin, e->pot_pn, e->total, e->kin / e->pot);
}

void energy(double t, double (*x)[3], double (*v)[3], double* m, int n, double init_ene, double eps)
{
    pn1_energy_t e;
    legacy_energy("energy", x, v, m, n, eps, &e);
    const double de = (init_ene - e.total) / e.total;
    print_block(t, &e);
    std::printf(" DE = %e %g\n", de, t);
}

void initial_energy(double (*x)[3], double (*v)[3], double* m, int n, double* init_ene, double eps)
{
    pn1_energy_t e;
    legacy_energy("initial_energy", x, v, m, n, eps, &e);
    print_block(0.0, &e);
    *init_ene = e.total;
}

}  // extern "C"

#include "pn1_ens.inc"
