// pn1_ctx.hpp -- the context behind pn1_ctx* (private to libpn1), shared by
// the translation units of the library: pn1_api.hip (force, energy, stage
// driver, legacy shapes, ensembles) and pn1_adj.hip (adjoint).
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "pn1.h"
#include "pn1_common.hpp"

struct pn1_adj;   // adjoint buffers (pn1_adj.hip)

struct pn1_ctx {
    int dev = -1;
    hipStream_t st = nullptr;
    int cap = 0, n = 0;
    bool have_state = false;
    double eps = 1e-3, c = 1.0, dt = 1.0 / 1024.0;
    int mode = PN1_FP64;
    double m0 = 0.0;            // host copy of m[0]
    double mmax = 0.0;          // largest star mass (mixed form's pair thresholds)
    bool eqm = false;           // every star has the same mass
    double tauf = 0.0;          // BH-close threshold of the mixed form (PN1_FLAG_ETA)
    int jsplit_req = 0, tile_req = 0, variant_req = -1;
    // launch plan for the current n
    int S = 1, nIblk = 0, pair_grid = 0, nblk = 1, tile = 64;
    // device buffers
    double *X = nullptr, *Y = nullptr, *H1 = nullptr, *H2 = nullptr;
    double *m = nullptr, *acc = nullptr, *gblk = nullptr, *gs = nullptr;
    double *bhblk = nullptr, *eblk = nullptr, *eout = nullptr;
    pn1dev::JRec64* j64 = nullptr;
    pn1dev::JRec32* j32 = nullptr;
    pn1dev::Thr* thr = nullptr; // mixed form's per-star pair thresholds
    float* kr = nullptr;        // mixed form's per-star cancellation ratios
    // mixed form: stars taken out of the FP32 pass (flags, per-block counts,
    // their list and number) and the FP64 sums of their pairs
    int *fl = nullptr, *fblk = nullptr, *flist = nullptr, *fn = nullptr;
    double* xp = nullptr;
    // hard stars (PN1_HARD_K): their list, FP64 rows [NACC + 1][cap] and the
    // segment sums of the split rows [HCAP][HSEG][NACC + 1]; hgrid
    // workgroups of k_hard_rows take the work units in turn
    int* hlist = nullptr;
    double* xh = nullptr;
    double* hseg = nullptr;
    int hgrid = 1024;
    // acc and kr hold the last mixed evaluation of the uploaded run (cleared
    // by pn1_upload and by FP64 evaluations): the mixed form's thresholds
    // read them
    bool acc_ok = false;
    void* part = nullptr;
    size_t part_bytes = 0;
    double* h_stage = nullptr;  // pinned, 7 * cap doubles
    hipGraphExec_t gexec = nullptr;
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    int prof_launches = 0;
    double prof_ms = 0.0;
    pn1_adj* adj = nullptr;     // allocated by the first adjoint call
    unsigned gen = 0;           // bumped when the state, masses or parameters change
    char err[512] = {0};
};

namespace pn1host {

int seterr(pn1_ctx* c, int code, const char* fmt, ...);
pn1dev::Phys make_phys(const pn1_ctx* c);
// Launches one force evaluation (stage 0: plain, on X; 1..3: the stage
// driver's calls, with their state updates) on the context's stream.
int enqueue_force(pn1_ctx* c, int stage, const pn1dev::Phys& P);
// Mixed form after an upload: one evaluation with the pair thresholds off
// (no previous accelerations yet), so that the next has them.
int ensure_prev_acc(pn1_ctx* c, const pn1dev::Phys& P);
// SoA device [k][cap] -> AoS host [n][3] for k in 0..2, through the stage.
int fetch3(pn1_ctx* c, const double* dsrc, double* out);
// Checks that a state is resident and makes the context's device current;
// *rc is PN1_OK, or the error when it returns false.
bool ready(pn1_ctx* c, int* rc);
void drop_graph(pn1_ctx* c);
void adj_free(pn1_ctx* c);   // pn1_adj.hip

}  // namespace pn1host

// hipMalloc / hipHostMalloc whose failure is returned as a status only: HIP
// also keeps a failed call's code for the next hipGetLastError(), which the
// launch checks read, so a refused allocation would otherwise fail the next,
// valid launch.  *p is null on failure.
inline hipError_t dev_malloc(void** p, size_t bytes)
{
    const hipError_t r = hipMalloc(p, bytes);
    if (r != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
    }
    return r;
}

inline hipError_t host_malloc(void** p, size_t bytes)
{
    const hipError_t r = hipHostMalloc(p, bytes);
    if (r != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
    }
    return r;
}

// Opens a timed region on stream `st`: two events handed to `ev` at once
// (so a later error leaks neither), the first recorded now; the caller
// records *end after the launches it times.
inline hipError_t prof_open(std::vector<std::pair<hipEvent_t, hipEvent_t>>& ev, hipStream_t st,
                            hipEvent_t* end)
{
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t r = hipEventCreate(&a);
    if (r != hipSuccess) return r;
    if ((r = hipEventCreate(&b)) != hipSuccess) {
        (void)hipEventDestroy(a);
        return r;
    }
    ev.emplace_back(a, b);
    *end = b;
    return hipEventRecord(a, st);
}

#define HIPCHK(ctx, call)                                                                 \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return seterr((ctx), PN1_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

