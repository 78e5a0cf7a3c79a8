// pn1_ens.hpp -- the ensemble behind pn1_ens* (private to libpn1), shared by
// pn1_api.hip (force, energy, steps: pn1_ens.inc) and pn1_adj.hip (adjoint).
//
// Device layout: one state of capacity cap (SoA, as pn1_ctx) holding all
// clusters back to back; per cluster its BH first.  Tables built at upload:
//   coff[nc + 1]   first particle of each cluster
//   cm0[nc]        BH mass of each cluster
//   cmmax[nc]      largest star mass of each cluster
//   blk[nb], cblk[nc + 1]   256-star blocks of the O(n) kernels, per cluster
//   items[nitems]  pair work items in 8 per-XCD queues (qoff[9]): whole
//                  clusters go to the least-loaded queue, largest first;
//                  each queue is sorted by decreasing pair count
//   poff[nc]       first partial-sum slot of each cluster
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "pn1.h"
#include "pn1_ens_types.hpp"

struct pn1_ens_adj;   // adjoint buffers (pn1_adj.hip)

struct pn1_ens {
    int dev = -1;
    hipStream_t st = nullptr;
    int cap = 0, capc = 0, n = 0, nc = 0;
    bool have_state = false;
    double eps = 1e-3, c = 1.0, dt = 1.0 / 1024.0;
    int mode = PN1_FP64;
    int L_req = 0, tile_req = 0, variant_req = -1;
    int L = 512, tile = 64;
    double *X = nullptr, *Y = nullptr, *H1 = nullptr, *H2 = nullptr, *m = nullptr, *acc = nullptr;
    double *gblk = nullptr, *gs = nullptr, *bhblk = nullptr, *eblk = nullptr, *eout = nullptr;
    double* cm0 = nullptr;
    double* cmmax = nullptr;      // each cluster's largest star mass
    pn1dev::Thr* thr = nullptr;   // mixed form's per-star pair thresholds
    float* kr = nullptr;          // mixed form's per-star cancellation ratios
    bool acc_ok = false;          // acc, kr hold the last mixed evaluation
    bool eqm = false;             // each cluster's stars have one mass
    double* ctauf = nullptr;      // each cluster's BH-close threshold (PN1_FLAG_ETA)
    // mixed form: stars taken out of the FP32 pass (flags, per-block counts
    // [nb_cap], lists at each cluster's offset, per-cluster numbers) and the
    // FP64 sums of their pairs [NACC + 1][cap]
    int *fl = nullptr, *fblk = nullptr, *flist = nullptr, *fn = nullptr;
    double* xp = nullptr;
    // hard stars (PN1_HARD_K): per-cluster lists, all clusters' (star,
    // cluster) pairs [cap] (their number in counter[NXCD]), FP64 rows
    // [NACC + 1][cap] and the segment sums of the split rows
    // [HCAP][HSEG][NACC + 1]
    int* hlist = nullptr;
    int2* hall = nullptr;
    double* xh = nullptr;
    double* hseg = nullptr;
    int fblk_cap = 0, gblk_cap = 0, bhblk_cap = 0, eblk_cap = 0;
    pn1dev::JRec64* j64 = nullptr;
    pn1dev::JRec32* j32 = nullptr;
    int *coff = nullptr, *cblk = nullptr, *qoff = nullptr;
    pn1dev::EBlk* blk = nullptr;
    pn1dev::EItem* items = nullptr;
    long long* poff = nullptr;
    unsigned int* counter = nullptr;
    int nb = 0, nitems = 0, nb_cap = 0, items_cap = 0;
    long long slots = 0;
    void* part = nullptr;
    size_t part_bytes = 0;
    double* h_stage = nullptr;
    int grid64[2] = {0, 0}, grid32[4] = {0, 0, 0, 0}, gride = 0;   // per kernel variant
    hipGraphExec_t gexec = nullptr;
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    int prof_launches = 0;
    double prof_ms = 0.0;
    pn1_ens_adj* adj = nullptr;   // allocated by the first adjoint call
    unsigned gen = 0;             // bumped when the state or eps, c, dt change
    std::vector<int> off_h;       // host copy of the cluster offsets
    char err[512] = {0};
};

namespace pn1host {

int eerr(pn1_ens* e, int code, const char* fmt, ...);
pn1dev::Phys ens_phys(const pn1_ens* e);
// One force evaluation of every cluster (stage 0: plain, on X; 1..3: the
// stage driver's calls with their state updates), on the ensemble's stream.
int ens_enqueue_force(pn1_ens* e, int stage, const pn1dev::Phys& P);
// Mixed form after an upload: one evaluation with the pair thresholds off.
int ens_ensure_prev_acc(pn1_ens* e, const pn1dev::Phys& P);
int ens_fetch3(pn1_ens* e, const double* dsrc, double* out);
bool ens_ready(pn1_ens* e, int* rc);
void edrop_graph(pn1_ens* e);
void ens_adj_free(pn1_ens* e);   // pn1_adj.hip

}  // namespace pn1host

#define EHIPCHK(e, call)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return eerr((e), PN1_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)
