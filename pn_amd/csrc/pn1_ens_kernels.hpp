// pn1_ens_kernels.hpp -- device code for ensembles of independent clusters.
//
// All clusters live in one state (SoA stride cap).  Cluster c occupies
// particles [coff[c], coff[c + 1]); its BH is particle coff[c], its stars
// follow.  One force evaluation of the whole ensemble is five launches:
//
//   k_ens_prep   per-star j-records and per-block separable sums; blocks
//                come from a block table, each inside one cluster
//   k_ens_gsum   one workgroup per cluster: that cluster's separable sums;
//                also resets the work counters of the pair kernel and the
//                count of the mixed form's hard stars
//   k_ens_pair*  ONE persistent launch for all clusters: workgroups take
//                (cluster, 256-star i-block, j-chunk) items from per-XCD
//                queues (atomic counters), items sorted by decreasing cost
//                on the host; every item writes its partial sums to its own
//                slot, so the result does not depend on which workgroup ran
//                which item
//   k_ens_fin    per star, as k_fin, with its cluster's sums and BH
//   k_ens_bh     one workgroup per cluster: its BH
//
// The pair arithmetic is the single-cluster kernels' (body64 / body32),
// and so is the order of each lane's j loop: with one chunk per cluster
// the results equal a single-cluster run with an unsplit j range bit for
// bit.
//
// Included by pn1_api.hip and pn1_pair32.hip (the mixed-form pair kernels):
// the kernels that are not templates are declared inline.
#pragma once
#include "pn1_ens_types.hpp"
#include "pn1_kernels.hpp"

namespace pn1dev {

// Lane geometry of an item: the lane's local star, its particle index
// (clamped into the cluster for lanes past its end) and the wave's own
// window [w0, w1) of the chunk (masked self pairs).
struct EGeom {
    int li, i, w0, w1;
    bool live;   // the wave holds at least one of the cluster's stars
};

__device__ __forceinline__ EGeom ens_geom(const EItem& w)
{
    EGeom g;
    g.li = w.ib * WG + (int)threadIdx.x;
    g.i = w.off + 1 + (g.li < w.nst ? g.li : w.nst - 1);
    const int wl = w.ib * WG + (__builtin_amdgcn_readfirstlane((int)threadIdx.x) & ~63);
    const int wb = w.off + 1 + wl;
    g.w0 = max(w.j0, wb);
    g.w1 = min(w.j1, wb + 64);
    g.live = wl < w.nst;
    return g;
}

template <int PP>
__device__ __forceinline__ void ens_pair64(const JRec64* __restrict__ jr,
                                           const double* __restrict__ st,
                                           double* __restrict__ part,
                                           const EItem* __restrict__ items,
                                           unsigned int* __restrict__ counter,
                                           const int* __restrict__ qoff, Phys P)
{
    int q = 0;
    for (;;) {
        const int it = ens_take(counter, qoff, q);
        if (it < 0) break;
        const EItem w = items[it];
        const EGeom g = ens_geom(w);
        const I64 I = load_i64(jr, st, P.cap, g.i, P.eps2);
        // a wave past the cluster's last star (a partly filled last i
        // block) skips the loop: it would only repeat that star's sums
        if (!g.live) continue;
        double a[NACC];
#pragma unroll
        for (int k = 0; k < NACC; ++k) a[k] = 0.0;
        if (g.w0 < g.w1) {
            run64<false, PP>(jr, w.j0, g.w0, g.i, I, a, P.eps2);
            run64<true, PP>(jr, g.w0, g.w1, g.i, I, a, P.eps2);
            run64<false, PP>(jr, g.w1, w.j1, g.i, I, a, P.eps2);
        } else {
            run64<false, PP>(jr, w.j0, w.j1, g.i, I, a, P.eps2);
        }
        if (g.li < w.nst) {
            double* pp = part + w.pslot * NACC + g.li;
#pragma unroll
            for (int k = 0; k < NACC; ++k) stream_out(a[k], &pp[(size_t)k * w.nst]);
        }
    }
}

template <int PP, bool TS, bool EQ>
__device__ __forceinline__ void ens_pair32(const JRec32* __restrict__ jr,
                                           const JRec64* __restrict__ j64,
                                           const Thr* __restrict__ thr,
                                           const double* __restrict__ st,
                                           float* __restrict__ part,
                                           const EItem* __restrict__ items,
                                           unsigned int* __restrict__ counter,
                                           const int* __restrict__ qoff,
                                           const double* __restrict__ cmmax, Phys P, int tile)
{
    int q = 0;
    for (;;) {
        const int it = ens_take(counter, qoff, q);
        if (it < 0) break;
        const EItem w = items[it];
        const EGeom g = ens_geom(w);
        const I32 I = load_i32(j64, thr, st, P.cap, g.i, w.off, P.eps2);
        if (!g.live) continue;                 // as in k_ens_pair64
        float a[NACC32], sN[3], cN[3];
#pragma unroll
        for (int k = 0; k < NACC32; ++k) a[k] = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) sN[k] = cN[k] = 0.0f;
        int cnt = 0;
        const float e2 = P.eps2f;
        const double e2d = P.eps2, rs = 1.0 / cmmax[w.c];
        if (g.w0 < g.w1) {
            run32<false, PP, TS, EQ>(jr, j64, w.j0, g.w0, g.i, I, a, e2, e2d, rs, sN, cN, tile, cnt);
            run32<true, PP, TS, EQ>(jr, j64, g.w0, g.w1, g.i, I, a, e2, e2d, rs, sN, cN, tile, cnt);
            run32<false, PP, TS, EQ>(jr, j64, g.w1, w.j1, g.i, I, a, e2, e2d, rs, sN, cN, tile, cnt);
        } else {
            run32<false, PP, TS, EQ>(jr, j64, w.j0, w.j1, g.i, I, a, e2, e2d, rs, sN, cN, tile, cnt);
        }
        fold_newton(a, sN, cN);
        if (g.li < w.nst) {
            float* pp = part + w.pslot * MROWS + g.li;
#pragma unroll
            for (int k = 0; k < NACC32; ++k)
                if (TS || k != A_TS) stream_out(k <= A_NZ ? sN[k] : a[k], &pp[(size_t)k * w.nst]);
#pragma unroll
            for (int k = 0; k < 3; ++k) stream_out(cN[k], &pp[(size_t)(NACC32 + k) * w.nst]);
        }
    }
}

// The kernels.  The persistent loop keeps an item's geometry live beside
// the pair sums, which left to itself the compiler pays for with registers
// (66 to 99 VGPRs, 4 to 7 waves per SIMD against the single-cluster
// kernels' 60 and 72).  The forms with one j-record at a time are held to
// the single-cluster kernels' occupancy: 8 waves per SIMD for the mixed
// kernel (64 VGPRs, no spills), 7 for the FP64 one; the two-in-flight forms
// (variant bit 0) are left to the compiler, which would spill them there.
template <int PP>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(7, 8)))
void k_ens_pair64(const JRec64* __restrict__ jr, const double* __restrict__ st,
                  double* __restrict__ part, const EItem* __restrict__ items,
                  unsigned int* __restrict__ counter, const int* __restrict__ qoff, Phys P)
{
    ens_pair64<PP>(jr, st, part, items, counter, qoff, P);
}

template <int PP>
__global__ __launch_bounds__(WG)
void k_ens_pair64x(const JRec64* __restrict__ jr, const double* __restrict__ st,
                   double* __restrict__ part, const EItem* __restrict__ items,
                   unsigned int* __restrict__ counter, const int* __restrict__ qoff, Phys P)
{
    ens_pair64<PP>(jr, st, part, items, counter, qoff, P);
}

template <int PP, bool TS, bool EQ>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_ens_pair32(const JRec32* __restrict__ jr, const JRec64* __restrict__ j64,
                  const Thr* __restrict__ thr, const double* __restrict__ st,
                  float* __restrict__ part, const EItem* __restrict__ items,
                  unsigned int* __restrict__ counter, const int* __restrict__ qoff,
                  const double* __restrict__ cmmax, Phys P, int tile)
{
    ens_pair32<PP, TS, EQ>(jr, j64, thr, st, part, items, counter, qoff, cmmax, P, tile);
}

template <int PP, bool TS, bool EQ>
__global__ __launch_bounds__(WG)
void k_ens_pair32x(const JRec32* __restrict__ jr, const JRec64* __restrict__ j64,
                   const Thr* __restrict__ thr, const double* __restrict__ st,
                   float* __restrict__ part, const EItem* __restrict__ items,
                   unsigned int* __restrict__ counter, const int* __restrict__ qoff,
                   const double* __restrict__ cmmax, Phys P, int tile)
{
    ens_pair32<PP, TS, EQ>(jr, j64, thr, st, part, items, counter, qoff, cmmax, P, tile);
}

// M: the mixed form's inputs (M.j32 null: FP64), with each cluster's largest
// star mass cmmax[c] and BH-close threshold ctauf[c] in place of M.mmax,
// M.tauf; fl, fblk: as k_prep.
inline __global__ __launch_bounds__(WG) void k_ens_prep(const double* __restrict__ st,
                                                const double* __restrict__ m,
                                                const double* __restrict__ cm0,
                                                const int* __restrict__ coff,
                                                const EBlk* __restrict__ blk, Phys P,
                                                JRec64* __restrict__ j64,
                                                double* __restrict__ gblk, MixPrep M,
                                                const double* __restrict__ cmmax,
                                                const double* __restrict__ ctauf,
                                                int* __restrict__ fl, int* __restrict__ fblk)
{
    const EBlk b = blk[blockIdx.x];
    const int off = coff[b.c], nst = coff[b.c + 1] - off - 1;
    const int li = b.first + (int)threadIdx.x;
    double g[NGS];
#pragma unroll
    for (int k = 0; k < NGS; ++k) g[k] = 0.0;
    if (M.j32) {
        M.mmax = cmmax[b.c];
        M.tauf = ctauf[b.c];
    }
    bool f = false, hard = false;
    if (li < nst) {
        f = prep_star(st, m, P.cap, off + 1 + li, off, cm0[b.c], P.eps2, j64, g, M, hard);
        hard = hard && M.hard_rows;
        if (M.j32) fl[off + 1 + li] = (f ? 1 : 0) | (hard ? 2 : 0);
    }
    block_sum<NGS>(g, gblk + (size_t)blockIdx.x * NGS);
    if (M.j32) {
        const int cnt = __syncthreads_count(f);
        const int hcnt = __syncthreads_count(hard);
        if (threadIdx.x == 0) {
            fblk[2 * blockIdx.x] = cnt;
            fblk[2 * blockIdx.x + 1] = hcnt;
        }
    }
}

// One workgroup per cluster: its BH-close stars, in increasing order, at
// flist[coff[c] ..), their number in fn[c]; its hard stars likewise at
// hlist[coff[c] ..), their number in hn[c], and appended as (star,
// cluster) pairs to the evaluation's list of every cluster's hard stars,
// hall[0 .. *hcount) (k_ens_gsum zeroes *hcount).  The clusters' order in
// hall depends on which workgroup reserves its range first; each hard row
// is computed on its own, so the results do not.
inline __global__ __launch_bounds__(WG) void k_ens_flist(const int* __restrict__ fl,
                                                 const int* __restrict__ fblk,
                                                 const int* __restrict__ cblk,
                                                 const int* __restrict__ coff,
                                                 int* __restrict__ flist, int* __restrict__ fn,
                                                 int* __restrict__ hlist, int* __restrict__ hn,
                                                 int2* __restrict__ hall,
                                                 unsigned int* __restrict__ hcount)
{
    __shared__ int s_at;
    const int c = blockIdx.x;
    list_flagged<0, 2>(fl, fblk, cblk[c], cblk[c + 1], coff[c] + 1, coff[c + 1],
                       flist + coff[c], fn + c);
    const int nh = list_flagged<1, 2>(fl, fblk, cblk[c], cblk[c + 1], coff[c] + 1,
                                      coff[c + 1], hlist + coff[c], hn + c);
    if (nh == 0) return;                     // uniform: every thread holds the count
    if (threadIdx.x == 0) s_at = (int)atomicAdd(hcount, (unsigned int)nh);
    __syncthreads();                         // s_at, and the list written above
    for (int k = (int)threadIdx.x; k < nh; k += WG)
        hall[s_at + k] = make_int2(hlist[coff[c] + k], c);
}

// FP64 rows of every cluster's hard stars (as k_hard_rows, j over the
// star's own cluster): every entry of hall in HSEG units (into hseg), or,
// beyond HCAP entries, every one whole; the workgroups take the units in
// turn.  With at most HCAP hard stars in the ensemble, each row is summed
// as a single-cluster run sums it.
inline __global__ __launch_bounds__(HWG) void k_ens_hard_rows(const JRec64* __restrict__ jr,
                                                      const double* __restrict__ st,
                                                      const int* __restrict__ coff,
                                                      const int2* __restrict__ hall,
                                                      const unsigned int* __restrict__ hcount,
                                                      double* __restrict__ xh,
                                                      double* __restrict__ hseg, Phys P)
{
    const int nh = (int)*hcount, ns = hard_split(nh), nu = ns * HSEG + (nh - ns);
    for (int u = blockIdx.x; u < nu; u += gridDim.x) {
        const int2 w = hall[u < ns * HSEG ? u / HSEG : ns + (u - ns * HSEG)];
        hard_unit(jr, st, P.cap, P.eps2, u, ns, w.x, coff[w.y] + 1, coff[w.y + 1], xh, hseg);
    }
}

// The split rows' segment sums, in segment order (as k_hard_sum).
inline __global__ __launch_bounds__(WG) void k_ens_hard_sum(const int2* __restrict__ hall,
                                                    const unsigned int* __restrict__ hcount,
                                                    const double* __restrict__ hseg,
                                                    double* __restrict__ xh, int cap)
{
    const int t = blockIdx.x * WG + threadIdx.x;
    const int ns = hard_split((int)*hcount);
    if (t < ns * (NACC + 1)) hard_sum(hseg, t, hall[t / (NACC + 1)].x, cap, xh);
}

// Per star block: its stars' FP64 pairs with their cluster's flagged stars
// (k_pair_fl per cluster), into xp [NACC + 1][cap] at the star's index.
inline __global__ __launch_bounds__(WG) void k_ens_pair_fl(const JRec64* __restrict__ jr,
                                                   const double* __restrict__ st,
                                                   const int* __restrict__ coff,
                                                   const EBlk* __restrict__ blk,
                                                   const int* __restrict__ flist,
                                                   const int* __restrict__ fn,
                                                   double* __restrict__ xp, Phys P)
{
    const EBlk b = blk[blockIdx.x];
    const int nf = fn[b.c];
    const int off = coff[b.c], nst = coff[b.c + 1] - off - 1;
    const int li = b.first + (int)threadIdx.x;
    if (nf == 0 || li >= nst) return;
    const int i = off + 1 + li;
    flagged_pairs(jr, st, P.cap, i, P.eps2, flist + off, nf, xp, (size_t)P.cap, i);
}

// One workgroup per cluster: rows [cblk[c], cblk[c+1]) of `rows` (N wide).
template <int N>
__device__ __forceinline__ void cluster_rows_sum(const double* __restrict__ rows,
                                                 const int* __restrict__ cblk, int c,
                                                 double (&g)[N])
{
#pragma unroll
    for (int k = 0; k < N; ++k) g[k] = 0.0;
    for (int b = cblk[c] + (int)threadIdx.x; b < cblk[c + 1]; b += WG)
#pragma unroll
        for (int k = 0; k < N; ++k) g[k] += rows[(size_t)b * N + k];
}

inline __global__ __launch_bounds__(WG) void k_ens_gsum(const double* __restrict__ gblk,
                                                const int* __restrict__ cblk,
                                                double* __restrict__ gs,
                                                unsigned int* __restrict__ counter)
{
    const int c = blockIdx.x;
    if (c == 0 && threadIdx.x <= NXCD) counter[threadIdx.x] = 0u;   // queues, hard stars
    double g[NGS];
    cluster_rows_sum<NGS>(gblk, cblk, c, g);
    block_sum<NGS>(g, gs + (size_t)c * NGS);
}

template <int MODE, int STAGE>
__global__ __launch_bounds__(WG) void k_ens_fin(const double* in, double* X, double* Y,
                                               double* H1, double* H2,
                                               const double* __restrict__ m,
                                               const double* __restrict__ cm0,
                                               const int* __restrict__ coff,
                                               const EBlk* __restrict__ blk,
                                               const void* __restrict__ part,
                                               const long long* __restrict__ poff,
                                               const double* __restrict__ gs,
                                               double* __restrict__ acc,
                                               double* __restrict__ bhblk,
                                               const double* __restrict__ cmmax,
                                               float* __restrict__ kout,
                                               const int* __restrict__ fn,
                                               const double* __restrict__ xp,
                                               const int* __restrict__ fl,
                                               const double* __restrict__ xh, Phys P,
                                               EConst E)
{
    const EBlk b = blk[blockIdx.x];
    const int off = coff[b.c], nst = coff[b.c + 1] - off - 1;
    const int li = b.first + (int)threadIdx.x;
    double cb[3] = {0.0, 0.0, 0.0};
    if (li < nst) {
        double S_[NACC], ts, a[3];
        if (MODE == 1 && (fl[off + 1 + li] & 2)) {
            load_row(xh, (size_t)P.cap, off + 1 + li, S_, ts);
        } else {
            const void* base = MODE == 0
                                   ? (const void*)((const double*)part + poff[b.c] * NACC)
                                   : (const void*)((const float*)part + poff[b.c] * MROWS);
            gather_acc<MODE, (STAGE <= 1)>(base, ens_nchunks(nst, E.L), (size_t)nst, li,
                             MODE == 1 ? cmmax[b.c] : 1.0, S_, ts);
            if (MODE == 1 && fn[b.c] > 0) add_flagged(xp, (size_t)P.cap, off + 1 + li, S_, ts);
        }
        fin_star<STAGE>(in, X, Y, H1, H2, m, P.cap, off + 1 + li, off, S_,
                        gs + (size_t)b.c * NGS, fin_consts(P, cm0[b.c]), acc, cb, a);
        if (MODE == 1 && kout) kout[off + 1 + li] = cancel_ratio(ts, a);
    }
    block_sum<3>(cb, bhblk + (size_t)blockIdx.x * 3);
}

template <int STAGE>
__global__ __launch_bounds__(WG) void k_ens_bh(const double* __restrict__ bhblk,
                                              const int* __restrict__ cblk,
                                              const int* __restrict__ coff,
                                              const double* __restrict__ gs, double* X,
                                              double* Y, double* H1, double* H2,
                                              double* __restrict__ acc, Phys P)
{
    const int c = blockIdx.x;
    double s[3];
    cluster_rows_sum<3>(bhblk, cblk, c, s);
    __shared__ double tot[3];
    block_sum<3>(s, tot);
    if (threadIdx.x == 0) {
        const double t[3] = {tot[0], tot[1], tot[2]};
        bh_finish<STAGE>(t, gs + (size_t)c * NGS, coff[c], P.cap, P.lam, P.dt, X, Y, H1, H2,
                         acc);
    }
}

// ---------------------------------------------------------------------
// ensemble energies (FP64)
// ---------------------------------------------------------------------
inline __global__ __launch_bounds__(WG) void k_ens_epair64(const JRec64* __restrict__ jr,
                                                   const double* __restrict__ st,
                                                   double* __restrict__ part,
                                                   const EItem* __restrict__ items,
                                                   unsigned int* __restrict__ counter,
                                                   const int* __restrict__ qoff, Phys P)
{
    int q = 0;
    for (;;) {
        const int it = ens_take(counter, qoff, q);
        if (it < 0) break;
        const EItem w = items[it];
        const EGeom g = ens_geom(w);
        IE64 I;
        I.x = jr[g.i].x; I.y = jr[g.i].y; I.z = jr[g.i].z;
        I.ux = st[3 * P.cap + g.i] / PN1_SQ15;
        I.uy = st[4 * P.cap + g.i] / PN1_SQ15;
        I.uz = st[5 * P.cap + g.i] / PN1_SQ15;
        double a[NACCE];
#pragma unroll
        for (int k = 0; k < NACCE; ++k) a[k] = 0.0;
        if (g.w0 < g.w1) {
            erun64<false>(jr, w.j0, g.w0, g.i, I, a, P.eps2);
            erun64<true>(jr, g.w0, g.w1, g.i, I, a, P.eps2);
            erun64<false>(jr, g.w1, w.j1, g.i, I, a, P.eps2);
        } else {
            erun64<false>(jr, w.j0, w.j1, g.i, I, a, P.eps2);
        }
        if (g.li < w.nst) {
            double* pp = part + w.pslot * NACCE + g.li;
#pragma unroll
            for (int k = 0; k < NACCE; ++k) stream_out(a[k], &pp[(size_t)k * w.nst]);
        }
    }
}

inline __global__ __launch_bounds__(WG) void k_ens_efin(const double* __restrict__ st,
                                                const double* __restrict__ m,
                                                const double* __restrict__ cm0,
                                                const int* __restrict__ coff,
                                                const EBlk* __restrict__ blk,
                                                const double* __restrict__ part,
                                                const long long* __restrict__ poff,
                                                const double* __restrict__ gs,
                                                double* __restrict__ eblk, Phys P, EConst E)
{
    const EBlk b = blk[blockIdx.x];
    const int off = coff[b.c], nst = coff[b.c + 1] - off - 1;
    const int li = b.first + (int)threadIdx.x;
    double t[NEB] = {0.0, 0.0, 0.0, 0.0};
    if (li < nst) {
        const double* pp = part + poff[b.c] * NACCE + li;
        double A[NACCE] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const int S = ens_nchunks(nst, E.L);
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int k = 0; k < NACCE; ++k) A[k] += pp[(size_t)(s * NACCE + k) * nst];
        efin_star(st, m, P.cap, off + 1 + li, off, cm0[b.c], P.eps2, A,
                  gs[(size_t)b.c * NGS + G_G1], t);
    }
    block_sum<NEB>(t, eblk + (size_t)blockIdx.x * NEB);
}

// out[c] = {kin, pot, pot_pn/c^2, pot_pn2/c^2, total} of cluster c
inline __global__ __launch_bounds__(WG) void k_ens_ered(const double* __restrict__ eblk,
                                                const int* __restrict__ cblk,
                                                const int* __restrict__ coff,
                                                const double* __restrict__ st,
                                                const double* __restrict__ m,
                                                const double* __restrict__ gs,
                                                double* __restrict__ out, Phys P)
{
    const int c = blockIdx.x;
    double s[NEB];
    cluster_rows_sum<NEB>(eblk, cblk, c, s);
    __shared__ double tot[NEB];
    block_sum<NEB>(s, tot);
    if (threadIdx.x == 0) {
        const int b0 = coff[c], cap = P.cap;
        const double v0x = st[3 * cap + b0], v0y = st[4 * cap + b0], v0z = st[5 * cap + b0];
        const double kin = 0.5 * (m[b0] * (v0x * v0x + v0y * v0y + v0z * v0z) + tot[0]);
        const double pot = tot[1] - m[b0] * gs[(size_t)c * NGS + G_G1];
        const double pn = P.lam * tot[2], pn2 = P.lam * tot[3];
        double* o = out + (size_t)c * 5;
        o[0] = kin;
        o[1] = pot;
        o[2] = pn;
        o[3] = pn2;
        o[4] = kin + pot + pn + pn2;
    }
}

}  // namespace pn1dev
