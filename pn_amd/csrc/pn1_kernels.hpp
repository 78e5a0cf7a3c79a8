// pn1_kernels.hpp -- device code of libpn1 (gfx950, wave64).
//
// One force evaluation of Limit-sticky9's force() (Limit-sticky9.c:21-424)
// is five launches on one stream:
//
//   k_prep    O(n)   per-star BH geometry (d, |d|^2, e = (|d|^2+eps^2)^-1/2)
//                    written once as an AoS j-record, plus per-block
//                    partials of the separable sums (sum m e, sum m e^3 d, ...)
//   k_gsum    1 WG   fixed-order reduction of those partials
//   k_pair*   O(n^2) the star-star pair sums, one rsqrt per pair
//   k_fin     O(n)   star accelerations from the pair sums and the O(n)
//                    closed forms; the stage update of the RK driver; per-block
//                    partials of the BH's 1PN acceleration
//   k_bh      1 WG   BH acceleration and its stage update
//
// Pair kernels: i-particle per lane in VGPRs; the j loop has wave-uniform
// bounds, so each j-record is fetched with scalar loads into SGPRs and
// broadcast to all 64 lanes as an instruction operand (no LDS traffic in
// the inner loop).  The j range is split into S chunks (one partial per
// chunk, reduced in fixed order by k_fin), and workgroups are mapped so that
// each XCD works on one chunk, whose records then stay in that XCD's L2.
//
// All reductions use a fixed order, so results are bitwise reproducible
// for a given launch shape.
//
// Two translation units include this header: pn1_api.hip launches every
// kernel but the mixed-form pair kernels, pn1_pair32.hip those.  The kernels
// that are not templates are therefore declared inline.
#pragma once
#include "pn1_common.hpp"

namespace pn1dev {
// Stage update of the reference's RK driver for particle p, given its
// acceleration a (Limit-sticky9.c:853-891).  STAGE 1..3 follow the three
// force calls; STAGE 0 (plain force evaluation) does nothing.  FMA
// contraction is off so each line rounds exactly as the reference's.
template <int STAGE>
__device__ __forceinline__ void stage_update(int p, int cap, const double (&a)[3],
                                             double* X, double* Y, double* H1,
                                             double* H2, double dt)
{
#pragma clang fp contract(off)
    if (STAGE == 0) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const size_t ix = (size_t)k * cap + p, iv = (size_t)(3 + k) * cap + p;
        const double x = X[ix], v = X[iv];
        if (STAGE == 1) {
            const double x1 = x + 0.5 * v * dt;
            const double v1 = v + 0.5 * a[k] * dt;
            Y[ix] = x1; Y[iv] = v1; H1[ix] = x1; H1[iv] = v1;
        } else if (STAGE == 2) {
            const double v1o = Y[iv];            // v1 of stage 1
            const double x1 = x + 0.5 * v1o * dt;
            const double v1 = v + 0.5 * a[k] * dt;
            Y[ix] = x1; Y[iv] = v1; H2[ix] = x1; H2[iv] = v1;
        } else {
            const double v1o = Y[iv];            // v1 of stage 2
            const double hx3 = x + v1o * dt;
            const double hv3 = v + a[k] * dt;
            X[ix] = (x + hx3 + 2.0 * (H1[ix] + H2[ix])) / 6.0;
            X[iv] = (v + hv3 + 2.0 * (H1[iv] + H2[iv])) / 6.0;
        }
    }
}

// ---------------------------------------------------------------------
// k_prep: per-star BH geometry + separable sums
// ---------------------------------------------------------------------
// Mixed form's pair thresholds of a star at x from its BH whose previous
// evaluation gave the acceleration aprev and cancellation ratio kprev
// (null: none known), for star masses up to mmax (see Thr).
__device__ __forceinline__ Thr pair_thresholds(double x, double y, double z,
                                               const double* aprev, const float* kprev,
                                               size_t cap, int p, double mmax, bool rows,
                                               bool& hard)
{
    Thr t;
    hard = false;
    if (!aprev) {
        t.rc2 = 0.0f;
        t.dn2 = __FLT_MAX__;
        return t;
    }
    hard = !(kprev[p] <= (float)PN1_HARD_K);                   // NaN too
    if (hard && rows) {          // the FP32 lane's sums are not used
        t.rc2 = t.dn2 = 0.0f;
        return t;
    }
    const double ax = aprev[p], ay = aprev[cap + p], az = aprev[2 * cap + p];
    const double A = sqrt(ax * ax + ay * ay + az * az);
    const double X = sqrt(x * x + y * y + z * z);
    const double rc2 = mmax / (PN1_CLOSE_KAPPA * A);            // +inf if A == 0
    const double dn3 = 6.0 * 0x1p-24 * X * mmax / (PN1_LAZY_ETA * A);
    const double dn2 = cbrt(dn3 * dn3);
    t.rc2 = hard ? __FLT_MAX__ : (float)fmin(rc2, (double)__FLT_MAX__);
    // close pairs are found inside the lo-part branch: dn2 >= rc2
    t.dn2 = fmaxf((float)fmin(dn2, (double)__FLT_MAX__), t.rc2);
    return t;
}

// Star p (SoA stride cap) of a cluster whose BH sits at index b0 and has
// mass m0: writes its j-records and adds its terms of the separable sums
// to g.  Mixed form (M.j32 set): also its pair thresholds, from its previous
// acceleration M.aprev (SoA [3][cap]) and cancellation ratio M.kprev, and
// its mixed record in units of the cluster's largest star mass M.mmax;
// returns whether the star is close enough to the BH to be taken out of the
// FP32 pass (see PN1_FLAG_ETA), in which case its mixed record is neutral;
// hard: whether its pair terms cancel beyond PN1_HARD_K.
__device__ __forceinline__ bool prep_star(const double* __restrict__ st,
                                          const double* __restrict__ m, int cap, int p,
                                          int b0, double m0, double eps2,
                                          JRec64* __restrict__ j64, double (&g)[NGS],
                                          const MixPrep& M, bool& hard)
{
#pragma clang fp contract(off)
    hard = false;
    const double x = st[p], y = st[cap + p], z = st[2 * cap + p];
    const double vx = st[3 * cap + p], vy = st[4 * cap + p], vz = st[5 * cap + p];
    const double mp = m[p];
    const double dx = x - st[b0], dy = y - st[cap + b0], dz = z - st[2 * cap + b0];
    const double D2 = dx * dx + dy * dy + dz * dz;
    const double e = 1.0 / sqrt(D2 + eps2);     // as Limit-sticky9.c:106-108
    const double E3 = e * e * e;
    const double q = mp * E3;
    const double V2 = vx * vx + vy * vy + vz * vz;
    JRec64 r;
    r.x = x; r.y = y; r.z = z;
    r.sx = PN1_SQ15 * vx; r.sy = PN1_SQ15 * vy; r.sz = PN1_SQ15 * vz;
    r.m = mp; r.D2 = D2;
    r.A = 1.25 * m0 * e - 2.0 * V2; r.B = 0.25 * m0 * E3;
    const double Q = 3.5 * m0 * q;
    r.MB = mp * r.B;
    r.Qdx = Q * dx; r.Qdy = Q * dy; r.Qdz = Q * dz;
    r.e = e; r.pad = 0.0;
    j64[p] = r;
    g[G_G1] = mp * e;
    g[G_M1X] = q * dx; g[G_M1Y] = q * dy; g[G_M1Z] = q * dz;
    g[G_PX] = mp * vx; g[G_PY] = mp * vy; g[G_PZ] = mp * vz;
    if (!M.j32) return false;
    const bool flagged = M.lam * mp * e * e > M.tauf;
    JRec32 h;
    if (flagged) {
        // far away (w ~ 1e-18: every FP32 pair term vanishes), no mass and
        // no 1PN fields (t = |r|^2 - D2_i is then ~1e36)
        h.xh = h.yh = h.zh = PN1_FAR;
        h.xl = h.yl = h.zl = 0.0f;
        h.sx = h.sy = h.sz = 0.0f;
        h.m = 0.0f;
        h.A = h.B = h.MB = 0.0f;
        h.Qdx = h.Qdy = h.Qdz = 0.0f;
    } else {
        // double-single position relative to the BH: the FP32 hi parts then
        // round at the cluster's own scale, wherever the cluster sits
        h.xh = (float)dx; h.yh = (float)dy; h.zh = (float)dz;
        h.xl = (float)(dx - (double)h.xh);
        h.yl = (float)(dy - (double)h.yh);
        h.zl = (float)(dz - (double)h.zh);
        h.sx = (float)r.sx; h.sy = (float)r.sy; h.sz = (float)r.sz;
        h.m = M.mmax > 0.0 ? (float)(mp / M.mmax) : 0.0f;     // no scale: massless stars
        h.A = (float)r.A; h.B = (float)r.B; h.MB = (float)r.MB;
        h.Qdx = (float)r.Qdx; h.Qdy = (float)r.Qdy; h.Qdz = (float)r.Qdz;
    }
    M.j32[p] = h;
    M.thr[p] = pair_thresholds(dx, dy, dz, M.aprev, M.kprev, (size_t)cap, p, M.mmax,
                               M.hard_rows != 0, hard);
    return flagged;
}

// Mixed form: fl per star, bit 0 whether it is taken out of the FP32 pass
// (BH-close), bit 1 whether it gets an FP64 row of its own (hard); fblk per
// block the number of each, [2 b] and [2 b + 1].  The BH mass is m[0].
inline __global__ __launch_bounds__(WG) void k_prep(const double* __restrict__ st,
                                            const double* __restrict__ m, Phys P,
                                            JRec64* __restrict__ j64,
                                            double* __restrict__ gblk, MixPrep M,
                                            int* __restrict__ fl, int* __restrict__ fblk)
{
    const int p = blockIdx.x * WG + threadIdx.x + 1;
    double g[NGS];
#pragma unroll
    for (int k = 0; k < NGS; ++k) g[k] = 0.0;
    bool f = false, hard = false;
    if (p < P.n) {
        f = prep_star(st, m, P.cap, p, 0, m[0], P.eps2, j64, g, M, hard);
        hard = hard && M.hard_rows;
        if (M.j32) fl[p] = (f ? 1 : 0) | (hard ? 2 : 0);
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

// Exclusive prefix sum of one value per thread over the workgroup, in
// thread order; *total receives the sum.
__device__ __forceinline__ int block_excl_scan(int v, int* total)
{
    __shared__ int wsum[WG / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int base = 0;
#pragma unroll
    for (int w = 0; w < WG / 64; ++w) base += w < wid ? wsum[w] : 0;
    *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return base + incl - v;
}

// The stars with bit BIT of fl set in blocks [b0, b1) (block b covers the
// stars p0 + (b - b0) * WG .. + WG - 1, its count of them at
// fblk[b * STRIDE + BIT]) listed in increasing order at out[0 ..), their
// number in *nout and returned to every thread.  One workgroup.
template <int BIT = 0, int STRIDE = 1>
__device__ __forceinline__ int list_flagged(const int* __restrict__ fl,
                                             const int* __restrict__ fblk, int b0, int b1,
                                             int p0, int pend, int* __restrict__ out,
                                             int* __restrict__ nout)
{
    int base = 0;
    for (int bb = b0; bb < b1; bb += WG) {
        const int b = bb + (int)threadIdx.x;
        const int cnt = b < b1 ? fblk[b * STRIDE + BIT] : 0;
        int total;
        const int at = base + block_excl_scan(cnt, &total);
        if (cnt > 0) {
            // the block's flags 16 at a time: 16 loads in flight, not one
            int k = at;
            const int q0 = p0 + (b - b0) * WG, qe = min(q0 + WG, pend);
            for (int q = q0; q < qe; q += 16) {
                int f[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) f[u] = fl[min(q + u, qe - 1)];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (q + u < qe && ((f[u] >> BIT) & 1)) out[k++] = q + u;
            }
        }
        base += total;
    }
    if (threadIdx.x == 0) *nout = base;
    return base;
}

// The BH-close stars (flist, fn[0]) and the hard stars (hlist, fn[1]) of
// k_prep's flags.
inline __global__ __launch_bounds__(WG) void k_flist(const int* __restrict__ fl,
                                             const int* __restrict__ fblk, int nblk, int n,
                                             int* __restrict__ flist, int* __restrict__ hlist,
                                             int* __restrict__ fn)
{
    list_flagged<0, 2>(fl, fblk, 0, nblk, 1, n, flist, fn);
    list_flagged<1, 2>(fl, fblk, 0, nblk, 1, n, hlist, fn + 1);
}

// Fixed-order sum of nblk rows of N doubles (one workgroup).
template <int N>
__global__ __launch_bounds__(WG) void k_rows_sum(const double* __restrict__ rows, int nblk,
                                                double* __restrict__ out)
{
    double g[N];
#pragma unroll
    for (int k = 0; k < N; ++k) g[k] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += WG)
#pragma unroll
        for (int k = 0; k < N; ++k) g[k] += rows[(size_t)b * N + k];
    block_sum<N>(g, out);
}

// ---------------------------------------------------------------------
// force pair kernel, FP64
// ---------------------------------------------------------------------
// Per-lane state of the i particle: position, g = 4 v, h = 4 v / sqrt(3/2),
// |d|^2 and |d|^2 + eps^2.
struct I64 {
    double x, y, z, gx, gy, gz, hx, hy, hz, D2e;
};

__device__ __forceinline__ I64 load_i64(const JRec64* __restrict__ jr,
                                        const double* __restrict__ st, int cap, int i,
                                        double eps2)
{
    const JRec64& r = jr[i];
    const double vx = st[3 * cap + i], vy = st[4 * cap + i], vz = st[5 * cap + i];
    I64 I;
    I.x = r.x; I.y = r.y; I.z = r.z;
    I.gx = 4.0 * vx; I.gy = 4.0 * vy; I.gz = 4.0 * vz;
    I.hx = PN1_4OSQ15 * vx; I.hy = PN1_4OSQ15 * vy; I.hz = PN1_4OSQ15 * vz;
    I.D2e = r.D2 + eps2;
    return I;
}

template <bool MASK>
__device__ __forceinline__ void body64(const JRec64& J, int j, int i, const I64& I,
                                       double (&a)[NACC], double eps2)
{
    double mj = J.m, MBj = J.MB, Qdx = J.Qdx, Qdy = J.Qdy, Qdz = J.Qdz;
    if (MASK) {
        const bool self = (j == i);
        mj = self ? 0.0 : mj;
        MBj = self ? 0.0 : MBj;
        Qdx = self ? 0.0 : Qdx;
        Qdy = self ? 0.0 : Qdy;
        Qdz = self ? 0.0 : Qdz;
    }
    const double rx = J.x - I.x, ry = J.y - I.y, rz = J.z - I.z;
    const double R2e = __builtin_fma(rx, rx, __builtin_fma(ry, ry, __builtin_fma(rz, rz, eps2)));
    const double w = rsq64(R2e);
    const double w2 = w * w;
    const double mw = mj * w;
    const double mw3 = mw * w2;
    a[A_NX] = __builtin_fma(mw3, rx, a[A_NX]);
    a[A_NY] = __builtin_fma(mw3, ry, a[A_NY]);
    a[A_NZ] = __builtin_fma(mw3, rz, a[A_NZ]);
    a[A_PHI] += mw;
    a[A_W3] += mw3;
    a[A_TX] = __builtin_fma(w, Qdx, a[A_TX]);
    a[A_TY] = __builtin_fma(w, Qdy, a[A_TY]);
    a[A_TZ] = __builtin_fma(w, Qdz, a[A_TZ]);
    const double t = R2e - I.D2e;
    a[A_ZE] = __builtin_fma(MBj, t, a[A_ZE]);
    const double sr = __builtin_fma(J.sx, rx, __builtin_fma(J.sy, ry, J.sz * rz));
    // one record operand per instruction (one scalar register pair per VALU op)
    const double tv = sr * w;
    double br = __builtin_fma(tv, tv, J.A);
    br = __builtin_fma(J.B, t, br);
    br = __builtin_fma(I.hx, J.sx, __builtin_fma(I.hy, J.sy, __builtin_fma(I.hz, J.sz, br)));
    const double Kr = mw3 * br;
    a[A_RX] = __builtin_fma(Kr, rx, a[A_RX]);
    a[A_RY] = __builtin_fma(Kr, ry, a[A_RY]);
    a[A_RZ] = __builtin_fma(Kr, rz, a[A_RZ]);
    const double gr = __builtin_fma(I.gx, rx, __builtin_fma(I.gy, ry, I.gz * rz));
    const double Ku = mw3 * __builtin_fma(-PN1_SQ6, sr, gr);
    a[A_UX] = __builtin_fma(Ku, J.sx, a[A_UX]);
    a[A_UY] = __builtin_fma(Ku, J.sy, a[A_UY]);
    a[A_UZ] = __builtin_fma(Ku, J.sz, a[A_UZ]);
    a[A_KS] += Ku;
}

// j loop.  PP bit 0 clear: one record at a time; set: two records in
// flight, the scalar loads of one issued before the other is used (j + 1 <
// j1 <= n, so jr[j + 2] stays within the cap + 2 records allocated).
// PP bit 1 (mixed form only): one Newton step on the FP32 rsqrt.
template <bool MASK, int PP>
__device__ __forceinline__ void run64(const JRec64* __restrict__ jr, int j0, int j1, int i,
                                      const I64& I, double (&a)[NACC], double eps2)
{
    if ((PP & 1) == 0) {
        for (int j = j0; j < j1; ++j) body64<MASK>(jr[j], j, i, I, a, eps2);
        return;
    }
    if (j0 >= j1) return;
    JRec64 ra = jr[j0];
    int j = j0;
    for (; j + 1 < j1; j += 2) {
        const JRec64 rb = jr[j + 1];
        body64<MASK>(ra, j, i, I, a, eps2);
        ra = jr[j + 2];
        body64<MASK>(rb, j + 1, i, I, a, eps2);
    }
    if (j < j1) body64<MASK>(ra, j, i, I, a, eps2);
}

// i blocks ib0 .. ib0 + nIblk - 1 (ib0 = 0: every star; a later first block
// evaluates one i shard of the stars against all of them).
template <int PP>
__global__ __launch_bounds__(WG) void k_pair64(const JRec64* __restrict__ jr,
                                              const double* __restrict__ st,
                                              double* __restrict__ part, Phys P, int S,
                                              int nIblk, int ib0)
{
    int s, ib;
    map_block(blockIdx.x, S, nIblk, s, ib);
    if (ib >= nIblk) return;
    ib += ib0;
    const int n = P.n;
    const int li = ib * WG + threadIdx.x + 1;
    const int i = li < n ? li : n - 1;
    const I64 I = load_i64(jr, st, P.cap, i, P.eps2);
    const JRange R = j_range(s, S, n, ib);
    const double eps2 = P.eps2;
    double a[NACC];
#pragma unroll
    for (int k = 0; k < NACC; ++k) a[k] = 0.0;
    if (R.w0 < R.w1) {
        run64<false, PP>(jr, R.j0, R.w0, i, I, a, eps2);
        run64<true, PP>(jr, R.w0, R.w1, i, I, a, eps2);
        run64<false, PP>(jr, R.w1, R.j1, i, I, a, eps2);
    } else {
        run64<false, PP>(jr, R.j0, R.j1, i, I, a, eps2);
    }
    if (li < n) {
#pragma unroll
        for (int k = 0; k < NACC; ++k) stream_out(a[k], &part[(size_t)(s * NACC + k) * P.cap + li]);
    }
}

// ---------------------------------------------------------------------
// force pair kernel, mixed: FP32 pair arithmetic on double-single
// positions, FP32 tile sums folded into Kahan-compensated FP32 totals, and
// the Newtonian term of the few pairs that can dominate a star's
// acceleration in FP64
// ---------------------------------------------------------------------
struct I32 {
    float xh, yh, zh, xl, yl, zl, gx, gy, gz, hx, hy, hz, D2e, rc2, dn2;
};

// Star i: its double-single position relative to its BH (particle b0; as
// in its mixed record, which is neutral if the star is close to the BH, so
// it is rebuilt from the FP64 record), velocity, |d_i|^2 and pair
// thresholds.
__device__ __forceinline__ I32 load_i32(const JRec64* __restrict__ j64,
                                        const Thr* __restrict__ thr,
                                        const double* __restrict__ st, int cap, int i, int b0,
                                        double eps2)
{
    const JRec64& r = j64[i];
    const double vx = st[3 * cap + i], vy = st[4 * cap + i], vz = st[5 * cap + i];
    const double D2 = r.D2;
    const Thr t = thr[i];
    double dx, dy, dz;
    {
#pragma clang fp contract(off)
        dx = r.x - st[b0];
        dy = r.y - st[cap + b0];
        dz = r.z - st[2 * cap + b0];
    }
    I32 I;
    I.xh = (float)dx; I.yh = (float)dy; I.zh = (float)dz;
    I.xl = (float)(dx - (double)I.xh);
    I.yl = (float)(dy - (double)I.yh);
    I.zl = (float)(dz - (double)I.zh);
    I.gx = (float)(4.0 * vx); I.gy = (float)(4.0 * vy); I.gz = (float)(4.0 * vz);
    I.hx = (float)(PN1_4OSQ15 * vx); I.hy = (float)(PN1_4OSQ15 * vy);
    I.hz = (float)(PN1_4OSQ15 * vz);
    I.D2e = (float)(D2 + eps2);
    I.rc2 = t.rc2;
    I.dn2 = t.dn2;
    return I;
}

// Folds the short-run FP32 Newtonian sums into their Kahan totals (total =
// sN - cN).
__device__ __forceinline__ void fold_newton(float (&a)[NACC32], float (&sN)[3], float (&cN)[3])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma clang fp contract(off)
        const float y = a[A_NX + k] - cN[k];
        const float u = sN[k] + y;
        cN[k] = (u - sN[k]) - y;
        sN[k] = u;
        a[A_NX + k] = 0.0f;
    }
}

// u = fl(s + y), and its rounding error (Knuth's TwoSum, exact for any
// magnitudes) taken off the compensation c of a total kept as s - c.
__device__ __forceinline__ void two_sum_into(float s, float y, float& u, float& c)
{
#pragma clang fp contract(off)
    u = s + y;
    const float bv = u - s;
    const float av = u - bv;
    c -= (s - av) + (y - bv);
}

// Close pairs (R2e < rc2, not the self pair), taken when any lane of the
// wave has one; `close` selects the lanes.  A close pair's Newtonian term
// (m_j / mscale) w^3 r is evaluated in FP64 (rscale = 1 / mscale), from
// the mass of j's FP64 record and the double-single positions of j (its
// mixed record Js) and of the lane, and goes into the Kahan total; the
// FP32 term
// mw3 r that the caller adds next is cancelled exactly: the short-run sum
// is folded, then set to -p with p = fl(mw3 r), so the caller's
// fma(mw3, r, -p) leaves the exact product residual e = mw3 r - p, and
// T64 - e enters the total (as hi + lo floats: hi by TwoSum, lo into the
// compensation).  Every lane computes; the other lanes keep their sums.
// The branch around it is wave-uniform, so j and the record stream stay
// scalar.
__device__ __forceinline__ void close_pairs(bool close, const JRec64& J, const JRec32& Js,
                                            const I32& I, double eps2, double rscale, float mw3,
                                            float rx, float ry, float rz, float (&a)[NACC32],
                                            float (&sN)[3], float (&cN)[3])
{
#pragma clang fp contract(off)
    // the lane's position (double-single, relative to the BH: hi + lo is it
    // to 2^-48), made opaque here so that its FP64 form is built in this
    // rarely taken branch rather than hoisted out of the j loop and kept
    // live across it; j's from its mixed record, the same way
    float ph[3] = {I.xh, I.yh, I.zh}, pl[3] = {I.xl, I.yl, I.zl};
#pragma unroll
    for (int k = 0; k < 3; ++k) __asm__ volatile("" : "+v"(ph[k]), "+v"(pl[k]));
    const double dx = ((double)Js.xh + (double)Js.xl) - ((double)ph[0] + (double)pl[0]);
    const double dy = ((double)Js.yh + (double)Js.yl) - ((double)ph[1] + (double)pl[1]);
    const double dz = ((double)Js.zh + (double)Js.zl) - ((double)ph[2] + (double)pl[2]);
    const float jm32 = Js.m;
    const double R2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, eps2)));
    const double w = rsq64(R2);
    // in the record's mass units; none for a star taken out of the FP32 pass
    // (its mixed record, mass jm32, is neutral)
    const double mw3d = (jm32 != 0.0f ? J.m * rscale : 0.0) * (w * (w * w));
    const double d[3] = {dx, dy, dz};
    const float r[3] = {rx, ry, rz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float p = mw3 * r[k];
        const float e = __builtin_fmaf(mw3, r[k], -p);
        const double D = mw3d * d[k] - (double)e;
        const float hi = (float)D;
        const float lo = (float)(D - (double)hi);
        // total = s - c; the short-run sum, then hi + lo, are added with
        // TwoSum, exact whatever the magnitudes (Kahan's update is exact
        // only while the running sum is the larger addend, and a star whose
        // terms cancel to a small |a| breaks that)
        float s1, c1 = cN[k];
        two_sum_into(sN[k], a[A_NX + k], s1, c1);
        float s2 = s1;
        two_sum_into(s1, hi, s2, c1);
        c1 -= lo;
        a[A_NX + k] = close ? -p : a[A_NX + k];
        sN[k] = close ? s2 : sN[k];
        cN[k] = close ? c1 : cN[k];
    }
}

// FP32 reciprocal square root of the pair kernels (RF: one Newton step,
// v_rsq_f32 ~1.6 ulp -> ~1 ulp).
template <bool RF>
__device__ __forceinline__ float rsq32(float R2e)
{
    float w = __builtin_amdgcn_rsqf(R2e);
    if (RF) {
        const float r = __builtin_fmaf(-(R2e * w), w, 1.0f);
        w = __builtin_fmaf(0.5f * w, r, w);
    }
    return w;
}

// One pair.  TS: also accumulate A_TS (the evaluations that refresh the
// stars' cancellation ratios).  EQ: every star of the cluster(s) has the
// same mass, so the record's mass m_j / mscale is 1 exactly (the self pair
// of the masked band keeps its zero mass) and m_j w is w.
template <bool MASK, bool RF, bool TS, bool EQ>
__device__ __forceinline__ void body32(const JRec32& J, const JRec64* __restrict__ j64, int j,
                                       int i, const I32& I, float (&a)[NACC32], float (&sN)[3],
                                       float (&cN)[3], float eps2, double eps2d, double rscale)
{
    float mj = J.m, MBj = J.MB, Qdx = J.Qdx, Qdy = J.Qdy, Qdz = J.Qdz;
    if (MASK) {
        const bool self = (j == i);
        mj = self ? 0.0f : mj;
        MBj = self ? 0.0f : MBj;
        Qdx = self ? 0.0f : Qdx;
        Qdy = self ? 0.0f : Qdy;
        Qdz = self ? 0.0f : Qdz;
    }
    // Separation from the hi parts of the positions; inside the lane's dn2
    // the lo parts are added, which gives the double-single separation
    // (x_hi,j - x_hi,i) + (x_lo,j - x_lo,i), so close pairs keep their digits.
    float rx = J.xh - I.xh, ry = J.yh - I.yh, rz = J.zh - I.zh;
    float R2e = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
    if (R2e < I.dn2) {
        rx += J.xl - I.xl;
        ry += J.yl - I.yl;
        rz += J.zl - I.zl;
        R2e = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
        // close pairs (dn2 >= rc2, so they are all in this branch): their
        // FP32 mw3 is formed here as below, bit for bit
        const bool close = R2e < I.rc2 && (!MASK || j != i);
        if (__builtin_amdgcn_ballot_w64(close)) {
            const float wc = rsq32<RF>(R2e);
            const float mw3c = (EQ && !MASK ? wc : mj * wc) * (wc * wc);
            close_pairs(close, j64[j], J, I, eps2d, rscale, mw3c, rx, ry, rz, a, sN, cN);
        }
    }
    const float w = rsq32<RF>(R2e);
    const float w2 = w * w;
    constexpr bool UNIT = EQ && !MASK;
    const float mw = UNIT ? w : mj * w;
    const float mw3 = mw * w2;
    a[A_PHI] += mw;
    a[A_W3] += mw3;
    if (TS) a[A_TS] = UNIT ? a[A_TS] + w2 : __builtin_fmaf(mw, w, a[A_TS]);
    a[A_TX] = __builtin_fmaf(w, Qdx, a[A_TX]);
    a[A_TY] = __builtin_fmaf(w, Qdy, a[A_TY]);
    a[A_TZ] = __builtin_fmaf(w, Qdz, a[A_TZ]);
    const float t = R2e - I.D2e;
    a[A_ZE] = __builtin_fmaf(MBj, t, a[A_ZE]);
    const float sr = __builtin_fmaf(J.sx, rx, __builtin_fmaf(J.sy, ry, J.sz * rz));
    // the bracket starts from (s_j.r w)^2 + A_j: one record operand per
    // instruction (gfx950 reads one scalar register per VALU op), no copy
    const float tv = sr * w;
    float br = __builtin_fmaf(tv, tv, J.A);
    br = __builtin_fmaf(J.B, t, br);
    br = __builtin_fmaf(I.hx, J.sx, __builtin_fmaf(I.hy, J.sy, __builtin_fmaf(I.hz, J.sz, br)));
    const float Kr = mw3 * br;
    a[A_RX] = __builtin_fmaf(Kr, rx, a[A_RX]);
    a[A_RY] = __builtin_fmaf(Kr, ry, a[A_RY]);
    a[A_RZ] = __builtin_fmaf(Kr, rz, a[A_RZ]);
    const float gr = __builtin_fmaf(I.gx, rx, __builtin_fmaf(I.gy, ry, I.gz * rz));
    const float Ku = mw3 * __builtin_fmaf((float)-PN1_SQ6, sr, gr);
    a[A_UX] = __builtin_fmaf(Ku, J.sx, a[A_UX]);
    a[A_UY] = __builtin_fmaf(Ku, J.sy, a[A_UY]);
    a[A_UZ] = __builtin_fmaf(Ku, J.sz, a[A_UZ]);
    a[A_KS] += Ku;
    a[A_NX] = __builtin_fmaf(mw3, rx, a[A_NX]);
    a[A_NY] = __builtin_fmaf(mw3, ry, a[A_NY]);
    a[A_NZ] = __builtin_fmaf(mw3, rz, a[A_NZ]);
}

// Mixed-form j loop: PP bit 0 clear: one body at a time; set: two bodies
// per iteration, records two ahead (as run64).
// One j-record with a single 64-byte scalar load (one wait per pair instead
// of one per group of fields the body reaches).
typedef float f32x16 __attribute__((ext_vector_type(16)));
__device__ __forceinline__ JRec32 load_rec(const JRec32* __restrict__ p)
{
    const f32x16 v = *reinterpret_cast<const f32x16*>(p);
    JRec32 r;
    r.xh = v[0]; r.yh = v[1]; r.zh = v[2];
    r.xl = v[3]; r.yl = v[4]; r.zl = v[5];
    r.sx = v[6]; r.sy = v[7]; r.sz = v[8];
    r.m = v[9]; r.A = v[10]; r.B = v[11]; r.MB = v[12];
    r.Qdx = v[13]; r.Qdy = v[14]; r.Qdz = v[15];
    return r;
}

// A wave-uniform pair counter folds the Newtonian sums into Kahan totals
// every `tile` pairs without restarting the record stream.
template <bool MASK, int PP, bool TS, bool EQ>
__device__ __forceinline__ void run32(const JRec32* __restrict__ jr,
                                      const JRec64* __restrict__ j64, int j0, int j1, int i,
                                      const I32& I, float (&a)[NACC32], float eps2, double eps2d,
                                      double rscale, float (&sN)[3], float (&cN)[3], int tile,
                                      int& cnt)
{
    constexpr bool RF = (PP & 2) != 0;
    if ((PP & 1) == 0) {
        for (int j = j0; j < j1; ++j) {
            body32<MASK, RF, TS, EQ>(load_rec(jr + j), j64, j, i, I, a, sN, cN, eps2, eps2d,
                                     rscale);
            if (++cnt >= tile) {
                cnt = 0;
                fold_newton(a, sN, cN);
            }
        }
        return;
    }
    if (j0 >= j1) return;
    JRec32 ra = jr[j0];
    int j = j0;
    for (; j + 1 < j1; j += 2) {
        const JRec32 rb = jr[j + 1];
        body32<MASK, RF, TS, EQ>(ra, j64, j, i, I, a, sN, cN, eps2, eps2d, rscale);
        ra = jr[j + 2];
        body32<MASK, RF, TS, EQ>(rb, j64, j + 1, i, I, a, sN, cN, eps2, eps2d, rscale);
        cnt += 2;
        if (cnt >= tile) {
            cnt = 0;
            fold_newton(a, sN, cN);
        }
    }
    if (j < j1) {
        body32<MASK, RF, TS, EQ>(ra, j64, j, i, I, a, sN, cN, eps2, eps2d, rscale);
        ++cnt;
    }
}


// The Newtonian sums are Kahan-compensated every `tile` pairs (at least
// 2; 64 by default): they carry the stars' accelerations, and with the
// close pairs taken out of the FP32 run sums, a run holds no term larger
// than PN1_CLOSE_KAPPA |a|.  The other sums enter scaled by 1/c^2 (or, for the
// BH, beside its FP64 Newtonian term) and accumulate in plain FP32 over
// the whole chunk.
// The one-record form is held to 8 waves per SIMD: under the max-ILP
// scheduler (pn1_pair32.hip) it would otherwise take 65 VGPRs.
template <int PP, bool TS, bool EQ>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu((PP & 1) ? 1 : 8, 8)))
void k_pair32(const JRec32* __restrict__ jr, const JRec64* __restrict__ j64,
              const Thr* __restrict__ thr, const double* __restrict__ st,
              float* __restrict__ part, Phys P, int S, int nIblk, int ib0, int tile)
{
    int s, ib;
    map_block(blockIdx.x, S, nIblk, s, ib);
    if (ib >= nIblk) return;
    ib += ib0;
    const int n = P.n;
    const int li = ib * WG + threadIdx.x + 1;
    const int i = li < n ? li : n - 1;
    const I32 I = load_i32(j64, thr, st, P.cap, i, 0, P.eps2);
    const JRange R = j_range(s, S, n, ib);
    const float eps2 = P.eps2f;
    const double eps2d = P.eps2, rscale = 1.0 / P.mmax;
    float a[NACC32], sumN[3], cmpN[3];
#pragma unroll
    for (int k = 0; k < NACC32; ++k) a[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) sumN[k] = cmpN[k] = 0.0f;
    int cnt = 0;
    if (R.w0 < R.w1) {
        run32<false, PP, TS, EQ>(jr, j64, R.j0, R.w0, i, I, a, eps2, eps2d, rscale, sumN, cmpN, tile, cnt);
        run32<true, PP, TS, EQ>(jr, j64, R.w0, R.w1, i, I, a, eps2, eps2d, rscale, sumN, cmpN, tile, cnt);
        run32<false, PP, TS, EQ>(jr, j64, R.w1, R.j1, i, I, a, eps2, eps2d, rscale, sumN, cmpN, tile, cnt);
    } else {
        run32<false, PP, TS, EQ>(jr, j64, R.j0, R.j1, i, I, a, eps2, eps2d, rscale, sumN, cmpN, tile, cnt);
    }
    fold_newton(a, sumN, cmpN);
    if (li < n) {
        // [s][MROWS][cap]: the sums, then the Newtonian compensations
        const size_t cap = (size_t)P.cap;
        float* ps = part + (size_t)s * MROWS * cap + li;
#pragma unroll
        for (int k = 0; k < NACC32; ++k)
            if (TS || k != A_TS) stream_out(k <= A_NZ ? sumN[k] : a[k], &ps[k * cap]);
#pragma unroll
        for (int k = 0; k < 3; ++k) stream_out(cmpN[k], &ps[(NACC32 + k) * cap]);
    }
}

// The same pairs in workgroups of four i-waves x four consecutive j
// chunks (1024 threads): wave w takes the i-stars of i-wave w % 4 of block
// ib against chunk 4 s4 + w / 4 of the split in S chunks, so every wave
// does what it does in k_pair32 (and its FP32 sums are the same bits), but
// the four chunk sums of each star are combined in LDS and written as one
// partial of the split in S / 4 chunks: a quarter of k_pair32's partial-sum
// traffic.  The combination is in FP64: the Newtonian totals s - c are
// written back as a hi/lo float pair (s = hi, c = -lo, so that k_fin's
// s - c is the FP64 total to ~2^-48), the other sums rounded once to FP32.
constexpr int QSUB = 4;
constexpr int WGQ = QSUB * WG;
template <int PP, bool TS, bool EQ>
__global__ __launch_bounds__(WGQ) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_pair32q(const JRec32* __restrict__ jr, const JRec64* __restrict__ j64,
               const Thr* __restrict__ thr, const double* __restrict__ st,
               float* __restrict__ part, Phys P, int S, int nIblk, int ib0, int tile)
{
    __shared__ float red[QSUB - 1][MROWS][WG];
    const int S4 = S / QSUB;
    int s4, ib;
    map_block(blockIdx.x, S4, s4, ib);
    if (ib >= nIblk) return;                  // the whole workgroup
    ib += ib0;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int iw = wave & 3, sub = wave >> 2, lane = threadIdx.x & 63;
    const int n = P.n;
    const int li = ib * WG + iw * 64 + lane + 1;
    const int i = li < n ? li : n - 1;
    const I32 I = load_i32(j64, thr, st, P.cap, i, 0, P.eps2);
    // chunk 4 s4 + sub of the S-chunk split (k_pair32's bounds), and the
    // self window of this wave's 64 stars
    JRange R = j_range(s4 * QSUB + sub, S, n, ib);
    {
        const int wb = ib * WG + iw * 64 + 1;
        R.w0 = max(R.j0, wb);
        R.w1 = min(R.j1, wb + 64);
    }
    const float eps2 = P.eps2f;
    const double eps2d = P.eps2, rscale = 1.0 / P.mmax;
    float a[NACC32], sumN[3], cmpN[3];
#pragma unroll
    for (int k = 0; k < NACC32; ++k) a[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) sumN[k] = cmpN[k] = 0.0f;
    int cnt = 0;
    if (R.w0 < R.w1) {
        run32<false, PP, TS, EQ>(jr, j64, R.j0, R.w0, i, I, a, eps2, eps2d, rscale, sumN, cmpN, tile, cnt);
        run32<true, PP, TS, EQ>(jr, j64, R.w0, R.w1, i, I, a, eps2, eps2d, rscale, sumN, cmpN, tile, cnt);
        run32<false, PP, TS, EQ>(jr, j64, R.w1, R.j1, i, I, a, eps2, eps2d, rscale, sumN, cmpN, tile, cnt);
    } else if (R.j0 < R.j1) {
        run32<false, PP, TS, EQ>(jr, j64, R.j0, R.j1, i, I, a, eps2, eps2d, rscale, sumN, cmpN, tile, cnt);
    }
    fold_newton(a, sumN, cmpN);
    const int col = iw * 64 + lane;
    if (sub > 0) {
#pragma unroll
        for (int k = 0; k < NACC32; ++k)
            if (TS || k != A_TS) red[sub - 1][k][col] = k <= A_NZ ? sumN[k] : a[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) red[sub - 1][NACC32 + k][col] = cmpN[k];
    }
    __syncthreads();
    if (sub == 0 && li < n) {
        const size_t cap = (size_t)P.cap;
        float* ps = part + (size_t)s4 * MROWS * cap + li;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double t = (double)sumN[k] - (double)cmpN[k];
#pragma unroll
            for (int q = 0; q < QSUB - 1; ++q)
                t += (double)red[q][k][col] - (double)red[q][NACC32 + k][col];
            const float hi = (float)t;
            stream_out(hi, &ps[k * cap]);
            stream_out(-(float)(t - (double)hi), &ps[(NACC32 + k) * cap]);
        }
#pragma unroll
        for (int k = A_NZ + 1; k < NACC32; ++k) {
            if (!TS && k == A_TS) continue;
            double t = (double)a[k];
#pragma unroll
            for (int q = 0; q < QSUB - 1; ++q) t += (double)red[q][k][col];
            stream_out((float)t, &ps[k * cap]);
        }
    }
}

// ---------------------------------------------------------------------
// k_fin: star accelerations, stage update, BH partials
// ---------------------------------------------------------------------
// The pairs of every star i with the stars taken out of the FP32 pass
// (flist[0 .. nf), see PN1_FLAG_ETA), in FP64 with the FP64 form's body:
// per star the NACC sums and sum m_j w^2, written to xp [NACC + 1][stride]
// at idx.  Nothing when nf is 0.
__device__ __forceinline__ void flagged_pairs(const JRec64* __restrict__ jr,
                                              const double* __restrict__ st, int cap, int i,
                                              double eps2, const int* __restrict__ flist,
                                              int nf, double* __restrict__ xp, size_t stride,
                                              int idx)
{
    const I64 I = load_i64(jr, st, cap, i, eps2);
    double a[NACC], ts = 0.0;
#pragma unroll
    for (int k = 0; k < NACC; ++k) a[k] = 0.0;
    for (int f = 0; f < nf; ++f) {
        const int j = flist[f];
        body64<true>(jr[j], j, i, I, a, eps2);
        if (j != i) {
            const double rx = jr[j].x - I.x, ry = jr[j].y - I.y, rz = jr[j].z - I.z;
            ts += jr[j].m / (rx * rx + ry * ry + rz * rz + eps2);
        }
    }
#pragma unroll
    for (int k = 0; k < NACC; ++k) xp[(size_t)k * stride + idx] = a[k];
    xp[(size_t)NACC * stride + idx] = ts;
}

inline __global__ __launch_bounds__(WG) void k_pair_fl(const JRec64* __restrict__ jr,
                                               const double* __restrict__ st,
                                               const int* __restrict__ flist,
                                               const int* __restrict__ fn,
                                               double* __restrict__ xp, Phys P)
{
    const int nf = *fn;
    const int i = blockIdx.x * WG + threadIdx.x + 1;
    if (nf == 0 || i >= P.n) return;
    flagged_pairs(jr, st, P.cap, i, P.eps2, flist, nf, xp, (size_t)P.cap, i);
}

// FP64 sums of hard star i (see PN1_HARD_K) over the records [j0, j1)
// with the FP64 form's body (BH-close stars included, with their full
// records): the NACC sums and sum m_j w^2, reduced over the workgroup's HWG
// lanes in a fixed order and written to out[k * stride + idx], k <= NACC.
// Two records per lane in flight (the loads of a row are not shared by
// other lanes, so their latency is hidden only by having many in flight).
constexpr int HWG = 256;
// A hard star's row is split in HSEG segments of the j range, each a work
// unit of its own: one workgroup streaming a whole row (16 MB of records
// at n = 131073) took ~460 us on one CU, on the critical path of every
// mixed evaluation; 32 units on 32 CUs take ~1/32 of that.  The segments'
// sums go to a scratch of HCAP stars x HSEG x (NACC + 1) and are added in
// segment order (k_hard_sum).  An evaluation with more than HCAP hard
// stars takes every row in one unit: whether a row is split depends only
// on the count, never on a star's place in the list (which in an ensemble
// follows the order the clusters' workgroups appended them).
constexpr int HSEG = 32;
constexpr int HCAP = 1024;

__device__ __forceinline__ int hard_split(int nh) { return nh <= HCAP ? nh : 0; }

__device__ __forceinline__ void hard_row(const JRec64* __restrict__ jr,
                                         const double* __restrict__ st, int cap, double eps2,
                                         int i, int j0, int j1, double* __restrict__ out,
                                         size_t stride, size_t idx)
{
    __shared__ double wsum[NACC + 1][HWG / 64];
    const I64 I = load_i64(jr, st, cap, i, eps2);
    double a[NACC], ts = 0.0;
#pragma unroll
    for (int k = 0; k < NACC; ++k) a[k] = 0.0;
    auto pair = [&](const JRec64& J, int j) {
        body64<true>(J, j, i, I, a, eps2);
        if (j != i) {
            const double rx = J.x - I.x, ry = J.y - I.y, rz = J.z - I.z;
            ts += J.m / (rx * rx + ry * ry + rz * rz + eps2);
        }
    };
    for (int j = j0 + (int)threadIdx.x; j < j1; j += 2 * HWG) {
        const int jb = j + HWG;
        const JRec64 A = jr[j];
        const JRec64 B = jr[jb < j1 ? jb : j];
        pair(A, j);
        if (jb < j1) pair(B, jb);
    }
    double v[NACC + 1];
#pragma unroll
    for (int k = 0; k < NACC; ++k) v[k] = a[k];
    v[NACC] = ts;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k <= NACC; ++k) {
        double t = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        if (lane == 0) wsum[k][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x <= NACC) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < HWG / 64; ++w) t += wsum[threadIdx.x][w];
        out[(size_t)threadIdx.x * stride + idx] = t;
    }
    __syncthreads();
}

// Work unit u of the rows of the hard stars, ns = hard_split(nh) of them
// split (unit h * HSEG + g: star h, segment g, into hseg), for a star of a
// cluster whose stars are [p0, pend); the units past ns * HSEG are whole
// rows (into xh).
__device__ __forceinline__ void hard_unit(const JRec64* __restrict__ jr,
                                          const double* __restrict__ st, int cap, double eps2,
                                          int u, int ns, int star, int p0, int pend,
                                          double* __restrict__ xh, double* __restrict__ hseg)
{
    if (u < ns * HSEG) {
        const int g = u % HSEG, len = (pend - p0 + HSEG - 1) / HSEG;
        const int j0 = min(pend, p0 + g * len), j1 = min(pend, j0 + len);
        hard_row(jr, st, cap, eps2, star, j0, j1, hseg + (size_t)u * (NACC + 1), 1, 0);
    } else {
        hard_row(jr, st, cap, eps2, star, p0, pend, xh, (size_t)cap, (size_t)star);
    }
}

// The hard stars hlist[0 .. fn[1]): all in HSEG units each, or, beyond
// HCAP of them, all whole; the workgroups take the units in turn.
inline __global__ __launch_bounds__(HWG) void k_hard_rows(const JRec64* __restrict__ jr,
                                                  const double* __restrict__ st,
                                                  const int* __restrict__ hlist,
                                                  const int* __restrict__ fn,
                                                  double* __restrict__ xh,
                                                  double* __restrict__ hseg, Phys P)
{
    const int nh = fn[1], ns = hard_split(nh), nu = ns * HSEG + (nh - ns);
    for (int u = blockIdx.x; u < nu; u += gridDim.x) {
        const int h = u < ns * HSEG ? u / HSEG : ns + (u - ns * HSEG);
        hard_unit(jr, st, P.cap, P.eps2, u, ns, hlist[h], 1, P.n, xh, hseg);
    }
}

// Row value k of split hard star h (t = h * (NACC + 1) + k, h < ns): its
// HSEG segment sums added in segment order, into xh at the star.
__device__ __forceinline__ void hard_sum(const double* __restrict__ hseg, int t, int star,
                                         int cap, double* __restrict__ xh)
{
    const int h = t / (NACC + 1), k = t % (NACC + 1);
    const double* q = hseg + (size_t)h * HSEG * (NACC + 1) + k;
    double s = 0.0;
    for (int g = 0; g < HSEG; ++g) s += q[(size_t)g * (NACC + 1)];
    xh[(size_t)k * cap + star] = s;
}

inline __global__ __launch_bounds__(WG) void k_hard_sum(const int* __restrict__ hlist,
                                                const int* __restrict__ fn,
                                                const double* __restrict__ hseg,
                                                double* __restrict__ xh, int cap)
{
    const int t = blockIdx.x * WG + threadIdx.x;
    const int ns = hard_split(fn[1]);
    if (t < ns * (NACC + 1)) hard_sum(hseg, t, hlist[t / (NACC + 1)], cap, xh);
}

// Sum of the S chunk partials of one star.  `part` points at the
// cluster's partial area, `stride` is its per-sum stride and `idx` the
// star's index in it.  FP64: [S][NACC][stride]; mixed: [S][MROWS][stride]
// floats (the NACC32 sums, then the Newtonian Kahan compensations), whose
// sums built from the record's mass m_j / mscale are scaled back by mscale;
// ts is the mixed form's A_TS sum (0 for FP64, and without TS: the stages
// whose pair kernels do not write that row).
template <int MODE, bool TS = true>
__device__ __forceinline__ void gather_acc(const void* __restrict__ part, int S, size_t stride,
                                           int idx, double mscale, double (&A)[NACC], double& ts)
{
#pragma unroll
    for (int k = 0; k < NACC; ++k) A[k] = 0.0;
    ts = 0.0;
    if (MODE == 0) {
        const double* pp = (const double*)part;
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int k = 0; k < NACC; ++k) A[k] += pp[(size_t)(s * NACC + k) * stride + idx];
    } else {
        const float* pp = (const float*)part;
        for (int s = 0; s < S; ++s) {
            const float* ps = pp + (size_t)s * MROWS * stride + idx;
#pragma unroll
            for (int k = 0; k < NACC; ++k) A[k] += (double)ps[(size_t)k * stride];
            if (TS) ts += (double)ps[(size_t)A_TS * stride];
#pragma unroll
            for (int k = 0; k < 3; ++k) A[A_NX + k] -= (double)ps[(size_t)(NACC32 + k) * stride];
        }
#pragma unroll
        for (int k = 0; k < NACC; ++k)
            if (k <= A_UZ || k == A_KS || k == A_PHI || k == A_W3) A[k] *= mscale;
        ts *= mscale;
    }
}

// A hard star's FP64 row (k_hard_rows).
__device__ __forceinline__ void load_row(const double* __restrict__ xh, size_t stride, int idx,
                                         double (&A)[NACC], double& ts)
{
#pragma unroll
    for (int k = 0; k < NACC; ++k) A[k] = xh[(size_t)k * stride + idx];
    ts = xh[(size_t)NACC * stride + idx];
}

// Adds one star's FP64 sums of its pairs with the stars taken out of the
// FP32 pass (k_pair_fl) to its gathered sums.
__device__ __forceinline__ void add_flagged(const double* __restrict__ xp, size_t stride, int idx,
                                            double (&A)[NACC], double& ts)
{
#pragma unroll
    for (int k = 0; k < NACC; ++k) A[k] += xp[(size_t)k * stride + idx];
    ts += xp[(size_t)NACC * stride + idx];
}

// Mixed form: the star's cancellation ratio for the next evaluation's
// thresholds, sum m_j w^2 / |a| (+inf for a = 0).
__device__ __forceinline__ float cancel_ratio(double ts, const double (&a)[3])
{
    return (float)(ts / sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]));
}

// Run constants shared by the O(n) passes of one cluster.
struct FinC {
    double m0, eps, eps2, ws, lam, dt;
};

// Star p (BH of its cluster at index b0): its acceleration from its pair
// sums S_ and the cluster's separable sums gs, the stage update, and its
// share cb of the BH's 1PN acceleration (a1 + a1c).
template <int STAGE>
__device__ __forceinline__ void fin_star(const double* in, double* X, double* Y, double* H1,
                                         double* H2, const double* __restrict__ m, int cap,
                                         int p, int b0, const double (&S_)[NACC],
                                         const double* __restrict__ gs, const FinC& C,
                                         double* __restrict__ acc, double (&cb)[3],
                                         double (&a)[3])
{
    const double m0 = C.m0, mp = m[p], eps2 = C.eps2;
    const double x[3] = {in[p], in[cap + p], in[2 * cap + p]};
    const double v[3] = {in[3 * cap + p], in[4 * cap + p], in[5 * cap + p]};
    const double x0[3] = {in[b0], in[cap + b0], in[2 * cap + b0]};
    double d[3], D2, e;
    {
#pragma clang fp contract(off)
        for (int k = 0; k < 3; ++k) d[k] = x[k] - x0[k];
        D2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        e = 1.0 / sqrt(D2 + eps2);
    }
    const double E3 = e * e * e, q = mp * E3;
    const double G1 = gs[G_G1];
    const double M1[3] = {gs[G_M1X], gs[G_M1Y], gs[G_M1Z]};
    const double Pm[3] = {gs[G_PX], gs[G_PY], gs[G_PZ]};
    const double V2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double vd = v[0] * d[0] + v[1] * d[1] + v[2] * d[2];

    // ---- star i: Newton + B + S + sum over j != i of the ac pair term
    const double Pp[3] = {Pm[0] - mp * v[0], Pm[1] - mp * v[1], Pm[2] - mp * v[2]};
    const double vP = v[0] * Pp[0] + v[1] * Pp[1] + v[2] * Pp[2];
    const double Pd = Pp[0] * d[0] + Pp[1] * d[1] + Pp[2] * d[2];
    const double c1 = -m0 * E3 * (4.0 * S_[A_PHI] + 1.25 * (G1 - mp * e)) + E3 * S_[A_ZE];
    const double cB = 4.0 * m0 * e - V2;   // also the i-only part of the K_r bracket
    const double cS = 5.0 * m0 * mp * e * e * e * e - 4.0 * mp * E3 * V2;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double newton = S_[A_NX + k] - m0 * E3 * d[k];
        const double B = -m0 * E3 * (d[k] * cB + 4.0 * vd * v[k]);
        const double Sg = cS * d[k] + 7.0 * mp * E3 * vd * v[k];
        const double t4 = E3 * (4.0 * vP * d[k] - 3.0 * Pd * v[k] - 4.0 * vd * Pp[k]);
        const double ru = S_[A_UX + k] / PN1_SQ15 - v[k] * S_[A_KS];
        const double t3 = S_[A_TX + k] - 3.5 * m0 * e * (M1[k] - q * d[k]);
        const double rr = S_[A_RX + k] + cB * S_[A_NX + k];
        const double pn = B + Sg + rr + ru + c1 * d[k] + t3 + t4;
        a[k] = newton + C.lam * pn;
        acc[(size_t)k * cap + p] = a[k];
    }

    // ---- this star's share of the BH's 1PN acceleration: a1 + a1c
    // the j = i bracket of the a1c pair term in the reference's operation
    // order (Limit-sticky9.c:139-140 at |r|^2 = 0)
    double self_br;
    {
#pragma clang fp contract(off)
        const double ws = C.ws;
        self_br = 4.0 * e + 1.25 * ws - 0.25 * D2 / eps2 * ws + 0.25 * D2 / eps2 * ws;
    }
    const double vP0 = v[0] * Pm[0] + v[1] * Pm[1] + v[2] * Pm[2];
    const double Pd0 = Pm[0] * d[0] + Pm[1] * d[1] + Pm[2] * d[2];
    const double mE3 = mp * E3;
    const double cA1 = mE3 * (5.0 * m0 * e - 2.0 * V2 + 1.5 * vd * vd / D2);
    const double cC0 = 4.0 * mp * mp * e * e * e * e;
    // xi_i = sum' m_j w^3 (D2_i - D2_j) = -2 d_i.N - phi + eps^2 W3
    const double xi = -2.0 * (d[0] * S_[A_NX] + d[1] * S_[A_NY] + d[2] * S_[A_NZ])
                      - S_[A_PHI] + eps2 * S_[A_W3];
    const double cC2 = mE3 * (4.0 * (G1 - mp * e) + 1.25 * S_[A_PHI] + 0.25 * xi
                              + mp * self_br);
    const double cQ2 = mp * mp * E3 * (4.0 * e + 1.25 * C.eps);
    const double vm = mE3 / m0, vm2 = mp * mp * E3 / m0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double A1 = cA1 * d[k] + 3.0 * mE3 * vd * v[k];
        const double C2 = cC2 * d[k] - 3.5 * e * mp * S_[A_NX + k]
                          - vm * (4.0 * vP0 * d[k] - 3.0 * vd * Pm[k] - 4.0 * Pd0 * v[k]);
        const double Q2 = cQ2 * d[k] - vm2 * (4.0 * V2 * d[k] - 7.0 * vd * v[k]);
        cb[k] = A1 + cC0 * d[k] + C2 + Q2;
    }
    stage_update<STAGE>(p, cap, a, X, Y, H1, H2, C.dt);
}

__device__ __forceinline__ FinC fin_consts(const Phys& P, double m0)
{
    FinC C;
    C.m0 = m0;
    C.eps = P.eps;
    C.eps2 = P.eps2;
    C.ws = P.ws;
    C.lam = P.lam;
    C.dt = P.dt;
    return C;
}

// Stars p0 .. p1 - 1 (p0 = 1, p1 = n: every star; a sub-range is one i shard
// of a plain evaluation, STAGE 0); the BH mass is m[0].
// kout (mixed form, evaluations that accumulated A_TS; else null): each
// star's cancellation ratio, for the next evaluations' pair thresholds.
// Mixed form: plus the FP64 sums xp of the pairs with the stars taken out of
// the FP32 pass (when fn[0] > 0); a hard star (fl bit 1) takes its FP64 row
// xh instead of both.
template <int MODE, int STAGE>
__global__ __launch_bounds__(WG) void k_fin(const double* in, double* X, double* Y, double* H1,
                                           double* H2, const double* __restrict__ m,
                                           const void* __restrict__ part, int S,
                                           const double* __restrict__ gs,
                                           double* __restrict__ acc,
                                           double* __restrict__ bhblk, float* __restrict__ kout,
                                           const int* __restrict__ fn,
                                           const double* __restrict__ xp,
                                           const int* __restrict__ fl,
                                           const double* __restrict__ xh, Phys P, int p0,
                                           int p1)
{
    const int p = p0 + blockIdx.x * WG + threadIdx.x;
    double cb[3] = {0.0, 0.0, 0.0};
    if (p < p1) {
        double S_[NACC], ts, a[3];
        if (MODE == 1 && (fl[p] & 2)) {
            load_row(xh, (size_t)P.cap, p, S_, ts);
        } else {
            gather_acc<MODE, (STAGE <= 1)>(part, S, (size_t)P.cap, p, P.mmax, S_, ts);
            if (MODE == 1 && fn[0] > 0) add_flagged(xp, (size_t)P.cap, p, S_, ts);
        }
        fin_star<STAGE>(in, X, Y, H1, H2, m, P.cap, p, 0, S_, gs, fin_consts(P, m[0]), acc, cb, a);
        if (MODE == 1 && kout) kout[p] = cancel_ratio(ts, a);
    }
    block_sum<3>(cb, bhblk + (size_t)blockIdx.x * 3);
}

// BH at index b0: Newton sum_j m_j e_j^3 d_j plus lam times the star shares
// tot; stage update.  Thread 0 only.
template <int STAGE>
__device__ __forceinline__ void bh_finish(const double (&tot)[3], const double* __restrict__ gs,
                                          int b0, int cap, double lam, double dt, double* X,
                                          double* Y, double* H1, double* H2,
                                          double* __restrict__ acc)
{
    double a0[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        a0[k] = gs[G_M1X + k] + lam * tot[k];
        acc[(size_t)k * cap + b0] = a0[k];
    }
    stage_update<STAGE>(b0, cap, a0, X, Y, H1, H2, dt);
}

template <int STAGE>
__global__ __launch_bounds__(WG) void k_bh(const double* __restrict__ bhblk, int nblk,
                                          const double* __restrict__ gs, double* X, double* Y,
                                          double* H1, double* H2, double* __restrict__ acc,
                                          Phys P)
{
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += WG)
#pragma unroll
        for (int k = 0; k < 3; ++k) s[k] += bhblk[(size_t)b * 3 + k];
    __shared__ double tot[3];
    block_sum<3>(s, tot);
    if (threadIdx.x == 0) {
        const double t[3] = {tot[0], tot[1], tot[2]};
        bh_finish<STAGE>(t, gs, 0, P.cap, P.lam, P.dt, X, Y, H1, H2, acc);
    }
}

// ---------------------------------------------------------------------
// energy: pair sums (FP64 in both modes), per-star terms, totals
// ---------------------------------------------------------------------
// Energy pair body.  The record holds s_j = sqrt(3/2) v_j; the lane's
// velocity is pre-divided by sqrt(3/2) so that (r.v_i)(r.s_j) needs no
// rescaling, and k_efin divides sum m_j w s_j by sqrt(3/2).
struct IE64 {
    double x, y, z, ux, uy, uz;
};

template <bool MASK>
__device__ __forceinline__ void ebody64(const JRec64& J, int j, int i, const IE64& I,
                                        double (&a)[NACCE], double eps2)
{
    double mj = J.m;
    if (MASK) mj = (j == i) ? 0.0 : mj;
    const double rx = J.x - I.x, ry = J.y - I.y, rz = J.z - I.z;
    const double R2e = __builtin_fma(rx, rx, __builtin_fma(ry, ry, __builtin_fma(rz, rz, eps2)));
    const double w = rsq64(R2e);
    const double mw = mj * w;
    a[E_PHI] += mw;
    a[E_PVX] = __builtin_fma(mw, J.sx, a[E_PVX]);
    a[E_PVY] = __builtin_fma(mw, J.sy, a[E_PVY]);
    a[E_PVZ] = __builtin_fma(mw, J.sz, a[E_PVZ]);
    const double rvi = __builtin_fma(I.ux, rx, __builtin_fma(I.uy, ry, I.uz * rz));
    const double rvj = __builtin_fma(J.sx, rx, __builtin_fma(J.sy, ry, J.sz * rz));
    a[E_GAM] = __builtin_fma(mw * (w * w), rvi * rvj, a[E_GAM]);
}

template <bool MASK>
__device__ __forceinline__ void erun64(const JRec64* __restrict__ jr, int j0, int j1, int i,
                                       const IE64& I, double (&a)[NACCE], double eps2)
{
    if (j0 >= j1) return;
    JRec64 cur = jr[j0];
    for (int j = j0; j < j1; ++j) {
        const JRec64 nxt = jr[j + 1];
        ebody64<MASK>(cur, j, i, I, a, eps2);
        cur = nxt;
    }
}

inline __global__ __launch_bounds__(WG) void k_epair64(const JRec64* __restrict__ jr,
                                               const double* __restrict__ st,
                                               double* __restrict__ part, Phys P, int S,
                                               int nIblk)
{
    int s, ib;
    map_block(blockIdx.x, S, nIblk, s, ib);
    if (ib >= nIblk) return;
    const int n = P.n, cap = P.cap;
    const int li = ib * WG + threadIdx.x + 1;
    const int i = li < n ? li : n - 1;
    IE64 I;
    I.x = jr[i].x; I.y = jr[i].y; I.z = jr[i].z;
    I.ux = st[3 * cap + i] / PN1_SQ15;
    I.uy = st[4 * cap + i] / PN1_SQ15;
    I.uz = st[5 * cap + i] / PN1_SQ15;
    const JRange R = j_range(s, S, n, ib);
    const double eps2 = P.eps2;
    double a[NACCE];
#pragma unroll
    for (int k = 0; k < NACCE; ++k) a[k] = 0.0;
    if (R.w0 < R.w1) {
        erun64<false>(jr, R.j0, R.w0, i, I, a, eps2);
        erun64<true>(jr, R.w0, R.w1, i, I, a, eps2);
        erun64<false>(jr, R.w1, R.j1, i, I, a, eps2);
    } else {
        erun64<false>(jr, R.j0, R.j1, i, I, a, eps2);
    }
    if (li < n) {
#pragma unroll
        for (int k = 0; k < NACCE; ++k) stream_out(a[k], &part[(size_t)(s * NACCE + k) * cap + li]);
    }
}

// Per-star energy terms of Limit-sticky9.c:452-550 with the self pair
// excluded, for star p of the cluster whose BH is at b0, from its energy
// pair sums A: {m v^2, -1/2 m phi, pot_pn term, pot_pn2 term}.
__device__ __forceinline__ void efin_star(const double* __restrict__ st,
                                          const double* __restrict__ m, int cap, int p, int b0,
                                          double m0, double eps2, const double (&A)[NACCE],
                                          double G1, double (&t)[NEB])
{
    const double mp = m[p];
    const double x[3] = {st[p], st[cap + p], st[2 * cap + p]};
    const double v[3] = {st[3 * cap + p], st[4 * cap + p], st[5 * cap + p]};
    const double v0[3] = {st[3 * cap + b0], st[4 * cap + b0], st[5 * cap + b0]};
    double d[3], e;
    {
#pragma clang fp contract(off)
        for (int k = 0; k < 3; ++k) d[k] = x[k] - st[(size_t)k * cap + b0];
        e = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps2);
    }
    const double V2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double v0v = v0[0] * v[0] + v0[1] * v[1] + v0[2] * v[2];
    const double dv0 = d[0] * v0[0] + d[1] * v0[1] + d[2] * v0[2];
    const double dv = d[0] * v[0] + d[1] * v[1] + d[2] * v[2];
    const double vPhv = (v[0] * A[E_PVX] + v[1] * A[E_PVY] + v[2] * A[E_PVZ]) / PN1_SQ15;
    t[0] = mp * V2;
    t[1] = -0.5 * mp * A[E_PHI];
    t[2] = 0.375 * mp * V2 * V2 + 1.5 * m0 * e * mp * V2 + 0.5 * m0 * m0 * e * e * mp
           + 0.5 * m0 * e * mp * (3.0 * V2 - 7.0 * v0v - dv0 * dv * e * e);
    t[3] = mp * (0.25 * (6.0 * V2 * A[E_PHI] - 7.0 * vPhv - A[E_GAM]) + m0 * e * A[E_PHI]
                 + 0.5 * m0 * e * (G1 - mp * e));
}

inline __global__ __launch_bounds__(WG) void k_efin(const double* __restrict__ st,
                                            const double* __restrict__ m,
                                            const double* __restrict__ part, int S,
                                            const double* __restrict__ gs,
                                            double* __restrict__ eblk, Phys P)
{
    const int p = blockIdx.x * WG + threadIdx.x + 1;
    double t[NEB] = {0.0, 0.0, 0.0, 0.0};
    if (p < P.n) {
        double A[NACCE] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int k = 0; k < NACCE; ++k) A[k] += part[(size_t)(s * NACCE + k) * P.cap + p];
        efin_star(st, m, P.cap, p, 0, P.m0, P.eps2, A, gs[G_G1], t);
    }
    block_sum<NEB>(t, eblk + (size_t)blockIdx.x * NEB);
}

// out = {kin, pot, pot_pn/c^2, pot_pn2/c^2, total}
inline __global__ __launch_bounds__(WG) void k_ered(const double* __restrict__ eblk, int nblk,
                                            const double* __restrict__ st,
                                            const double* __restrict__ m,
                                            const double* __restrict__ gs,
                                            double* __restrict__ out, Phys P)
{
    double s[NEB] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += WG)
#pragma unroll
        for (int k = 0; k < NEB; ++k) s[k] += eblk[(size_t)b * NEB + k];
    __shared__ double tot[NEB];
    block_sum<NEB>(s, tot);
    if (threadIdx.x == 0) {
        const int cap = P.cap;
        const double v0x = st[3 * cap], v0y = st[4 * cap], v0z = st[5 * cap];
        const double kin = 0.5 * (m[0] * (v0x * v0x + v0y * v0y + v0z * v0z) + tot[0]);
        const double pot = tot[1] - P.m0 * gs[G_G1];
        const double pn = P.lam * tot[2], pn2 = P.lam * tot[3];
        out[0] = kin;
        out[1] = pot;
        out[2] = pn;
        out[3] = pn2;
        out[4] = kin + pot + pn + pn2;
    }
}

}  // namespace pn1dev
