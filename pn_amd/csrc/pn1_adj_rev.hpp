// pn1_adj_rev.hpp -- one reverse sweep of the pair term phi_ij that gives
// the i-side and the j-side gradient together (the single-pass force VJP).
//
// phi_ij (adj_pair_phi, Limit-sticky9.c:39-51, 118-174, 278-382) carries a
// factor m_j in every term: phi = m_j psi with
//   psi = w3 G + lam H,  w = (|r|^2 + eps^2)^-1/2, w2 = w^2, w3 = w^3,
//   G = (u_i.r) E + lam b6 (u_i.dv) - Kh (u_0.r),
//       E = 1 + lam (m0 bb2 + b5)
//   H = -Ka bb1 + Kb e_j^3 (w - e_i)(u_i.d_j) + Kc (v_i.v_j) + Kd (v_j.d_i)
//       + Ke (u_i.v_j) + Kf br + Kg (u_0.v_j)
// with r = x_j - x_i, dv = v_j - v_i, t = |d_i|^2 - |r|^2,
//   b5 = |v_i|^2 - 2|dv|^2 + 1.5 (v_j.r)^2 w2,  b6 = 4 v_i.r - 3 v_j.r,
//   bb1 = 4w + 1.25 e_j + 0.25 t e_j^3,  bb2 = 4 e_i + 1.25 e_j - 0.25 t e_j^3,
//   br = 4 e_j + 1.25 w + 0.25 (|d_i|^2 - |d_j|^2) w3,
// and the i-only coefficients (m_i, e_i, d_i, v_i, u_i; m0, 1/m0, u_0)
//   Ka = m0 e_i^3 (u_i.d_i)            Kb = 3.5 m0
//   Kc = 4 e_i^3 (u_i.d_i - m_i/m0 u_0.d_i)
//   Kd = -3 e_i^3 (u_i.v_i) + 4 m_i e_i^3/m0 (u_0.v_i)
//   Ke = -4 e_i^3 (v_i.d_i)            Kf = m_i e_i^3 (u_0.d_i)
//   Kg = 3 m_i e_i^3/m0 (v_i.d_i)      Kh = 3.5 lam m_i e_i.
// This is adj_pair_phi regrouped (the same terms, summed in another order).
//
// rev_pair differentiates psi by hand in reverse order, seeded with
// P = m_j (0 for a masked pair).  The gradient splits three ways:
//   * pair-level occurrences of x_i, v_i, x_j, v_j (through r, dv and the
//     dot products) go straight into the per-pair vector accumulators;
//   * the i-only scalars (e_i^3, u_i.d_i, ..., |d_i|^2, m0) enter psi
//     through the K's, E, bb2, t and Kh: their adjoints are linear in a
//     handful of pair sums (SA ... RB2), accumulated per pair and chained to
//     (x_i, v_i, m0) once per lane by rev_ifin;
//   * the j-only scalars |d_j|^2, e_j, e_j^3 fold into one accumulator rho
//     (the adjoint of |d_j|^2, through e_j and e_j^3), and u_0.v_j into u0v;
//     rev_jfin chains them to (x_j, v_j).
// Positions enter only through differences, so with d = x - x_BH the BH
// position gradient is minus the stars' sum (handled by the caller); the
// pair term does not depend on the BH velocity.
#pragma once
#include "pn1_adj_kernels.hpp"

namespace pn1dev {

// Lane constants of star i (read-only in the j loop).
struct RevI {
    double x[3], v[3], d[3], u[3];
    double rb2, rbe, vvi, uivi;
    double Ka, Kc, Kd, Ke, Kf, Kg, Kh;
};

// Per-lane sums over j of the i side.
struct RevIAcc {
    double gx[3], gv[3];
    double SA, SB, SBB, SC, SD, SE, SF, SG, SH, SM, SR, SW, RB2;
};

// The j record as it travels between lanes, and its sums over i.
struct RevJ {
    double x[3], v[3], rc2, rce, m;
};
struct RevJAcc {
    double gx[3], gv[3], rho, u0v;
};

// Uniform values of one evaluation.
struct RevU {
    double x0[3], u0[3], m0, im0, Kb, eps2, lam;
};

__host__ __device__ __forceinline__ RevU rev_uniform(const ABH& B, double eps2, double lam)
{
    RevU U;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        U.x0[k] = B.x0[k];
        U.u0[k] = B.u0[k];
    }
    U.m0 = B.m0;
    U.im0 = 1.0 / B.m0;
    U.Kb = 3.5 * B.m0;
    U.eps2 = eps2;
    U.lam = lam;
    return U;
}

__host__ __device__ __forceinline__ RevI rev_setup(const ARec& me, const RevU& U)
{
    RevI I;
    I.x[0] = me.x; I.x[1] = me.y; I.x[2] = me.z;
    I.v[0] = me.vx; I.v[1] = me.vy; I.v[2] = me.vz;
    I.d[0] = me.dx; I.d[1] = me.dy; I.d[2] = me.dz;
    I.u[0] = me.ux; I.u[1] = me.uy; I.u[2] = me.uz;
    I.rb2 = me.D2;
    I.rbe = me.e;
    const double rbe3 = me.e * me.e * me.e;
    const double uidi = dot(I.u, I.d), u0di = dot(U.u0, I.d);
    const double u0vi = dot(U.u0, I.v), vidi = dot(I.v, I.d);
    I.uivi = dot(I.u, I.v);
    I.vvi = dot(I.v, I.v);
    const double mi = me.m, mim0 = mi * U.im0;
    I.Ka = U.m0 * rbe3 * uidi;
    I.Kc = 4.0 * rbe3 * (uidi - mim0 * u0di);
    I.Kd = rbe3 * (-3.0 * I.uivi + 4.0 * mim0 * u0vi);
    I.Ke = -4.0 * rbe3 * vidi;
    I.Kf = mi * rbe3 * u0di;
    I.Kg = 3.0 * mim0 * rbe3 * vidi;
    I.Kh = 3.5 * U.lam * mi * me.e;
    return I;
}

__host__ __device__ __forceinline__ RevJ rev_jrec(const ARec& r)
{
    RevJ J;
    J.x[0] = r.x; J.x[1] = r.y; J.x[2] = r.z;
    J.v[0] = r.vx; J.v[1] = r.vy; J.v[2] = r.vz;
    J.rc2 = r.D2;
    J.rce = r.e;
    J.m = r.m;
    return J;
}

__host__ __device__ __forceinline__ void rev_zero(RevIAcc& a)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) a.gx[k] = a.gv[k] = 0.0;
    a.SA = a.SB = a.SBB = a.SC = a.SD = a.SE = a.SF = a.SG = 0.0;
    a.SH = a.SM = a.SR = a.SW = a.RB2 = 0.0;
}

__host__ __device__ __forceinline__ void rev_zero(RevJAcc& a)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) a.gx[k] = a.gv[k] = 0.0;
    a.rho = a.u0v = 0.0;
}

// One pair (i, j) with seed P: adds P d psi / d(...) to both sides.
__host__ __device__ __forceinline__ void rev_pair(const RevI& I, const RevJ& J, const RevU& U,
                                                  double P, RevIAcc& ai, RevJAcc& aj)
{
    const double lam = U.lam;
    // ---- forward ----
    double r[3], dv[3], dj[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r[k] = J.x[k] - I.x[k];
        dv[k] = J.v[k] - I.v[k];
        dj[k] = J.x[k] - U.x0[k];
    }
    const double rce = J.rce, rce2 = rce * rce, rce3 = rce2 * rce;
    const double rbc2 = dot(r, r);
    const double w = rsq64(rbc2 + U.eps2);
    const double w2 = w * w, w3 = w2 * w;
    const double vivj = dot(I.v, J.v), vjdi = dot(J.v, I.d), vjr = dot(J.v, r);
    const double vir = dot(I.v, r), dvdv = dot(dv, dv);
    const double uir = dot(I.u, r), uidj = dot(I.u, dj), uivj = dot(I.u, J.v);
    const double uiw = uivj - I.uivi;
    const double u0r = dot(U.u0, r), u0vj = dot(U.u0, J.v);
    const double b5 = I.vvi - 2.0 * dvdv + 1.5 * vjr * vjr * w2;
    const double b6 = 4.0 * vir - 3.0 * vjr;
    const double t = I.rb2 - rbc2, tb = I.rb2 - J.rc2;
    const double bb1 = 4.0 * w + 1.25 * rce + 0.25 * t * rce3;
    const double bb2 = 4.0 * I.rbe + 1.25 * rce - 0.25 * t * rce3;
    const double br = 4.0 * rce + 1.25 * w + 0.25 * tb * w3;
    const double wme = w - I.rbe;
    const double E = 1.0 + lam * (U.m0 * bb2 + b5);
    const double G = uir * E + lam * b6 * uiw - I.Kh * u0r;
    // ---- reverse: psi = w3 G + lam H, seeded with P ----
    const double Hb = lam * P, Gb = P * w3;
    double w3b = P * G;
    const double uirb = Gb * E;
    const double Gl = Gb * lam;
    const double bb2b = Gl * uir * U.m0;
    const double b5b = Gl * uir;
    const double b6b = Gl * uiw;
    const double uiwb = Gl * b6;
    const double u0rb = -Gb * I.Kh;
    const double bb1b = -Hb * I.Ka;
    const double HKb = Hb * U.Kb;
    double rce3b = HKb * wme * uidj;
    double wb = HKb * rce3 * uidj;
    const double uidjb = HKb * rce3 * wme;
    const double vivjb = Hb * I.Kc, vjdib = Hb * I.Kd, brb = Hb * I.Kf;
    double uivjb = Hb * I.Ke;
    const double u0vjb = Hb * I.Kg;
    // bb1, bb2, br
    wb += 4.0 * bb1b + 1.25 * brb;
    double rceb = 1.25 * (bb1b + bb2b) + 4.0 * brb;
    const double tbar = 0.25 * rce3 * (bb1b - bb2b);
    rce3b += 0.25 * t * (bb1b - bb2b);
    const double tbb = 0.25 * w3 * brb;
    w3b += 0.25 * tb * brb;
    // b5, b6, u_i.dv
    const double dvdvb = -2.0 * b5b;
    const double vjrb = 3.0 * vjr * w2 * b5b - 3.0 * b6b;
    double w2b = 1.5 * vjr * vjr * b5b;
    const double virb = 4.0 * b6b;
    uivjb += uiwb;
    // w3 = w2 w, w2 = w w, w = (rbc2 + eps^2)^-1/2
    w2b += w3b * w;
    wb += w3b * w2 + 2.0 * w * w2b;
    const double rbc2b = -tbar - 0.5 * w3 * wb;
    // j-only scalars: |d_j|^2 directly (tb) and through e_j, e_j^3
    rceb += 3.0 * rce2 * rce3b;
    aj.rho += -tbb - 0.5 * rce3 * rceb;
    aj.u0v += u0vjb;
    // ---- vectors ----
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double rb = 2.0 * rbc2b * r[k] + vjrb * J.v[k] + virb * I.v[k]
                          + uirb * I.u[k] + u0rb * U.u0[k];
        const double dvb = 2.0 * dvdvb * dv[k];
        const double vjb = vivjb * I.v[k] + vjdib * I.d[k] + vjrb * r[k] + uivjb * I.u[k];
        aj.gx[k] += rb + uidjb * I.u[k];
        aj.gv[k] += dvb + vjb;
        ai.gx[k] += vjdib * J.v[k] - rb;
        ai.gv[k] += vivjb * J.v[k] + virb * r[k] - dvb;
    }
    // ---- i-only scalar sums ----
    ai.SA += P * bb1;
    ai.SB += P * rce3 * wme * uidj;
    ai.SBB += P * rce3 * uidj;
    ai.SC += P * vivj;
    ai.SD += P * vjdi;
    ai.SE += P * uivj;
    ai.SF += P * br;
    ai.SG += P * u0vj;
    ai.SH += Gb * u0r;
    ai.SM += Gb * uir * bb2;
    ai.SR += Gb * uir;
    ai.SW += Gb * b6;
    ai.RB2 += tbar + tbb;
}

// i side of the lane's star: g = d/d(x_i, v_i, m0) of the pair sums,
// from the accumulators and the record `me`.
__host__ __device__ __forceinline__ void rev_ifin(const ARec& me, const RevU& U,
                                                  const RevIAcc& a, double (&g)[NADJ1])
{
    const double lam = U.lam, m0 = U.m0, im0 = U.im0, mi = me.m;
    const double vi[3] = {me.vx, me.vy, me.vz}, di[3] = {me.dx, me.dy, me.dz};
    const double ui[3] = {me.ux, me.uy, me.uz};
    const double rbe = me.e, rbe2 = rbe * rbe, rbe3 = rbe2 * rbe;
    const double uidi = dot(ui, di), u0di = dot(U.u0, di), u0vi = dot(U.u0, vi);
    const double vidi = dot(vi, di), uivi = dot(ui, vi);
    const double mim0 = mi * im0;
    // adjoints of the K's
    const double Kab = -lam * a.SA, Kbb = lam * a.SB, Kcb = lam * a.SC, Kdb = lam * a.SD;
    const double Keb = lam * a.SE, Kfb = lam * a.SF, Kgb = lam * a.SG, Khb = -a.SH;
    // m0 directly (E, Ka, Kb) and through 1/m0 (Kc, Kd, Kg)
    double m0b = lam * a.SM + 3.5 * Kbb + Kab * rbe3 * uidi;
    const double im0b = mi * rbe3 * (-4.0 * Kcb * u0di + 4.0 * Kdb * u0vi + 3.0 * Kgb * vidi);
    m0b -= im0 * im0 * im0b;
    const double rbe3b = Kab * m0 * uidi + Kcb * 4.0 * (uidi - mim0 * u0di)
                         + Kdb * (-3.0 * uivi + 4.0 * mim0 * u0vi) - 4.0 * Keb * vidi
                         + Kfb * mi * u0di + 3.0 * Kgb * mim0 * vidi;
    const double uidib = Kab * m0 * rbe3 + 4.0 * Kcb * rbe3;
    const double u0dib = -4.0 * Kcb * rbe3 * mim0 + Kfb * mi * rbe3;
    const double uivib = -3.0 * Kdb * rbe3 - lam * a.SW;
    const double u0vib = 4.0 * Kdb * mi * rbe3 * im0;
    const double vidib = -4.0 * Keb * rbe3 + 3.0 * Kgb * mim0 * rbe3;
    const double rbeb = 3.5 * lam * mi * Khb + 4.0 * lam * m0 * a.SR - lam * U.Kb * a.SBB
                        + 3.0 * rbe2 * rbe3b;
    const double rb2b = a.RB2 - 0.5 * rbe3 * rbeb;
    const double vvib = lam * a.SR;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double db = 2.0 * rb2b * di[k] + uidib * ui[k] + u0dib * U.u0[k] + vidib * vi[k];
        g[k] = a.gx[k] + db;
        g[3 + k] = a.gv[k] + 2.0 * vvib * vi[k] + uivib * ui[k] + u0vib * U.u0[k]
                   + vidib * di[k];
    }
    g[6] = m0b;
}

// j side of the record J: g = d/d(x_j, v_j) of the pair sums.
__host__ __device__ __forceinline__ void rev_jfin(const RevJ& J, const RevU& U, const RevJAcc& a,
                                                  double (&g)[NADJ2])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        g[k] = a.gx[k] + 2.0 * a.rho * (J.x[k] - U.x0[k]);
        g[3 + k] = a.gv[k] + a.u0v * U.u0[k];
    }
}


// =====================================================================
// Device: the single-pass pair kernel.
//
// A workgroup (4 waves) takes 256 stars i (one per lane, fixed) and a
// chunk of j records, in tiles of 64.  Each wave stages the tile's records
// in LDS (read-only for the tile).  Step t pairs lane l with record
// jt + ((l + t) mod 64), read from LDS, while that record's running j-side
// sums move one lane down after every step (a DPP lane rotation).  So each
// lane meets every record of the tile once and, after 64 steps, every
// record's sums are back on its home lane, summed over the wave's 64 stars.
// The four waves' sums are added in wave order through LDS and written per
// (i block, j); the i side is written per (j chunk, i).  Every order is
// fixed, so the result is bitwise reproducible.
// =====================================================================
// The value v of lane (lane + 1) mod 64: DPP wave_rol:1 on each half
// (v_mov_b32_dpp, a VALU move with no LDS round trip, so nothing waits on
// it; measured 12% faster for this kernel than ds_bpermute).  The records
// themselves are read from LDS instead (another 3%: their 18 moves per step
// leave the VALU).
__device__ __forceinline__ double lane_next(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x134, 0xf, 0xf, true);
    const unsigned hi =
        (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x134, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ void rev_next(RevJAcc& a)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        a.gx[k] = lane_next(a.gx[k]);
        a.gv[k] = lane_next(a.gv[k]);
    }
    a.rho = lane_next(a.rho);
    a.u0v = lane_next(a.u0v);
}

// One tile of records [jt, min(jt + 64, jend)) against the lane's star i
// (masked when !iok, and for j = i).  The i side accumulates into ai; g
// gets the tile's j-side sums of the lane's home record jt + lane.  tl is
// the wave's LDS tile, field-major (at step t the lanes read one field of
// records (l + t) mod 64, consecutive doubles: no bank conflicts).  Called
// by every thread of the workgroup (it holds a barrier); the caller's
// barriers after the tile keep the next tile's stores behind these reads.
constexpr int REVJ_F = 9;
// `live` (wave-uniform) is false for a wave wholly past the last star of a
// partly filled i block: its pairs all carry weight 0, so it skips the
// steps and its j-side sums stay zero (it still stages and meets the
// workgroup's barriers).
__device__ __forceinline__ void rev_tile(const ARec* __restrict__ rec, const RevI& I,
                                         const RevU& U, int i, bool iok, bool live, int jt,
                                         int jend, RevIAcc& ai, double (&g)[NADJ2],
                                         double (*tl)[64])
{
    const int lane = (int)(threadIdx.x & 63);
    const int jl = jt + lane;
    const RevJ J0 = rev_jrec(rec[jl < jend ? jl : jt]);
    tl[0][lane] = J0.x[0]; tl[1][lane] = J0.x[1]; tl[2][lane] = J0.x[2];
    tl[3][lane] = J0.v[0]; tl[4][lane] = J0.v[1]; tl[5][lane] = J0.v[2];
    tl[6][lane] = J0.rc2; tl[7][lane] = J0.rce; tl[8][lane] = J0.m;
    __syncthreads();
    RevJAcc aj;
    rev_zero(aj);
    if (live) {
        for (int t = 0; t < 64; ++t) {
            const int slot = (lane + t) & 63;
            RevJ J;
            J.x[0] = tl[0][slot]; J.x[1] = tl[1][slot]; J.x[2] = tl[2][slot];
            J.v[0] = tl[3][slot]; J.v[1] = tl[4][slot]; J.v[2] = tl[5][slot];
            J.rc2 = tl[6][slot]; J.rce = tl[7][slot]; J.m = tl[8][slot];
            const int j = jt + slot;
            const double Pm = (iok && j < jend && j != i) ? J.m : 0.0;
            rev_pair(I, J, U, Pm, ai, aj);
            rev_next(aj);
        }
    }
    rev_jfin(J0, U, aj, g);
}


// Block (s, ib): stars 1 + ib*WG + [0, WG), records [1 + s L, 1 + (s+1) L).
// part rows: i side (s*NADJ1 + k)*cap + i, s < S; j side
// (S*NADJ1 + ib*NADJ2 + k)*cap + j.
__global__ __launch_bounds__(WG) void k_adj_rev(const ARec* __restrict__ rec,
                                               const double* __restrict__ y,
                                               const double* __restrict__ mu,
                                               const double* __restrict__ m,
                                               double* __restrict__ part, Phys P, int L, int S)
{
    const int s = (int)blockIdx.x % S, ib = (int)blockIdx.x / S;
    const int n = P.n;
    const size_t cap = (size_t)P.cap;
    const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int li = ib * WG + (int)threadIdx.x + 1;
    const bool iok = li < n;
    const bool live = ib * WG + (__builtin_amdgcn_readfirstlane((int)threadIdx.x) & ~63) + 1 < n;
    const int i = iok ? li : n - 1;
    const ABH B = load_abh(y, mu, m, P.cap, 0);
    const RevU U = rev_uniform(B, P.eps2, P.lam);
    const ARec me = rec[i];
    const RevI I = rev_setup(me, U);
    RevIAcc ai;
    rev_zero(ai);
    const int j0 = 1 + s * L, j1 = min(n, j0 + L);
    double* pj = part + (size_t)S * NADJ1 * cap;
    __shared__ double sh[NADJ2][4][64];
    __shared__ double tls[4][REVJ_F][64];
    for (int jt = j0; jt < j1; jt += 64) {
        double g[NADJ2];
        rev_tile(rec, I, U, i, iok, live, jt, j1, ai, g, tls[wv]);
#pragma unroll
        for (int k = 0; k < NADJ2; ++k) sh[k][wv][lane] = g[k];
        __syncthreads();
        const int jl = jt + lane;
        if (wv == 0 && jl < j1) {
#pragma unroll
            for (int k = 0; k < NADJ2; ++k)
                pj[((size_t)ib * NADJ2 + k) * cap + jl] =
                    ((sh[k][0][lane] + sh[k][1][lane]) + sh[k][2][lane]) + sh[k][3][lane];
        }
        __syncthreads();
    }
    double g[NADJ1];
    rev_ifin(me, U, ai, g);
    if (iok) {
#pragma unroll
        for (int k = 0; k < NADJ1; ++k) part[((size_t)s * NADJ1 + k) * cap + li] = g[k];
    }
}

}  // namespace pn1dev
