// pn1_common.hpp -- definitions shared by the device code of libpn1
// (force, energy, ensemble and adjoint kernels): launch constants, the
// j-record layouts, run constants and reduction / mapping helpers.
#pragma once
#include <hip/hip_runtime.h>

namespace pn1dev {

constexpr int WG = 256;     // threads per workgroup (4 waves of 64)
constexpr int NACC = 16;    // force pair accumulators
constexpr int NACCE = 5;    // energy pair accumulators
constexpr int NGS = 7;      // separable O(n) sums
constexpr int NEB = 4;      // energy per-block partials

// sqrt(3/2): stored velocities are pre-scaled by it, so that the pair term
// 1.5 (v_j.r)^2 w^2 is a plain square of (s_j.r) w with s_j = sqrt(3/2) v_j.
#define PN1_SQ15 1.2247448713915890491
// 4 / sqrt(3/2) and 3 / sqrt(3/2) = sqrt(6)
#define PN1_4OSQ15 3.2659863237109041
#define PN1_SQ6 2.4494897427831781

// force pair accumulators (per star i, sums over stars j != i; t = |r|^2 - D2_i,
// s_j = sqrt(3/2) v_j, K_u = m_j w^3 (4 v_i.r - 3 v_j.r)):
//   N    = sum m_j w^3 r                                      (Newton)
//   Rr   = sum K_r r,  K_r = m_j w^3 [m0(4e_i + 1.25e_j + 0.25e_j^3 t)
//                                      + |v_i|^2 - 2|u|^2 + 1.5 (v_j.r)^2 w^2]
//          using |v_i|^2 - 2|u|^2 = -|v_i|^2 - 2|v_j|^2 + 4 v_i.v_j; the
//          i-only part al_i = 4 m0 e_i - |v_i|^2 is left out of the pair sum
//          and added by k_fin as al_i N
//   U    = sum K_u s_j,  KS = sum K_u      (sum K_u u = U / sqrt(3/2) - v_i KS)
//   T3   = sum 3.5 m0 m_j e_j^3 w d_j      (k_fin subtracts e_i times its w-free sum)
//   phi  = sum m_j w,  W3 = sum m_j w^3,  zeta = sum m_j 0.25 m0 e_j^3 t
//          (the BH's a1c needs xi = sum m_j w^3 (D2_i - D2_j); with d_j = d_i + r
//          and w^3 |r|^2 = w - eps^2 w^3 it is -2 d_i.N - phi + eps^2 W3, which
//          k_fin forms in FP64: one add per pair instead of a subtract and an FMA,
//          and no |d_j|^2 in the j-record)
// zeta shares t with K_r, as the reference's two e_j^3 terms share
// r1b2 - rbc2 (Limit-sticky9.c:302, 305); the two cancel strongly when star
// j is close to the BH, and the shared rounding cancels with them.  The
// (w - e_i) c_j term (:307-308) is summed over d_j = x_j - x_BH, never over
// r + d_i, which would cancel catastrophically in the same situation.
enum { A_NX, A_NY, A_NZ, A_RX, A_RY, A_RZ, A_UX, A_UY, A_UZ, A_TX, A_TY, A_TZ,
       A_KS, A_PHI, A_W3, A_ZE };
// The mixed form keeps one more sum, A_TS = sum m_j w^2 (>= sum of the
// Newtonian terms' magnitudes, which bounds how far the FP32 rounding of
// those terms can reach into a star's acceleration), and writes per chunk
// and star MROWS floats: the NACC32 sums, then the three Kahan
// compensations of the Newtonian sums.
constexpr int A_TS = NACC;
constexpr int NACC32 = NACC + 1;
constexpr int MROWS = NACC32 + 3;
// energy pair accumulators: phi, sum m_j w s_j, sum m_j w^3 (r.v_i)(r.s_j)
enum { E_PHI, E_PVX, E_PVY, E_PVZ, E_GAM };
// separable sums over stars: sum m e, sum m e^3 d, sum m v
enum { G_G1, G_M1X, G_M1Y, G_M1Z, G_PX, G_PY, G_PZ };

// j-record, FP64 form, 128 bytes (two 16-dword scalar loads):
//   x            position                 s = sqrt(3/2) v
//   m, D2        mass, |d|^2, d = x - x_BH
//   A            1.25 m0 e - 2 |v|^2       B = 0.25 m0 e^3
//   MB = m B     Qd = Q d, Q = 3.5 m0 m e^3 (the pair term Q w d is then one FMA
//                per component)            e = (|d|^2 + eps^2)^-1/2
struct __align__(16) JRec64 {
    double x, y, z, sx, sy, sz, m, D2, A, B, MB, Qdx, Qdy, Qdz, e, pad;
};
// j-record, mixed form: double-single position (hi, lo), the rest rounded
// to FP32, and Qd = Q d (the pair term Q w d is then one FMA per component).
// m is the star's mass divided by the cluster's largest star mass (mscale):
// 1 exactly when the masses are equal, so the masses carry no rounding that
// every FP32 pair term would share; the sums built from it (Newton, Rr, U,
// KS, phi, W3, A_TS) are scaled back by mscale in FP64.
// 64 bytes: one 16-dword scalar load.
struct __align__(16) JRec32 {
    float xh, yh, zh, xl, yl, zl, sx, sy, sz, m, A, B, MB, Qdx, Qdy, Qdz;
};
static_assert(sizeof(JRec64) == 128, "JRec64 layout");
static_assert(sizeof(JRec32) == 64, "JRec32 layout");

// Everything a kernel needs about the run, passed by value.
struct Phys {
    double eps, eps2, lam, dt, m0;
    double ws;       // 1/sqrt(eps^2): the reference's self-pair w
    double mmax;     // largest |m| of the stars (mixed form; 0: all massless)
    float eps2f, m0f;
    int n, cap;      // particles incl. the BH; SoA stride
};

// Mixed form: per-star pair thresholds, from the star's previous
// acceleration |a| (the last evaluation of the context), compared with the
// FP32 R2e = |r|^2 + eps^2 of each pair.
//   rc2  a pair with R2e < rc2 can pull harder than PN1_CLOSE_KAPPA |a|
//        (m_j / R2e > kappa |a|).  Its Newtonian term is evaluated in FP64
//        and added to the compensated sum: a neighbour whose pull is several
//        times the star's net acceleration would otherwise bring the FP32
//        rounding of its one term (~3 u) into the result at that multiple.
//   dn2  below it the separation needs the lo parts of the double-single
//        positions, which are taken relative to the cluster's BH (x below:
//        the star's offset from it).  With the hi parts alone each position
//        is off by up to u |x| (u = 2^-24), which moves a pair's Newtonian
//        term by up to 6 u |x| m_j / |r|^3; dn2 is where that reaches
//        PN1_LAZY_ETA |a| (|r|^3 = 6 u |x| mmax / (eta |a|)).  Farther pairs
//        keep the hi parts only: their position errors act as a displacement
//        of the star by ~u |x| in the smooth far field, and a few u of that
//        field.  (Absolute positions would make that displacement grow with
//        the cluster's distance from the origin: a cluster 1000 units away
//        measured 1.4e-6.)
// A star whose previous evaluation had mscale * A_TS > PN1_HARD_K |a| (its
// pair terms cancel to 1/30 of their magnitudes or less) takes every pair in
// FP64: the FP32 rounding of the terms, ~0.1 u of their magnitudes once
// averaged, would otherwise near 1e-6 of |a|.  With MixPrep::hard_rows the
// star's whole row is evaluated in FP64 by a pass of its own (k_hard_rows,
// one workgroup per such star) and its FP32 lane does no extra work
// (rc2 = dn2 = 0; k_fin takes the FP64 row instead); otherwise its FP32 lane
// takes the FP64 branch for every pair (rc2 = FLT_MAX), which makes its whole
// wave several times slower than the others.
// Without a previous evaluation (after an upload) the thresholds are
// rc2 = 0 and dn2 = FLT_MAX: no FP64 pairs, lo parts everywhere.
#define PN1_CLOSE_KAPPA 0.25
#define PN1_LAZY_ETA 1e-8
#define PN1_HARD_K 30.0
struct __align__(8) Thr {
    float rc2, dn2;
};

// Mixed form: stars close to the BH.  A star j at distance |d_j| from the
// BH enters every other star's 1PN pair terms with factors e_j^3 ~ |d_j|^-3
// (Limit-sticky9.c:302-308): 0.25 m0 e_j^3 (|r|^2 - |d_i|^2) and
// 3.5 m0 m_j e_j^3 (w - e_i) c_j, each of which cancels (against the
// zeta term and against e_i c_j) to a small result only after the sums.
// In FP32 sums their rounding reaches ~6 u lam m_j e_j^2 |d_i| of a star
// i's acceleration (|a_i| >= m0 / |d_i|^2).  A star with
// lam m_j e_j^2 > tauf, tauf = PN1_FLAG_ETA / (6 u D), D twice the
// farthest star from the BH at upload, is taken out of the FP32 pass (its
// mixed record is neutralised: far away, zero mass, zero 1PN fields) and
// its pairs with every star are evaluated in FP64 by k_pair_fl, whose sums
// k_fin adds.  Such stars are rare (none in the benchmark's Plummer sphere
// at N = 131072 within |d| < 3e-3).
#define PN1_FLAG_ETA 1e-8
#define PN1_FAR 1e18f
// What k_prep needs for the mixed form besides the FP64 records.
struct MixPrep {
    JRec32* j32;            // null: FP64 form (nothing below is used)
    Thr* thr;
    const double* aprev;    // previous evaluation (both null after an upload)
    const float* kprev;
    double mmax, tauf;      // the cluster's largest |m| of its stars; BH-close threshold
    double lam;
    int hard_rows;          // stars beyond PN1_HARD_K get FP64 rows of their own
};

// ---------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------
// FP64 reciprocal square root.  v_rsq_f64 alone is good to ~2^-24
// relative (measured on gfx950), so one correction of third order follows:
// with r = 1 - x y^2, 1/sqrt(x) = y (1 - r)^(-1/2) = y (1 + r/2 + 3r^2/8 + ...),
// truncation ~ (5/16) r^3 ~ 1e-22; the result is within ~1 ulp.  The host
// pass (CPU checks of the device functions) uses 1 / sqrt.
__host__ __device__ __forceinline__ double rsq64(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    const double h = x * y;
    const double r = __builtin_fma(-h, y, 1.0);
    const double t = __builtin_fma(0.375, r, 0.5);
    return __builtin_fma(y * r, t, y);
#else
    return 1.0 / __builtin_sqrt(x);
#endif
}

// Fixed-order workgroup sum of N doubles; thread 0 writes out[0..N).
// Wave level: xor butterfly (every lane ends with the same bits); then the
// four wave totals are added in wave order.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* out)
{
    __shared__ double sh[N][WG / 64];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double t = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        v[k] = t;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) sh[k][wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) out[k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
    }
}

// Store of a pair kernel's partial sum, read once by the finishing kernel:
// non-temporal, so that the partials streaming out do not evict the
// j-records from L2 (measured on MI355X at n = 131073: the FP64 pair
// kernel's fetches fall from 1.03 GB to 0.18 GB per launch).
template <typename T>
__device__ __forceinline__ void stream_out(T v, T* p)
{
    __builtin_nontemporal_store(v, p);
}

// Tapered j split.  A split in S >= 2 PN1_TAPER_N chunks (S a multiple of
// 8) ends in PN1_TAPER_N short chunks, each 1 / PN1_TAPER_F as long as the
// others, and a pair launch runs the long chunks of all its i blocks before
// any short one.  Equal workgroups leave the chip idling at the end of a
// launch for about half a workgroup's duration, waiting for the last ones;
// with short workgroups last, that wait shrinks in proportion (measured on
// MI355X at n = 131073, S = 64: ~1.5% shorter mixed and FP64 steps).
#define PN1_TAPER_N 16
#define PN1_TAPER_F 8

__device__ __forceinline__ bool tapered(int S)
{
    return S >= 2 * PN1_TAPER_N && (S & 7) == 0;
}

// Workgroup -> (j chunk s, i block ib).  When S divides 8, consecutive
// workgroup ids (which the dispatcher deals round-robin to the 8 XCDs) are
// arranged so that every workgroup on XCD x works on chunk x % S; for
// larger S (multiples of 8) the chunk is bid % S, also x modulo 8.
__device__ __forceinline__ void map_block(int bid, int S, int& s, int& ib)
{
    if (S <= 8 && (8 % S) == 0) {
        const int xcd = bid & 7, k = bid >> 3, per = 8 / S;
        s = xcd % S;
        ib = k * per + xcd / S;
    } else {
        s = bid % S;
        ib = bid / S;
    }
}

// The same for a launch over nib i blocks with the tapered split: first
// the S - PN1_TAPER_N long chunks of every block, then the short ones
// (both parts keep chunk = XCD modulo 8).
__device__ __forceinline__ void map_block(int bid, int S, int nib, int& s, int& ib)
{
    if (!tapered(S)) {
        map_block(bid, S, s, ib);
        return;
    }
    const int Sm = S - PN1_TAPER_N, nm = Sm * nib;
    if (bid < nm) {
        s = bid % Sm;
        ib = bid / Sm;
    } else {
        const int t = bid - nm;
        s = Sm + t % PN1_TAPER_N;
        ib = t / PN1_TAPER_N;
    }
}

// Chunk bounds of the j range [1, n) split in S parts (tapered: S - N long
// chunks of length cm, N short ones of ceil(cm / F), (S - N + N / F) cm >=
// n - 1; the last short chunks may be empty), and the part of the chunk
// that can contain j == i for this wave's 64 lanes.
struct JRange {
    int j0, j1, w0, w1;
};
__device__ __forceinline__ JRange j_range(int s, int S, int n, int ib)
{
    JRange R;
    const int nst = n - 1;
    if (!tapered(S)) {
        const int clen = (nst + S - 1) / S;
        R.j0 = 1 + s * clen;
        R.j1 = min(n, R.j0 + clen);
    } else {
        constexpr int eq = PN1_TAPER_N / PN1_TAPER_F;   // short chunks per long one
        const int Sm = S - PN1_TAPER_N;
        const int cm = (nst + Sm + eq - 1) / (Sm + eq);
        const int ct = (cm + PN1_TAPER_F - 1) / PN1_TAPER_F;
        const int a = s < Sm ? 1 + s * cm : 1 + Sm * cm + (s - Sm) * ct;
        R.j0 = min(n, a);
        R.j1 = min(n, a + (s < Sm ? cm : ct));
    }
    const int wb = ib * WG + (__builtin_amdgcn_readfirstlane((int)threadIdx.x) & ~63) + 1;
    R.w0 = max(R.j0, wb);
    R.w1 = min(R.j1, wb + 64);
    return R;
}

}  // namespace pn1dev
