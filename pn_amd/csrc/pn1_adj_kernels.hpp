// pn1_adj_kernels.hpp -- device code of the force VJP (adjoint of force()).
//
// For a cotangent u (one 3-vector per particle) the VJP of the force is the
// gradient of  Phi = sum_p u_p . a_p(x, v; m0)  with respect to every x_p,
// v_p and the BH mass m0.  Phi splits into
//   pair terms  phi_ij, stars i, j >= 1 (self pair included, as in the
//               reference's sums): u_i . (Newton_ij + lam ac_ij)
//               + lam u_0 . a1c_ij       (Limit-sticky9.c:39-51, 118-174,
//                                          278-382)
//   star terms  psi_i, star i with the BH: Newton between star and BH both
//               ways, the BH's a1 / a1c prefix / a1c removal terms and the
//               star's BH 1PN, aa_cross and self-removal terms
//               (Limit-sticky9.c:54-115, 179-274, 386-422)
// Both are written once below as templates over the number type of each
// side.  The pair sums run in pn1_adj_rev.hpp: one hand-written reverse
// sweep of phi_ij per pair gives d/d(x_i, v_i, m0) and d/d(x_j, v_j)
// together (pairs j = i skipped).  The per-star terms run here with
// forward-mode duals seeded on the lane's own coordinates (the dual's
// tangents are the gradient):
//   fin     lanes i:   d (psi_i + phi_ii) / d(x_i, v_i, m0), plus the pair
//                      kernel's partial sums
// The self pair phi_ii is differentiated as one function of particle i (both
// sides carry the same duals, so r = x_i - x_i is zero with zero tangents):
// split into its i and j sides its partial derivatives would be +-m/eps^3
// terms cancelling to rounding.
// Every term depends on positions only through differences, so Phi is
// invariant under a common shift and the BH's position gradient is minus
// the sum of the stars'; no term depends on the BH's velocity.  All partial
// sums are added in a fixed order, so results are bitwise reproducible.
#pragma once
#include "pn1_common.hpp"
#include "pn1_dual.hpp"

namespace pn1dev {

constexpr int NADJ1 = 7;              // i side: x_i, v_i, m0
constexpr int NADJ2 = 6;              // j side: x_j, v_j

// Per-star adjoint record (FP64, 128 bytes): state, cotangent u, mass, BH
// geometry d = x - x_BH, D2 = |d|^2, e = (D2 + eps^2)^-1/2.
struct __align__(16) ARec {
    double x, y, z, vx, vy, vz, ux, uy, uz, m, D2, e, dx, dy, dz, pad;
};
static_assert(sizeof(ARec) == 128, "ARec layout");

// BH quantities every lane needs (uniform).
struct ABH {
    double x0[3], u0[3], m0;
};

// The BH of the cluster whose BH sits at index b0.
__device__ __forceinline__ ABH load_abh(const double* __restrict__ st,
                                        const double* __restrict__ mu,
                                        const double* __restrict__ m, int cap, int b0)
{
    ABH b;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b.x0[k] = st[(size_t)k * cap + b0];
        b.u0[k] = mu[(size_t)(3 + k) * cap + b0];
    }
    b.m0 = m[b0];
    return b;
}

template <class A, class B>
__host__ __device__ __forceinline__ auto dot(const A (&a)[3], const B (&b)[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// phi_ij for stars i, j.  i side: position, velocity, d_i = x_i - x_0,
// |d_i|^2, e_i, e_i^3, |v_i|^2; j side likewise; masses; m0 and 1/m0;
// cotangents u_i and u_0.  ac uses b = x_0 - x_i = -d_i, c = -d_j; a1c uses
// d_i, d_j; both use r = x_j - x_i.  1/(|r|^2 + eps^2) is written w^2.
template <class TI, class TJ, class TM>
__host__ __device__ __forceinline__ auto adj_pair_phi(const TI (&xi)[3], const TI (&vi)[3],
                                             const TI (&di)[3], const TI& rb2, const TI& rbe,
                                             const TI& rbe3, const TI& vvi,
                                             const TJ (&xj)[3], const TJ (&vj)[3],
                                             const TJ (&dj)[3], const TJ& rc2, const TJ& rce,
                                             const TJ& rce3, double mi, double mj, const TM& m0,
                                             const TM& im0, const double (&ui)[3],
                                             const double (&u0)[3], double eps2, double lam)
{
    using TR = decltype(xj[0] - xi[0]);
    TR r[3], w[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r[k] = xj[k] - xi[k];
        w[k] = vj[k] - vi[k];
    }
    const auto rbc2 = dot(r, r);
    const auto rbce = drsq(rbc2 + eps2);
    const auto rbce2 = rbce * rbce;
    const auto rbce3 = rbce2 * rbce;
    const auto mr3 = mj * rbce3;                       // Newton: m_j w^3
    const auto vivj = dot(vi, vj);
    const auto vjdi = dot(vj, di);
    const auto vidi = dot(vi, di);
    const auto vjr = dot(vj, r);
    const auto b5 = vvi - 2.0 * dot(w, w) + 1.5 * vjr * vjr * rbce2;
    const auto b6 = r[0] * (4.0 * vi[0] - 3.0 * vj[0]) + r[1] * (4.0 * vi[1] - 3.0 * vj[1])
                    + r[2] * (4.0 * vi[2] - 3.0 * vj[2]);
    const auto ui_r = dot(ui, r);
    const auto ui_di = dot(ui, di);
    const auto ui_dj = dot(ui, dj);
    const auto ui_vi = dot(ui, vi);
    const auto ui_vj = dot(ui, vj);
    const auto ui_w = dot(ui, w);
    const auto u0_r = dot(u0, r);
    const auto u0_di = dot(u0, di);
    const auto u0_vi = dot(u0, vi);
    const auto u0_vj = dot(u0, vj);
    // ac (Limit-sticky9.c:278-382), contracted with u_i
    const auto c1 = mj * rbe3 * m0;
    const auto bb1 = 4.0 * rbce + 1.25 * rce + 0.25 * (rb2 - rbc2) * rce3;
    const auto c2 = mr3 * m0;
    const auto bb2 = 4.0 * rbe + 1.25 * rce + 0.25 * (rbc2 - rb2) * rce3;
    const auto c3 = 3.5 * mj * m0 * rce3 * (rbce - rbe);
    const auto c4 = mj * rbe3;
    const auto uS = -(c1 * bb1) * ui_di + (c2 * bb2 + mr3 * b5) * ui_r + c3 * ui_dj
                    + c4 * (4.0 * vivj * ui_di - 3.0 * vjdi * ui_vi - 4.0 * vidi * ui_vj)
                    + mr3 * b6 * ui_w;
    // a1c (Limit-sticky9.c:118-174), contracted with u_0
    const auto mr3p = mi * rbe3;
    const auto br = 4.0 * rce + 1.25 * rbce + 0.25 * (rb2 - rc2) * rbce3;
    const auto cn = 3.5 * rbce3 * rbe * mi * mj;
    const auto cv = mr3p * mj * im0;
    const auto uT = mr3p * mj * br * u0_di - cn * u0_r
                    - cv * (4.0 * vivj * u0_di - 3.0 * vidi * u0_vj - 4.0 * vjdi * u0_vi);
    return mr3 * ui_r + lam * (uS + uT);
}

// psi_i for star i: d_i = x_i - x_0, |d_i|^2, e_i; velocity; masses.
template <class T, class TM>
__host__ __device__ __forceinline__ auto adj_star_psi(const T (&vi)[3], const T (&di)[3], const T& rb2,
                                             const T& rbe, double mi, const TM& m0,
                                             const TM& im0, const double (&ui)[3],
                                             const double (&u0)[3], double eps, double lam)
{
    const auto rbe3 = rbe * rbe * rbe;
    const auto vv = dot(vi, vi);
    const auto vd = dot(vi, di);
    const auto ui_di = dot(ui, di);
    const auto u0_di = dot(u0, di);
    const auto ui_vi = dot(ui, vi);
    const auto u0_vi = dot(u0, vi);
    const auto mr3 = mi * rbe3;
    const auto m0e3 = m0 * rbe3;
    // Newton: star <- BH, BH <- star (Limit-sticky9.c:39-51)
    const auto newton = -m0e3 * ui_di + mr3 * u0_di;
    // BH <- star 1PN a1 (:54-97), unsoftened 1/|d|^2
    const auto br3 = (5.0 * m0 * rbe - 2.0 * vv) + 1.5 * vd * vd * dinv(rb2);
    const auto t3 = mr3 * (br3 * u0_di + 3.0 * vd * u0_vi);
    // a1c prefix 4 m_i^2 e^4 d (:101-115)
    const auto t4 = 4.0 * mr3 * rbe * mi * u0_di;
    // a1c removal (:179-207), b = -d: subtracted
    const auto t5 = -mr3 * mi * u0_di * (4.0 * rbe + 1.25 * eps)
                    - mr3 * mi * im0 * (-4.0 * vv * u0_di + 7.0 * vd * u0_vi);
    // star <- BH 1PN (:215-238), b = -d
    const auto mri = m0 * rbe;
    const auto t6 = -m0e3 * (4.0 * mri - vv) * ui_di - 4.0 * m0e3 * vd * ui_vi;
    // aa_cross single-star term (:249-274)
    const auto t7 = 5.0 * m0e3 * mi * rbe * ui_di - mr3 * (4.0 * vv * ui_di - 7.0 * vd * ui_vi);
    // self-term removal (:386-422), b = -d: subtracted
    const auto br8 = 4.0 / eps + 1.25 * rbe + 0.25 * rb2 * rbe3;
    const auto t8 = -mr3 * m0 * br8 * ui_di + 3.5 * mr3 * m0 * (1.0 / eps - rbe) * ui_di
                    - mr3 * (-4.0 * vv * ui_di + 7.0 * vd * ui_vi);
    return newton + lam * (t3 + t4 - t5 + t6 + t7 - t8);
}

// Record of star p (BH at b0) of the state y (SoA, stride cap) with
// cotangent u = mu_v.
__device__ __forceinline__ void adj_prep_star(const double* __restrict__ y,
                                              const double* __restrict__ mu,
                                              const double* __restrict__ m, size_t cap,
                                              double eps2, int p, int b0,
                                              ARec* __restrict__ rec)
{
    ARec r;
    r.x = y[p]; r.y = y[cap + p]; r.z = y[2 * cap + p];
    r.vx = y[3 * cap + p]; r.vy = y[4 * cap + p]; r.vz = y[5 * cap + p];
    r.ux = mu[3 * cap + p]; r.uy = mu[4 * cap + p]; r.uz = mu[5 * cap + p];
    r.m = m[p];
    r.dx = r.x - y[b0]; r.dy = r.y - y[cap + b0]; r.dz = r.z - y[2 * cap + b0];
    r.D2 = r.dx * r.dx + r.dy * r.dy + r.dz * r.dz;
    r.e = rsq64(r.D2 + eps2);
    r.pad = 0.0;
    rec[p] = r;
}

__global__ __launch_bounds__(WG) void k_adj_prep(const double* __restrict__ y,
                                                const double* __restrict__ mu,
                                                const double* __restrict__ m, Phys P,
                                                ARec* __restrict__ rec)
{
    const int p = blockIdx.x * WG + threadIdx.x + 1;
    if (p < P.n) adj_prep_star(y, mu, m, (size_t)P.cap, P.eps2, p, 0, rec);
}

// d (psi_i + phi_ii) / d(x_i, v_i, m0) for star i with record `me`.  The
// self pair phi_ii is taken with both sides on the same variables: r = 0
// and its tangents cancel exactly, as in the reference's own sum.
__device__ __forceinline__ void adj_star_grad(const ARec& me, const ABH& B, double eps,
                                              double eps2, double lam, double (&g)[NADJ1])
{
    using D = Dual<NADJ1>;
    const D xi[3] = {dseed<NADJ1>(me.x, 0), dseed<NADJ1>(me.y, 1), dseed<NADJ1>(me.z, 2)};
    const D vi[3] = {dseed<NADJ1>(me.vx, 3), dseed<NADJ1>(me.vy, 4), dseed<NADJ1>(me.vz, 5)};
    const D m0 = dseed<NADJ1>(B.m0, 6);
    const D im0 = dinv(m0);
    const D di[3] = {xi[0] - B.x0[0], xi[1] - B.x0[1], xi[2] - B.x0[2]};
    const D rb2 = dot(di, di);
    const D rbe = drsq(rb2 + eps2);
    const double ume[3] = {me.ux, me.uy, me.uz};
    const D psi = adj_star_psi(vi, di, rb2, rbe, me.m, m0, im0, ume, B.u0, eps, lam);
    const D rbe3 = rbe * rbe * rbe;
    const D vvi = dot(vi, vi);
    const D self = adj_pair_phi(xi, vi, di, rb2, rbe, rbe3, vvi, xi, vi, di, rb2, rbe, rbe3,
                                me.m, me.m, m0, im0, ume, B.u0, eps2, lam);
#pragma unroll
    for (int k = 0; k < NADJ1; ++k) g[k] = psi.d[k] + self.d[k];
}

// out_x = g_x, out_v = mu_x + g_v for particle p; t = {g_x, d/dm0}.
__device__ __forceinline__ void adj_store(const double (&g)[NADJ1], const double* __restrict__ mu,
                                          double* __restrict__ out, size_t cap, int p,
                                          double (&t)[4])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        out[(size_t)k * cap + p] = g[k];
        out[(size_t)(3 + k) * cap + p] = mu[(size_t)k * cap + p] + g[3 + k];
        t[k] = g[k];
    }
    t[3] = g[6];
}

// Per star: the pair kernel's partials plus the star terms.  part holds S
// rows of NADJ1 i-side sums (one per j chunk) and then nIb rows of NADJ2
// j-side sums (one per 256-star i block), each row of stride cap.  Writes
// out_x = g_x and out_v = mu_x + g_v (the VJP of the stage function
// (x, v) -> (v, a)), and per-block partials {sum g_x, sum d/dm0}.
__global__ __launch_bounds__(WG) void k_adj_fin(const ARec* __restrict__ rec,
                                               const double* __restrict__ y,
                                               const double* __restrict__ mu,
                                               const double* __restrict__ m,
                                               const double* __restrict__ part, int S, int nIb,
                                               double* __restrict__ out,
                                               double* __restrict__ blk, Phys P)
{
    const int p = blockIdx.x * WG + threadIdx.x + 1;
    const size_t cap = (size_t)P.cap;
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    if (p < P.n) {
        const ABH B = load_abh(y, mu, m, P.cap, 0);
        double g[NADJ1];
        adj_star_grad(rec[p], B, P.eps, P.eps2, P.lam, g);
        for (int s = 0; s < S; ++s) {
#pragma unroll
            for (int k = 0; k < NADJ1; ++k) g[k] += part[(size_t)(s * NADJ1 + k) * cap + p];
        }
        const double* pj = part + (size_t)S * NADJ1 * cap;
        for (int b = 0; b < nIb; ++b) {
#pragma unroll
            for (int k = 0; k < NADJ2; ++k) g[k] += pj[(size_t)(b * NADJ2 + k) * cap + p];
        }
        adj_store(g, mu, out, cap, p, t);
    }
    block_sum<4>(t, blk + (size_t)blockIdx.x * 4);
}

// BH: out_x[0] = -(sum of the stars' g_x), out_v[0] = mu_x[0]; the m0
// derivative of this evaluation goes to gm0[1] and is added to gm0[0].
__global__ __launch_bounds__(WG) void k_adj_bh(const double* __restrict__ blk, int nblk,
                                              const double* __restrict__ mu,
                                              double* __restrict__ out,
                                              double* __restrict__ gm0, Phys P)
{
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += WG)
#pragma unroll
        for (int k = 0; k < 4; ++k) s[k] += blk[(size_t)b * 4 + k];
    __shared__ double tot[4];
    block_sum<4>(s, tot);
    if (threadIdx.x == 0) {
        const size_t cap = (size_t)P.cap;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            out[(size_t)k * cap] = -tot[k];
            out[(size_t)(3 + k) * cap] = mu[(size_t)k * cap];
        }
        gm0[1] = tot[3];
        gm0[0] += tot[3];
    }
}

// out = a A + b B + c C + d D over the 6 rows of n particles (stride cap);
// a null operand is skipped.  out may alias an operand.
__global__ __launch_bounds__(WG) void k_adj_lin(double* out, int n, int cap, double a,
                                               const double* A, double b, const double* B,
                                               double c, const double* C, double d,
                                               const double* D)
{
    const int p = blockIdx.x * WG + threadIdx.x;
    if (p >= n) return;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const size_t i = (size_t)k * cap + p;
        double r = a * A[i];
        if (B) r += b * B[i];
        if (C) r += c * C[i];
        if (D) r += d * D[i];
        out[i] = r;
    }
}

}  // namespace pn1dev
