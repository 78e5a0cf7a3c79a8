// CPU check of the single-pass pair adjoint (pn_amd/csrc/pn1_adj_rev.hpp).
//
// For random pairs (i, j) -- close pairs, stars close to the BH, c = 1 and
// c = 100 -- the hand-written reverse sweep (rev_pair, then rev_ifin /
// rev_jfin for the lane-level chains) must give the same 13 derivatives
//   d phi_ij / d(x_i, v_i, x_j, v_j, m0)
// as forward-mode duals (Dual<13>) of adj_pair_phi, the term-by-term pair
// function the GPU tests pin against the complex-step oracle.  A masked
// pair (seed 0) must give exact zeros.  Built by tests/test_adj_rev_host.py
// and run on the host; prints "worst <max relative error>".
#include <cmath>
#include <cstdio>
#include <random>

#include "pn1_adj_rev.hpp"

using namespace pn1dev;

static ARec make_rec(const double x[3], const double v[3], const double u[3], double m,
                     const double x0[3], double eps2)
{
    ARec r;
    r.x = x[0]; r.y = x[1]; r.z = x[2];
    r.vx = v[0]; r.vy = v[1]; r.vz = v[2];
    r.ux = u[0]; r.uy = u[1]; r.uz = u[2];
    r.m = m;
    r.dx = x[0] - x0[0]; r.dy = x[1] - x0[1]; r.dz = x[2] - x0[2];
    r.D2 = r.dx * r.dx + r.dy * r.dy + r.dz * r.dz;
    r.e = rsq64(r.D2 + eps2);
    r.pad = 0.0;
    return r;
}

int main(int argc, char** argv)
{
    const int trials = argc > 1 ? std::atoi(argv[1]) : 2000;
    std::mt19937_64 gen(7);
    std::normal_distribution<double> N(0.0, 1.0);
    std::uniform_real_distribution<double> U01(0.0, 1.0);
    double worst = 0.0;
    int nonzero_masked = 0;
    for (int trial = 0; trial < trials; ++trial) {
        double x0[3], u0[3], xi[3], vi[3], ui[3], xj[3], vj[3], uj[3];
        for (int k = 0; k < 3; ++k) {
            x0[k] = 0.1 * N(gen); u0[k] = N(gen);
            xi[k] = N(gen); vi[k] = 0.5 * N(gen); ui[k] = N(gen);
            vj[k] = 0.5 * N(gen); uj[k] = N(gen);
        }
        const double sep = trial % 4 == 0 ? 1e-3 : (trial % 4 == 1 ? 1e-1 : 1.0);
        for (int k = 0; k < 3; ++k) xj[k] = xi[k] + sep * N(gen);
        if (trial % 7 == 0)                       // star i close to the BH
            for (int k = 0; k < 3; ++k) xi[k] = x0[k] + 1e-2 * N(gen);
        const double mi = 0.01 + U01(gen), mj = 0.01 + U01(gen), m0 = 0.05 + U01(gen);
        const double eps = trial % 3 == 0 ? 1e-3 : 5e-2, eps2 = eps * eps;
        const double lam = trial % 2 ? 1.0 : 1e-4;
        const ARec I = make_rec(xi, vi, ui, mi, x0, eps2), J = make_rec(xj, vj, uj, mj, x0, eps2);

        // forward-mode duals: seeds x_i 0-2, v_i 3-5, x_j 6-8, v_j 9-11, m0 12
        using D = Dual<13>;
        D Dxi[3], Dvi[3], Dxj[3], Dvj[3];
        for (int k = 0; k < 3; ++k) {
            Dxi[k] = dseed<13>(xi[k], k);
            Dvi[k] = dseed<13>(vi[k], 3 + k);
            Dxj[k] = dseed<13>(xj[k], 6 + k);
            Dvj[k] = dseed<13>(vj[k], 9 + k);
        }
        const D Dm0 = dseed<13>(m0, 12), Dim0 = dinv(Dm0);
        const D di[3] = {Dxi[0] - x0[0], Dxi[1] - x0[1], Dxi[2] - x0[2]};
        const D dj[3] = {Dxj[0] - x0[0], Dxj[1] - x0[1], Dxj[2] - x0[2]};
        const D rb2 = dot(di, di), rbe = drsq(rb2 + eps2), rbe3 = rbe * rbe * rbe;
        const D vvi = dot(Dvi, Dvi);
        const D rc2 = dot(dj, dj), rce = drsq(rc2 + eps2), rce3 = rce * rce * rce;
        const D phi = adj_pair_phi(Dxi, Dvi, di, rb2, rbe, rbe3, vvi, Dxj, Dvj, dj, rc2, rce, rce3,
                                   mi, mj, Dm0, Dim0, ui, u0, eps2, lam);

        // reverse sweep
        ABH B;
        for (int k = 0; k < 3; ++k) {
            B.x0[k] = x0[k];
            B.u0[k] = u0[k];
        }
        B.m0 = m0;
        const RevU Uu = rev_uniform(B, eps2, lam);
        const RevI RI = rev_setup(I, Uu);
        const RevJ RJ = rev_jrec(J);
        RevIAcc ai;
        RevJAcc aj;
        rev_zero(ai);
        rev_zero(aj);
        rev_pair(RI, RJ, Uu, mj, ai, aj);
        double gi[NADJ1], gj[NADJ2];
        rev_ifin(I, Uu, ai, gi);
        rev_jfin(RJ, Uu, aj, gj);

        double ref[13], got[13];
        for (int k = 0; k < 6; ++k) {
            ref[k] = phi.d[k];         got[k] = gi[k];
            ref[6 + k] = phi.d[6 + k]; got[6 + k] = gj[k];
        }
        ref[12] = phi.d[12];
        got[12] = gi[6];
        double mx = 0.0, err = 0.0;
        for (int k = 0; k < 13; ++k) mx = std::fmax(mx, std::fabs(ref[k]));
        for (int k = 0; k < 13; ++k) err = std::fmax(err, std::fabs(ref[k] - got[k]));
        err /= mx;
        if (err > worst) worst = err;
        if (!(err <= 1e-12)) {
            std::printf("trial %d: relative error %.3e\n", trial, err);
            for (int k = 0; k < 13; ++k)
                std::printf("  %2d dual %+.15e reverse %+.15e\n", k, ref[k], got[k]);
        }

        // a masked pair contributes nothing
        RevIAcc zi;
        RevJAcc zj;
        rev_zero(zi);
        rev_zero(zj);
        rev_pair(RI, RJ, Uu, 0.0, zi, zj);
        double hi[NADJ1], hj[NADJ2];
        rev_ifin(I, Uu, zi, hi);
        rev_jfin(RJ, Uu, zj, hj);
        for (int k = 0; k < NADJ1; ++k) nonzero_masked += hi[k] != 0.0;
        for (int k = 0; k < NADJ2; ++k) nonzero_masked += hj[k] != 0.0;
    }
    std::printf("masked_nonzero %d\n", nonzero_masked);
    std::printf("worst %.3e\n", worst);
    return 0;
}
