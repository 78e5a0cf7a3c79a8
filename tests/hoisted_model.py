"""Numpy model of the decomposition the HIP kernels use (test infrastructure).

The reference evaluates three O(n^2) loops per force() call with seven
square roots per pair (Limit-sticky9.c:36-51, 116-174, 276-382).  The HIP
path evaluates ONE pair loop over stars with one rsqrt per pair and moves
everything separable into O(n) passes.  This module states that split in
plain numpy so the algebra can be checked against the oracle on the CPU
before any kernel runs; the kernels in pn_amd/csrc implement the same
formulas.

Notation (stars i, j in [1, n); index 0 is the black hole, BH):
  d_i = x_i - x_0, D2_i = |d_i|^2, e_i = (D2_i + eps^2)^-1/2, q_i = m_i e_i^3
  r = x_j - x_i, w = (|r|^2 + eps^2)^-1/2, u = v_j - v_i,  sum' = sum over j != i

Pair sums (per star i), the only O(n^2) work:
  Nst   = sum' m_j w^3 r                       (star-star Newton)
  PNr   = sum' K_r r,  K_r = m_j w^3 [m0 (4 e_i + 1.25 e_j + 0.25 e_j^3 (|r|^2 - D2_i))
                                      + V2_i - 2|u|^2 + 1.5 (v_j.r)^2 w^2]
          with V2_i - 2|u|^2 = -V2_i - 2 V2_j + 4 v_i.v_j
  RU, KS = sum' K_u v_j, sum' K_u,  K_u = m_j w^3 (r.(4 v_i - 3 v_j)),
          so that sum' K_u u = RU - v_i KS
  T3    = sum' 3.5 m0 q_j w d_j   (the O(n) pass subtracts e_i sum' 3.5 m0 q_j d_j)
  phi   = sum' m_j w,  W3 = sum' m_j w^3, and from them (xi_from_sums)
  Xi    = sum' m_j w^3 (D2_i - D2_j) = -2 d_i.Nst - phi + eps^2 W3
  zeta  = sum' m_j 0.25 m0 e_j^3 (|r|^2 - D2_i)
The (w - e_i) d_j term (Limit-sticky9.c:307-308) is summed over d_j, the
star-BH separation: re-expressing d_j as r + d_i instead would turn it
into large sums that cancel when a star sits close to the BH.
zeta and the 0.25 e_j^3 (|r|^2 - D2_i) part of K_r cancel each other for a
star j near the BH; both take the same per-pair |r|^2 - D2_i, as the
reference's two terms take the same r1b2 - rbc2 (Limit-sticky9.c:302, 305),
so the rounding of that difference cancels with them.
Energy pair sums: phi, Phv = sum' m_j w v_j, Gam = sum' m_j w^3 (r.v_i)(r.v_j).
"""
import numpy as np


def prep(x, v, m, eps):
    x0 = x[0]
    d = x[1:] - x0
    D2 = np.einsum("ij,ij->i", d, d)
    e = 1.0 / np.sqrt(D2 + eps * eps)
    return d, D2, e


def pair_sums(x, v, m, eps):
    """The O(n^2) accumulators of the force pair kernel (j = i skipped)."""
    m0 = m[0]
    xs, vs, ms = x[1:], v[1:], m[1:]
    d, D2, e = prep(x, v, m, eps)
    q = ms * e**3
    r = xs[None, :, :] - xs[:, None, :]            # [i, j, 3]
    R2 = np.einsum("ijk,ijk->ij", r, r)
    w = 1.0 / np.sqrt(R2 + eps * eps)
    np.fill_diagonal(w, 0.0)                       # skip j = i
    live = w > 0
    w2 = w * w
    mw = ms[None, :] * w
    mw3 = mw * w2
    V2 = np.einsum("ij,ij->i", vs, vs)
    vivj = vs @ vs.T
    vjr = np.einsum("jk,ijk->ij", vs, r)
    t = R2 - D2[:, None]
    # V2_i - 2|u|^2 = -V2_i - 2 V2_j + 4 v_i.v_j
    # the i-only part al_i = 4 m0 e_i - V2_i of the bracket is added in the
    # O(n) pass as al_i * Nst
    br = ((1.25 * m0 * e - 2 * V2)[None, :]
          + (0.25 * m0 * e**3)[None, :] * t + 4 * vivj + 1.5 * vjr**2 * w2)
    Kr = mw3 * br
    Ku = mw3 * (4 * np.einsum("ik,ijk->ij", vs, r) - 3 * vjr)
    qw = 3.5 * m0 * q[None, :] * w
    return dict(
        Nst=np.einsum("ij,ijk->ik", mw3, r),
        PNr=np.einsum("ij,ijk->ik", Kr, r),
        RU=Ku @ vs,
        KS=Ku.sum(1),
        T3=qw @ d,
        phi=mw.sum(1),
        W3=mw3.sum(1),
        Xi_direct=(mw3 * (D2[:, None] - D2[None, :])).sum(1),
        zeta=np.where(live, (0.25 * m0 * q)[None, :] * t, 0.0).sum(1),
    )


def xi_from_sums(d, Nst, phi, W3, eps):
    """Xi_i = sum' m_j w^3 (D2_i - D2_j) from sums the pair loop has anyway.
    With d_j = d_i + r, D2_i - D2_j = -2 d_i.r - |r|^2, and w^3 |r|^2 =
    w - eps^2 w^3 (w = (|r|^2 + eps^2)^-1/2), so
        Xi_i = -2 d_i.Nst_i - phi_i + eps^2 W3_i,  W3_i = sum' m_j w^3."""
    return -2.0 * np.einsum("ij,ij->i", d, Nst) - phi + eps * eps * W3


def force(x, v, m, eps, c):
    """a = a_N + a_PN / c^2 from the pair sums plus O(n) closed forms."""
    lam = 1.0 / (c * c)
    m0 = m[0]
    ms, vs = m[1:], v[1:]
    d, D2, e = prep(x, v, m, eps)
    E3 = e**3
    q = ms * E3
    V2 = np.einsum("ij,ij->i", vs, vs)
    vd = np.einsum("ij,ij->i", vs, d)
    G1 = (ms * e).sum()
    M1 = (q[:, None] * d).sum(0)
    P = (ms[:, None] * vs).sum(0)
    S = pair_sums(x, v, m, eps)
    Xi = xi_from_sums(d, S["Nst"], S["phi"], S["W3"], eps)

    a = np.zeros_like(x)
    # stars
    newton = S["Nst"] - (m0 * E3)[:, None] * d
    B = -(m0 * E3)[:, None] * (d * (4 * m0 * e - V2)[:, None] + 4 * vd[:, None] * vs)
    Sg = (5 * m0 * ms * e**4)[:, None] * d - (ms * E3)[:, None] * (
        4 * V2[:, None] * d - 7 * vd[:, None] * vs)
    G1p = G1 - ms * e
    Pp = P[None, :] - ms[:, None] * vs
    t1 = -(m0 * E3)[:, None] * d * (4 * S["phi"] + 1.25 * G1p)[:, None] \
        + (E3 * S["zeta"])[:, None] * d
    vP = np.einsum("ij,ij->i", vs, Pp)
    Pd = np.einsum("ij,ij->i", Pp, d)
    t4 = E3[:, None] * (4 * vP[:, None] * d - 3 * Pd[:, None] * vs - 4 * vd[:, None] * Pp)
    PNu = S["RU"] - vs * S["KS"][:, None]
    T3 = S["T3"] - (3.5 * m0 * e)[:, None] * (M1[None, :] - q[:, None] * d)
    PNr = S["PNr"] + (4 * m0 * e - V2)[:, None] * S["Nst"]
    X = PNr + PNu + t1 + T3 + t4
    a[1:] = newton + lam * (B + Sg + X)

    # black hole
    A1 = (ms * E3)[:, None] * d * (5 * m0 * e - 2 * V2 + 1.5 * vd**2 / D2)[:, None] \
        + (3 * ms * E3 * vd)[:, None] * vs
    C0 = (4 * ms**2 * e**4)[:, None] * d
    vP0 = vs @ P
    Pd0 = d @ P
    # the j = i term of the a1c pair bracket, in the reference's operation
    # order (Limit-sticky9.c:139-140 with |r|^2 + eps^2 = eps^2): its two
    # 0.25 D2/eps^2 terms cancel exactly but round on the way, and that
    # rounding is part of the reference's result
    ws = 1.0 / np.sqrt(eps * eps)
    eps2 = eps * eps
    self_br = ((4.0 * e + 1.25 * ws) - 0.25 * D2 / eps2 * ws) + 0.25 * D2 / eps2 * ws
    C2 = (ms * E3)[:, None] * d * (4 * (G1 - ms * e) + 1.25 * S["phi"] + 0.25 * Xi
                                   + ms * self_br)[:, None] \
        - (3.5 * e * ms)[:, None] * S["Nst"] \
        - (ms * E3 / m0)[:, None] * (4 * vP0[:, None] * d - 3 * vd[:, None] * P[None, :]
                                     - 4 * Pd0[:, None] * vs)
    Q2 = (ms**2 * E3)[:, None] * d * (4 * e + 1.25 * eps)[:, None] \
        - (ms**2 * E3 / m0)[:, None] * (4 * V2[:, None] * d - 7 * vd[:, None] * vs)
    a[0] = M1 + lam * (A1 + C0 + C2 + Q2).sum(0)
    return a


def energy_terms(x, v, m, eps, c):
    """{kin, pot, pot_pn/c^2, pot_pn2/c^2, total} from energy pair sums."""
    lam = 1.0 / (c * c)
    m0 = m[0]
    ms, vs = m[1:], v[1:]
    d, D2, e = prep(x, v, m, eps)
    xs = x[1:]
    r = xs[None, :, :] - xs[:, None, :]
    R2 = np.einsum("ijk,ijk->ij", r, r)
    w = 1.0 / np.sqrt(R2 + eps * eps)
    np.fill_diagonal(w, 0.0)
    mw = ms[None, :] * w
    phi = mw.sum(1)
    Phv = mw @ vs
    rvi = np.einsum("ijk,ik->ij", r, vs)
    rvj = np.einsum("ijk,jk->ij", r, vs)
    Gam = (mw * w * w * rvi * rvj).sum(1)
    V2 = np.einsum("ij,ij->i", vs, vs)
    G1 = (ms * e).sum()
    kin = 0.5 * (m * np.einsum("ij,ij->i", v, v)).sum()
    pot = -0.5 * (ms * phi).sum() - m0 * G1
    v0 = v[0]
    potpn = (0.375 * ms * V2 * V2 + 1.5 * m0 * e * ms * V2 + 0.5 * m0**2 * e**2 * ms
             + 0.5 * m0 * e * ms * (3 * V2 - 7 * (vs @ v0) - (d @ v0) * np.einsum(
                 "ij,ij->i", d, vs) * e**2)).sum()
    pn2 = (ms * (0.25 * (6 * V2 * phi - 7 * np.einsum("ij,ij->i", vs, Phv) - Gam)
                 + m0 * e * phi + 0.5 * m0 * e * (G1 - ms * e))).sum()
    total = kin + pot + lam * (potpn + pn2)
    return np.array([kin, pot, lam * potpn, lam * pn2, total])
