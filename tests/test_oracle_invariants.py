"""Closed forms, invariants and brute force that pin the CPU restatement
where the reference's own golden vectors leave it open (Philox, link
overrelaxation, correlators) and cross-check the rest with mathematics that a
dropped term, a wrong sign or a transposed operand would violate."""
import numpy as np
import pytest
from scipy import integrate, special

from helpers import (checkerboard, coords, from_mat, higgs_mat, lattice_fields,
                     plaquette_sum, rand_conf, to_mat)


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def dag(m):
    return np.conj(np.swapaxes(m, -1, -2))


def extract(m):
    """adjoint vector (a1,a2,a3) = (Im beta, Re beta, Im alpha) of an algebra
    element (f/a_force.f:99-107)"""
    return np.stack([m[..., 0, 1].imag, m[..., 0, 1].real, m[..., 0, 0].imag], axis=-1)


# ------------------------------------------------------------------ Philox

def test_philox_known_answers(O):
    """Random123 known-answer vectors for philox4x32-10."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, exp in kat:
        assert tuple(int(x) for x in O.philox4x32_10(ctr, key)) == exp


def test_philox_uniform_stream_statistics(O):
    """hot start from the production stream: Haar links (unit quaternions with
    E[c_k^2] = 1/4) and N(0, sigma^2) Higgs components."""
    L = O.OracleLattice(32, 32, 32)
    L.hotstart(5, 0.5)
    u, a = L.get_links(), L.get_higgs()
    assert np.max(np.abs(np.sum(u * u, -1) - 1)) < 1e-15
    n = u.size // 4
    for c in range(4):
        assert abs(np.mean(u[..., c] ** 2) - 0.25) < 5 * np.sqrt(1 / 12 / 8) / np.sqrt(n) * 4
        assert abs(np.mean(u[..., c])) < 5 * 0.5 / np.sqrt(n)
    assert abs(np.mean(a)) < 5 * 0.5 / np.sqrt(a.size)
    assert abs(np.var(a) - 0.25) < 5 * 0.25 * np.sqrt(2 / a.size)


# ------------------------------------------------------------- SU(2) algebra

def test_multall_brute_force(O):
    rng = np.random.default_rng(0)
    n = 64
    p = rng.normal(size=(n, 4))
    q = rng.normal(size=(n, 4))
    ah = rng.normal(size=(n, 3))
    s1 = rng.permutation(n).astype(np.int32)
    s2 = rng.permutation(n).astype(np.int32)
    P, Q = to_mat(p[s1]), to_mat(q[s2])
    exp = {"nn": P @ Q, "dn": dag(P) @ Q, "nd": P @ dag(Q), "dd": dag(P) @ dag(Q)}
    for k, m in exp.items():
        np.testing.assert_allclose(O.multall(k, p, q, s1, s2), from_mat(m), atol=1e-13)
    X = higgs_mat(ah[s1])
    np.testing.assert_allclose(O.multall("ad", ah, q, s1, s2), from_mat(X @ dag(Q)), atol=1e-13)
    np.testing.assert_allclose(O.multall("an", ah, q, s1, s2), from_mat(X @ Q), atol=1e-13)


# ----------------------------------------------------------------- geometry

@pytest.mark.parametrize("shape", [(2, 2, 2), (2, 4, 6), (4, 4, 4), (6, 4, 8), (8, 2, 4)])
def test_geometry(O, shape):
    """f/adresses.f tables == closed form == brute-force coordinate shifts."""
    L = O.OracleLattice(*shape)
    up, dn = L.site_tables()
    upc, dnc = L.site_tables(closed_form=True)
    assert np.array_equal(up, upc) and np.array_equal(dn, dnc)
    cs = coords(shape)
    ext = np.array(shape)
    for p in (0, 1):
        for mu in range(3):
            e = np.eye(3, dtype=int)[mu]
            for tab, sgn in ((up, 1), (dn, -1)):
                c = (cs[p] + sgn * e) % ext
                # index of c among the other parity's coordinates
                lut = {tuple(x): i for i, x in enumerate(cs[1 - p])}
                exp = np.array([lut[tuple(x)] for x in c])
                assert np.array_equal(tab[p, mu], exp)


# ------------------------------------------------------------------ staples

def _brute_staple(shape, u, mu):
    U, _ = lattice_fields(shape, u, np.zeros((2, u.shape[2], 3)))
    V = np.zeros(U.shape[1:], dtype=complex)
    for nu in range(3):
        if nu == mu:
            continue
        Un, Um = U[nu], U[mu]
        top = np.roll(Un, -1, axis=mu) @ dag(np.roll(Um, -1, axis=nu)) @ dag(Un)
        bot = dag(np.roll(np.roll(Un, -1, axis=mu), 1, axis=nu)) @ dag(np.roll(Um, 1, axis=nu)) \
            @ np.roll(Un, 1, axis=nu)
        V += top + bot
    return V


def test_staples_brute_force_and_plaquette_identity(O):
    shape = (4, 6, 4)
    u, a = rand_conf(shape, 3)
    L = O.OracleLattice(*shape)
    L.put_links(u)
    tot = 0.0
    for mu in range(3):
        Vb = _brute_staple(shape, u, mu)
        vb, _ = checkerboard(shape, np.stack([Vb] * 3), None)
        for eo in (0, 1):
            L.staples(eo, mu + 1)
            np.testing.assert_allclose(L.get_staple()[mu, eo], vb[mu, eo], atol=1e-13)
            tot += np.sum(np.real(np.trace(to_mat(u[mu, eo]) @ to_mat(L.get_staple()[mu, eo]),
                                           axis1=-2, axis2=-1))) / 2
    # every plaquette has four links: sum_links 1/2 Tr(U V) = 4 sum_P 1/2 Tr U_P
    assert abs(tot - 4 * plaquette_sum(shape, u)) < 1e-10


def test_staples_unit_config_and_gauge_covariance(O):
    shape = (4, 4, 6)
    L = O.OracleLattice(*shape)
    L.unitconf()
    for eo in (0, 1):
        for mu in (1, 2, 3):
            L.staples(eo, mu)
            v = L.get_staple()[mu - 1, eo]
            assert np.allclose(v, [4, 0, 0, 0])
    # V(x) -> g(x+mu) V(x) g(x)^+ under U_mu(x) -> g(x) U_mu(x) g(x+mu)^+
    u, _ = rand_conf(shape, 4)
    g = rand_conf(shape, 5)[0][0]  # one random SU(2) per site (p, s)
    U, _ = lattice_fields(shape, u, np.zeros((2, u.shape[2], 3)))
    G = lattice_fields(shape, np.stack([g] * 3), np.zeros((2, g.shape[1], 3)))[0][0]
    Ug = np.stack([G @ U[mu] @ dag(np.roll(G, -1, axis=mu)) for mu in range(3)])
    ug, _ = checkerboard(shape, Ug, None)
    L.put_links(u)
    L.staples(1, 2)
    v0 = L.get_staple()[1, 1].copy()
    L.put_links(ug)
    L.staples(1, 2)
    v1 = L.get_staple()[1, 1]
    Gc, _ = checkerboard(shape, np.stack([G] * 3), None)
    gx = to_mat(Gc[0, 1])
    up, dn = L.site_tables()
    gxp = to_mat(Gc[0, 0][up[1, 1]])
    np.testing.assert_allclose(to_mat(v1), gxp @ to_mat(v0) @ dag(gx), atol=1e-12)


# ----------------------------------------------------------------- heat_kpi

def test_heat_kpi_closed_form_values(O):
    """fed uniforms with r3 = 1/4 (cos^2 = 0), r4 = 0 (always accept),
    r5 = 1/2 (a3 = 0), r6 = 0 (pure a2 = 0, a1 = s): W = (1 + ln r1 b, 0; 0 + i s)."""
    n = 5
    b = np.array([0.1, 0.2, 0.3, 0.4, 0.05])
    r1 = np.array([0.9, 0.5, 0.99, 0.7, 0.0])  # r1 = 0 -> guarded 3.3e-6
    uni = np.concatenate([np.stack([r1, np.full(n, 0.6), np.full(n, 0.25)], 1).ravel(),
                          np.zeros(n), np.full(n, 0.5), np.zeros(n)])
    L = O.OracleLattice(4, 4, 4)
    L.rng_feed(uni)
    v, nacch = L.heat_kpi(b, np.zeros(n))
    r1g = np.where(r1 == 0, 0.0000033, r1)
    c = np.cos(2 * np.pi * 0.25) ** 2
    a0 = 1 + (np.log(r1g) + np.log(0.6) * c) * b
    s = np.sqrt(np.maximum(0, 1 - a0 ** 2))
    assert nacch == n
    np.testing.assert_allclose(v[:, 0], a0, rtol=1e-15)
    np.testing.assert_allclose(v[:, 1], 0.0, atol=1e-16)
    np.testing.assert_allclose(v[:, 2], 0.0, atol=1e-16)  # a0 sin(0)
    np.testing.assert_allclose(v[:, 3], s, rtol=1e-14)    # a0 cos(0)
    # r4 = 1 rejects everything with a0 > 0: a0 falls back to a0_old
    L.rng_feed(np.concatenate([uni[:3 * n], np.ones(n), np.full(n, 0.5), np.zeros(n)]))
    v, nacch = L.heat_kpi(b, np.full(n, 0.3))
    assert nacch == 0 and np.allclose(v[:, 0], 0.3)


def test_heat_kpi_distribution(O):
    """accepted a0 follow sqrt(1-a0^2) exp(a0/b): <a0> = I2(1/b)/I1(1/b)."""
    n = 200000
    b = 0.5
    L = O.OracleLattice(4, 4, 4)
    L.rng_philox(11, 0)
    v, nacch = L.heat_kpi(np.full(n, b), np.full(n, 2.0))  # rejected -> a0 = 2
    a0 = v[:, 0][v[:, 0] < 1.5]
    assert len(a0) == nacch
    exp = special.iv(2, 1 / b) / special.iv(1, 1 / b)
    err = np.std(a0) / np.sqrt(len(a0))
    assert abs(np.mean(a0) - exp) < 5 * err
    # direction uniform on the sphere of radius sqrt(1-a0^2)
    w = v[v[:, 0] < 1.5]
    s2 = 1 - w[:, 0] ** 2
    np.testing.assert_allclose(w[:, 1] ** 2 + w[:, 2] ** 2 + w[:, 3] ** 2, s2, atol=1e-12)
    for c in (1, 2, 3):
        assert abs(np.mean(w[:, c] ** 2 / s2) - 1 / 3) < 0.01


# ------------------------------------------------------------- link updates

def test_agauge_pure_gauge_limit(O):
    """beta_A = 0: every hopping Metropolis accepts, accept = (nacch/3V)/4,
    links stay unitary, and the plaquette moves towards its beta=9 value."""
    shape = (8, 8, 8)
    L = O.OracleLattice(*shape)
    L.set_couplings(9.0, 0.0, 1.0, 0.0)
    L.hotstart(3, 0.5)
    L.rng_philox(3, 0)
    acc = L.agauge(1)
    assert 0.0 < acc <= 0.25
    u = L.get_links()
    assert np.max(np.abs(np.sum(u * u, -1) - 1)) < 1e-13
    for _ in range(30):
        L.agauge(1)
    # 3D SU(2) at beta = 9: <1/2 Tr U_P> ~ 0.88 (weak coupling 1 - 3/(4 beta) ~ 0.92)
    assert 0.8 < L.observables()["plaq0"] < 0.95


def test_link_or_invariants(O):
    """beta_A = 0: overrelaxation accepts every link and keeps the Wilson
    action (brute force) unchanged; the same map applied twice with the same
    staple is the identity and preserves 1/2 Tr(U V)."""
    shape = (6, 4, 8)
    L = O.OracleLattice(*shape)
    L.set_couplings(9.0, 0.0, 1.0, 0.0)
    L.hotstart(8, 0.5)
    L.rng_philox(8, 0)
    s0 = plaquette_sum(shape, L.get_links())
    acc = L.link_or(1)
    s1 = plaquette_sum(shape, L.get_links())
    assert acc == 1.0
    assert abs(s1 - s0) < 1e-12 * abs(s0)
    # single-link algebra: U -> Vh^+ U^+ Vh^+
    rng = np.random.default_rng(1)
    U = to_mat(rand_conf((2, 2, 2), 9)[0][0, 0])
    V = to_mat(rng.normal(size=(4, 4)))
    Vh = V / np.sqrt(np.real(np.linalg.det(V)))[:, None, None]
    Up = dag(Vh) @ dag(U) @ dag(Vh)
    np.testing.assert_allclose(np.trace(Up @ V, axis1=1, axis2=2).real,
                               np.trace(U @ V, axis1=1, axis2=2).real, atol=1e-13)
    np.testing.assert_allclose(dag(Vh) @ dag(Up) @ dag(Vh), U, atol=1e-13)


def test_hopping_metropolis_detailed_balance_sign(O):
    """agauge's hopping action S_hop = beta_A Re[X(a)U X(a')U^+]_11 equals
    -beta_A a.R(U)a' with R(U)_ab = 1/2 Tr(sigma_a U sigma_b U^+)."""
    from helpers import PAULI
    rng = np.random.default_rng(2)
    u = rand_conf((2, 2, 2), 10)[0][0, 0]
    a, ap = rng.normal(size=(4, 3)), rng.normal(size=(4, 3))
    U = to_mat(u)
    m = higgs_mat(a) @ U @ higgs_mat(ap) @ dag(U)
    R = np.einsum("aij,njk,bkl,nlm->nabim", PAULI, U, PAULI, dag(U))
    R = 0.5 * np.real(np.trace(R, axis1=-2, axis2=-1))
    np.testing.assert_allclose(np.real(m[:, 0, 0]), -np.einsum("na,nab,nb->n", a, R, ap),
                               atol=1e-13)


# ------------------------------------------------------------------- Higgs

def test_a_force_brute_force_and_unit_links(O):
    shape = (4, 6, 4)
    u, a = rand_conf(shape, 12)
    L = O.OracleLattice(*shape)
    L.put_links(u)
    L.put_higgs(a)
    U, A = lattice_fields(shape, u, a)
    F = np.zeros(A.shape)
    for mu in range(3):
        F += extract(U[mu] @ higgs_mat(np.roll(A, -1, axis=mu)) @ dag(U[mu]))
        Ud = np.roll(U[mu], 1, axis=mu)
        F += extract(dag(Ud) @ higgs_mat(np.roll(A, 1, axis=mu)) @ Ud)
    _, Fc = checkerboard(shape, None, F)
    for eo in (0, 1):
        L.a_force(eo)
        np.testing.assert_allclose(L.get_force(), Fc[eo], atol=1e-13)
    L.unitconf()
    L.put_higgs(a)
    L.a_force(0)
    S = sum(np.roll(A, -1, axis=mu) + np.roll(A, 1, axis=mu) for mu in range(3))
    _, Sc = checkerboard(shape, None, S)
    np.testing.assert_allclose(L.get_force(), Sc[0], atol=1e-13)


def _decoupled_moments(b2, b4):
    w = lambda r, k: r ** (0.5 + k) * np.exp(-b2 * r - b4 * r * r)
    z = integrate.quad(w, 0, np.inf, args=(0,))[0]
    return (integrate.quad(w, 0, np.inf, args=(1,))[0] / z,
            integrate.quad(w, 0, np.inf, args=(2,))[0] / z)


@pytest.mark.parametrize("routine,b2,b4", [("ahiggs2", 0.7, 0.0), ("ahiggs2", 0.45, 0.3),
                                           ("ahiggs1", -0.45, 0.1), ("ahiggs1", 0.6, 0.05)])
def test_higgs_decoupled_single_site(O, routine, b2, b4):
    """beta_A = 0: each site samples rho^(1/2) exp(-b2 rho - b4 rho^2) drho;
    <rho>, <rho^2> from 1-D quadrature (ahiggs2 with b4 = 0: <rho> = 3/(2 b2),
    acceptance 1)."""
    shape = (16, 16, 8)
    L = O.OracleLattice(*shape)
    L.set_couplings(9.0, 0.0, b2, b4)
    L.unitconf()
    L.put_higgs(np.full((2, L.N, 3), 0.5))
    L.rng_philox(21, 0)
    d = 1.0
    r2, r4 = [], []
    for it in range(260):
        if routine == "ahiggs2":
            acc = L.ahiggs2(1)
            if b4 == 0:
                assert acc == 1.0
        else:
            d, acc = L.ahiggs1(d, 1)
        if it >= 60:
            ob = L.observables()
            r2.append(ob["a02"])
            r4.append(ob["a04"])
    m2, m4 = _decoupled_moments(b2, b4)
    if routine == "ahiggs2" and b4 == 0:
        assert abs(m2 - 1.5 / b2) < 1e-9
    for est, exp in ((np.array(r2), m2), (np.array(r4), m4)):
        # sweeps are correlated for the Metropolis; block the series
        blocks = est.reshape(20, -1).mean(1)
        err = blocks.std(ddof=1) / np.sqrt(len(blocks))
        assert abs(est.mean() - exp) < 5 * err + 1e-3 * abs(exp), (est.mean(), exp, err)


def test_ahiggs3_is_an_involution_at_zero_beta4(O):
    """beta_A = 0, beta_4 = 0: a -> -a, always accepted, twice = identity."""
    shape = (4, 4, 4)
    u, a = rand_conf(shape, 13)
    L = O.OracleLattice(*shape)
    L.set_couplings(9.0, 0.0, 0.7, 0.0)
    L.put_links(u)
    L.put_higgs(a)
    L.rng_philox(1, 0)
    assert L.ahiggs3(1) == 1.0
    np.testing.assert_array_equal(L.get_higgs(), -a)
    L.ahiggs3(1)
    np.testing.assert_array_equal(L.get_higgs(), a)


# ----------------------------------------------------------------- renorm

def test_renorm_unit_and_idempotent(O):
    shape = (4, 4, 4)
    u, _ = rand_conf(shape, 14)
    u *= 1.7
    L = O.OracleLattice(*shape)
    L.put_links(u)
    L.renorm()
    r = L.get_links()
    assert np.max(np.abs(np.sum(r * r, -1) - 1)) < 1e-15
    np.testing.assert_allclose(r, u / 1.7, atol=1e-15)
    L.renorm()
    assert np.max(np.abs(L.get_links() - r)) < 3e-16


# ------------------------------------------------------------ gauge fixing

def _divergence(shape, uf):
    """|d_mu A_mu|^2 per odd site with A_mu the anti-hermitian part of U_mu
    (the quantity f/or_gauge_fixing1.f:153-158 averages)."""
    U, _ = lattice_fields(shape, uf, np.zeros((2, uf.shape[2], 3)))
    D = sum(U[mu] - np.roll(U[mu], 1, axis=mu) for mu in range(3))
    return (D[..., 0, 0].imag ** 2 + np.abs(D[..., 0, 1]) ** 2)


def test_landau_gauge_invariants(O):
    shape = (6, 4, 8)
    u, a = rand_conf(shape, 15)
    L = O.OracleLattice(*shape)
    L.put_links(u)
    L.put_higgs(a)
    L.set_gauge_fixing(1.7, 1.0, 1e-12)
    da, it = L.landau_gauge()
    assert da <= 1e-12 and it > 0 and L.error() == 0
    uf, af = L.get_links_fixed(), L.get_higgs_fixed()
    # gauge-invariant quantities are unchanged
    assert abs(plaquette_sum(shape, uf) - plaquette_sum(shape, u)) < 1e-10
    np.testing.assert_allclose(np.sum(af * af, -1), np.sum(a * a, -1), atol=1e-12)
    assert np.max(np.abs(np.sum(uf * uf, -1) - 1)) < 1e-14
    # Landau gauge reached: the functional sum Re Tr U_f is larger and the
    # lattice divergence vanishes everywhere
    assert np.sum(uf[..., 0]) > np.sum(u[..., 0])
    assert np.max(_divergence(shape, uf)) < 1e-9
    # there is a gauge transform g with a_f = g a g^+ (rho and a.a' preserved)
    # and the hopping term is invariant
    def hop(uu, aa):
        U, A = lattice_fields(shape, uu, aa)
        return sum(np.sum(np.real((higgs_mat(A) @ U[mu] @ higgs_mat(np.roll(A, -1, axis=mu))
                                   @ dag(U[mu]))[..., 0, 0])) for mu in range(3))
    assert abs(hop(uf, af) - hop(u, a)) < 1e-10


def test_landau_gauge_iteration_cap(O):
    L = O.OracleLattice(4, 4, 4)
    u, a = rand_conf((4, 4, 4), 16)
    L.put_links(u)
    L.put_higgs(a)
    L.set_gauge_fixing(1.7, 1.0, 1e-14, 3)
    da, it = L.landau_gauge()
    assert it == 3 and L.error() == 2


# ------------------------------------------------------------- correlators

def _brute_correlators(shape, uf, af, a):
    lx, ly, lz = shape
    U, AF = lattice_fields(shape, uf, af)
    _, A = lattice_fields(shape, uf, a)
    am = AF.mean(axis=(0, 1))                                     # (lz, 3)
    Am = np.stack([np.stack([U[mu][..., 0, 0].imag, (-1j * U[mu][..., 0, 1])], -1)
                   for mu in range(3)]).mean(axis=(1, 2))         # (3, lz, 2) complex
    rho = (A ** 2).sum(-1).mean(axis=(0, 1))
    pol = np.cos(4 * np.sqrt((A ** 2).sum(-1))).mean(axis=(0, 1))
    ca, cu, c3, cr, cp = (np.zeros(lz) for _ in range(5))
    for z in range(lz):
        for x0 in range(lz):
            x1 = (x0 + z) % lz
            ca[z] += am[x0] @ am[x1]
            cu[z] += sum(np.real(Am[mu, x0, 0] * Am[mu, x1, 0] + Am[mu, x0, 1] * np.conj(Am[mu, x1, 1]))
                         for mu in (0, 1))
            c3[z] += np.real(Am[2, x0, 0] * Am[2, x1, 0] + Am[2, x0, 1] * np.conj(Am[2, x1, 1]))
            cr[z] += rho[x0] * rho[x1]
            cp[z] += pol[x0] * pol[x1]
    return ca / lz, cu / (2 * lz), c3 / lz, cr / lz, cp / lz, rho.mean(), pol.mean() ** 2, \
        AF.mean(axis=(0, 1, 2))


def test_correlators_brute_force(O):
    shape = (4, 6, 8)
    L = O.OracleLattice(*shape)
    uf, af = rand_conf(shape, 17)
    _, a = rand_conf(shape, 18)
    L.put_links_fixed(uf)
    L.put_higgs_fixed(af)
    L.put_higgs(a)
    ca, cu, c3 = L.correlator()
    mr, cr = L.scalarcorr()
    mp, cp = L.polyakov()
    b = _brute_correlators(shape, uf, af, a)
    for got, exp in zip((ca, cu, c3, cr, cp, mr, mp, L.observables()["a0f"]), b):
        np.testing.assert_allclose(got, exp, atol=1e-13)


def test_correlators_special_configs(O):
    """testconf: a = (1,1,1)/sqrt(3) -> C_a = 1; unit links -> C_u = C_3 = 0;
    constant Higgs -> scalar corr = rho^2, Polyakov mean^2 = cos(4|a|)^2."""
    shape = (4, 4, 6)
    L = O.OracleLattice(*shape)
    L.testconf()
    L.put_links_fixed(L.get_links())
    L.put_higgs_fixed(L.get_higgs())
    ca, cu, c3 = L.correlator()
    np.testing.assert_allclose(ca, 1.0, atol=1e-14)
    L.unitconf()
    L.put_links_fixed(L.get_links())
    ca, cu, c3 = L.correlator()
    assert np.all(cu == 0) and np.all(c3 == 0)
    a0 = np.array([0.3, -0.2, 0.5])
    L.put_higgs(np.broadcast_to(a0, (2, L.N, 3)))
    rho = a0 @ a0
    m, c = L.scalarcorr()
    np.testing.assert_allclose(c, rho ** 2, rtol=1e-14)
    np.testing.assert_allclose(m, rho, rtol=1e-14)
    m, c = L.polyakov()
    np.testing.assert_allclose(m, np.cos(4 * np.sqrt(rho)) ** 2, rtol=1e-14)
