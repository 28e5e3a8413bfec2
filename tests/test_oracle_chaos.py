"""Two properties of the Markov chain itself, measured on the CPU
restatement alone, that set how GPU and CPU chains can be compared.

1. Sensitivity to the last bit.  Two restatement chains of the production
   sweep (1 KP + 4 OR + ahiggs1, f/agauge.f, link overrelaxation,
   f/ahiggs1.f) that start from links 1 ulp apart and draw the same Philox
   uniforms drift apart by about 1.6e-13 after one sweep, 1.7e-12 after two
   and 9e-11 after six (16^3, hot start).  An implementation that rounds
   differently anywhere (fused multiply-adds, another libm, another but
   equal algebraic form) can therefore not stay closer than that, and the
   GPU chain comparisons use 1e-12 after one sweep, 1e-11 after two and the
   measured envelope times a small factor beyond (tests/test_gpu_parity.py).
   The Higgs field stays bit-identical while no accept/reject decision
   flips: the Metropolis proposal a + (r - 1/2) delta does not depend on
   the links.  Without the heatbath (overrelaxation and the ahiggs3
   reflection only) the map is far more chaotic: at 8^3, 2e-12 after one
   sweep, 1e-9 after two, 1e-6 after three.

2. Metastability at the production couplings.  Two restatement chains on
   8x8x16 from hot starts 101 and 202 (300 + 3000 sweeps) give plaquettes
   that differ by about 9 combined standard errors: the hot start leaves
   long-lived defects whose number depends on the start, and the
   Madras-Sokal errors, which see only the fast modes, do not cover the
   difference.  So chains are compared from one common thermalised start
   (tests/test_gpu_statistics.py)."""
import numpy as np

from helpers import rand_conf

PROD = (9.0, 0.175, -0.45, 0.025)


def _pair(oracle_lib, shape, perturb_seed, kp):
    u, a = rand_conf(shape, 3, 2.0)
    rng = np.random.default_rng(perturb_seed)
    up = u * (1.0 + 2.0 ** -52 * rng.choice([-1.0, 1.0], size=u.shape))
    out = []
    for links in (u, up):
        O = oracle_lib.OracleLattice(*shape)
        O.set_couplings(*PROD)
        O.put_links(links)
        O.put_higgs(a)
        O.philox(1003, 7)
        O.set_renorm_on_write(1)
        out.append(O)
    return out


def _sweep(O, d, kp):
    if kp:
        O.agauge(0)
    for k in range(4):
        O.link_or(0)
    d, _ = O.ahiggs1(d, 1)
    if not kp:
        O.ahiggs3(1)
    return d


def test_one_ulp_envelope_of_the_production_sweep(oracle_lib):
    A, B = _pair(oracle_lib, (16, 16, 16), 9, kp=True)
    dA = dB = 0.4
    dev = []
    for it in range(6):
        dA, dB = _sweep(A, dA, True), _sweep(B, dB, True)
        dev.append(float(np.abs(A.get_links() - B.get_links()).max()))
        assert np.array_equal(A.get_higgs(), B.get_higgs())  # no flip
        assert dA == dB
    print("1-ulp envelope after 1..6 sweeps:", ["%.2e" % x for x in dev])
    # the chain amplifies a 2e-16 perturbation to ~1e-13 in one sweep and
    # keeps growing: ~1.6e-13, 1.7e-12, ..., 9e-11
    assert 1e-14 < dev[0] < 1e-12
    assert 1e-13 < dev[1] < 1e-11
    assert 1e-12 < dev[5] < 1e-9
    assert dev[5] > 10 * dev[0]


def test_overrelaxation_alone_is_far_more_chaotic(oracle_lib):
    A, B = _pair(oracle_lib, (8, 8, 8), 10, kp=False)
    dA = dB = 0.4
    dev = []
    for it in range(3):
        dA, dB = _sweep(A, dA, False), _sweep(B, dB, False)
        dev.append(float(np.abs(A.get_higgs() - B.get_higgs()).max()))
    print("OR + ahiggs3 envelope after 1..3 sweeps:", ["%.2e" % x for x in dev])
    assert dev[2] > 1e3 * dev[0] > 0


def test_hot_starts_leave_metastable_plaquettes(oracle_lib):
    from lceqcd_amd.analysis import mean_error
    means = []
    for seed in (101, 202):
        O = oracle_lib.OracleLattice(8, 8, 16)
        O.set_couplings(*PROD)
        O.set_renorm_on_write(1)
        O.hotstart(seed, 0.5)
        O.philox(seed, 0)
        d = 0.4
        for it in range(300):
            d = _sweep(O, d, True)
        plaq = []
        for it in range(3000):
            O.agauge(0)
            for k in range(4):
                O.link_or(1 if k == 3 else 0)
            d, _ = O.ahiggs1(d, 1)
            plaq.append(O.observables()["plaq0"])
        means.append(mean_error(plaq))
    (m1, e1), (m2, e2) = means
    z = (m1 - m2) / np.hypot(e1, e2)
    print("plaquette from hot start 101: %.6f(%d), 202: %.6f(%d), z = %.2f"
          % (m1, round(e1 * 1e6), m2, round(e2 * 1e6), z))
    assert abs(z) > 5
