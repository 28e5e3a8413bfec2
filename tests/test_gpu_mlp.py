"""Learned Higgs proposal (lceq_mlp_*, lceq_higgs_mlp, lceq_sweep_mlp):

- the bf16 MFMA inference network against a CPU emulation of the same
  bf16 arithmetic (catches any operand-layout error, which is O(1));
- a bias-only network (a proposal independent of a) against the CPU
  restatement oracle/lceq_oracle.c:oc_higgs_mlp: same fields to 1e-12 and
  the same acceptance, which pins the frame, the Box-Muller draws, both
  proposal densities and the Metropolis-Hastings test;
- exactness for arbitrary weights: at beta_A = 0 the chain of an untrained
  network reproduces <rho>, <rho^2> of the single-site distribution
  rho^{1/2} exp(-beta_2 rho - beta_4 rho^2) by 1-D quadrature;
- training lowers the loss and gives a higher acceptance than the plain
  Metropolis step."""
import numpy as np
import pytest

from helpers import rand_conf

pytestmark = pytest.mark.gpu

PROD = (9.0, 0.175, -0.45, 0.025)


@pytest.fixture(scope="module")
def lceq():
    import lceqcd_amd
    return lceqcd_amd


def bf16(x):
    """round-to-nearest-even float32 -> bfloat16 -> float32"""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def forward(st, x):
    """The inference network in numpy with the GPU's bf16 roundings"""
    nin, nh, nout = 10, 64, 6
    o_b1 = nin * nh
    o_w2 = o_b1 + nh
    o_b2 = o_w2 + nh * nh
    o_w3 = o_b2 + nh
    o_b3 = o_w3 + nh * nout
    npar = o_b3 + nout
    w1 = bf16(st[:o_b1].reshape(nin, nh)).astype(np.float64)
    w2 = bf16(st[o_w2:o_b2].reshape(nh, nh)).astype(np.float64)
    w3 = bf16(st[o_w3:o_b3].reshape(nh, nout)).astype(np.float64)
    silu = lambda z: z / (np.float32(1) + np.exp(-z))
    xn = bf16((x - st[npar:npar + nin]) * st[npar + nin:npar + 2 * nin]).astype(np.float64)
    h1 = bf16(silu((xn @ w1).astype(np.float32) + st[o_b1:o_w2])).astype(np.float64)
    h2 = bf16(silu((h1 @ w2).astype(np.float32) + st[o_b2:o_w3])).astype(np.float64)
    return (h2 @ w3).astype(np.float32) + st[o_b3:npar]


def test_mfma_inference_matches_bf16_emulation(lceq):
    print("mlp inference vs emulation", flush=True)
    M = lceq.Mlp(seed=11)
    st = M.get_state()
    rng = np.random.default_rng(3)
    st[5254:5264] = rng.normal(size=10).astype(np.float32) * 0.3
    st[5264:5274] = (0.5 + rng.random(10)).astype(np.float32)
    st[640:704] = rng.normal(size=64).astype(np.float32) * 0.1   # b1
    st[5248:5254] = rng.normal(size=6).astype(np.float32) * 0.1  # b3
    M.set_state(st)
    x = rng.normal(size=(256 * 3 + 37, 10)).astype(np.float32) * 2.0
    y = M.eval(x)
    ref = forward(M.get_state(), x)
    err = np.abs(y - ref)
    scale = np.abs(ref).max()
    print("max err %.3g median %.3g (scale %.3g)" % (err.max(), np.median(err), scale), flush=True)
    # fp32 accumulation order inside MFMA can flip a bf16 rounding of a hidden
    # activation now and then; a wrong operand layout gives O(1) errors
    assert np.median(err) < 1e-5 * scale
    assert err.max() < 2e-2 * scale


def _bias_only_state(M):
    st = M.get_state()
    st[:5254] = 0.0
    st[5248:5254] = np.array([0.05, -0.02, 0.0, np.log(0.3), np.log(0.25), np.log(0.25)],
                             np.float32)
    st[5254:5264] = 0.0
    st[5264:5274] = 1.0
    return st


@pytest.mark.parametrize("spw", [32, 64])
@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 4, 8), (8, 8, 8), (10, 6, 8)])
def test_bias_only_network_equals_oracle(lceq, oracle_lib, shape, spw):
    print("mlp bias-only vs oracle", shape, spw, flush=True)
    M = lceq.Mlp(seed=2)
    M.set_sites_per_wave(spw)
    st = _bias_only_state(M)
    M.set_state(st)
    u, a = rand_conf(shape, 17, 0.4)
    G, O = lceq.Lattice(*shape), oracle_lib.OracleLattice(*shape)
    for L in (G, O):
        L.set_couplings(*PROD)
        L.put_links(u)
        L.put_higgs(a)
        L.rng_philox(77, 3)
    for k in range(3):
        ag = G.higgs_mlp(M)
        ao = O.higgs_mlp(st)
        assert ag == ao, (k, ag, ao)
        np.testing.assert_allclose(G.get_higgs(), O.get_higgs(), rtol=0, atol=1e-12)
    og, oo = G.observables(), O.observables()
    for key in ("a02", "a04", "a0kin"):
        np.testing.assert_allclose(og[key], oo[key], rtol=1e-12)


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 16, 16), (10, 6, 8)])
def test_sites_per_wave_gives_the_same_chain(lceq, shape):
    """32 and 64 sites per wave place a site in different MFMA columns and
    lanes; each column is computed on its own, so the chains agree bit for
    bit (untrained network: the proposal depends on a and F)."""
    print("mlp sites per wave", shape, flush=True)
    M = lceq.Mlp(seed=13)
    u, a = rand_conf(shape, 23, 0.6)
    runs = []
    for spw in (32, 64):
        M.set_sites_per_wave(spw)
        G = lceq.Lattice(*shape)
        G.set_couplings(*PROD)
        G.put_links(u)
        G.put_higgs(a)
        G.rng_philox(91, 0)
        acc = [G.higgs_mlp(M) for _ in range(3)]
        G.sweep_mlp(M, 2, n_or=2, measflag=1)
        runs.append((acc, G.get_higgs(), G.get_links(), G.history(2)))
        G.close()
    assert runs[0][0] == runs[1][0]
    for i in (1, 2):
        assert np.array_equal(runs[0][i], runs[1][i])
    # the observables are sums over blocks, whose number differs
    np.testing.assert_allclose(runs[0][3], runs[1][3], rtol=1e-12, atol=1e-15)
    with pytest.raises(lceq.LceqError, match="sites per wave"):
        M.set_sites_per_wave(48)


def _single_site_moments(b2, b4):
    from scipy.integrate import quad
    w = lambda r, k: r ** k * np.sqrt(r) * np.exp(-b2 * r - b4 * r * r)
    z = quad(w, 0, np.inf, args=(0,))[0]
    return quad(w, 0, np.inf, args=(1,))[0] / z, quad(w, 0, np.inf, args=(2,))[0] / z


def test_exact_for_untrained_network_at_beta_a_zero(lceq):
    """beta_A = 0 decouples the sites: rho = |a|^2 has density
    rho^{1/2} exp(-beta_2 rho - beta_4 rho^2).  An untrained network still
    proposes moves that depend on F (through the frame) and on a; the
    Metropolis-Hastings test with both densities must give the exact law."""
    print("mlp exactness beta_A = 0", flush=True)
    shape = (16, 16, 16)
    b2, b4 = 0.5, 0.1
    M = lceq.Mlp(seed=5)
    G = lceq.Lattice(*shape)
    G.set_couplings(9.0, 0.0, b2, b4)
    G.hotstart(9, 0.8)
    G.rng_philox(9, 0)
    rho, rho2, acc = [], [], []
    for it in range(600):
        acc.append(G.higgs_mlp(M))
        ob = G.observables()
        if it >= 100:
            rho.append(ob["a02"])
            rho2.append(ob["a04"])
    m1, m2 = _single_site_moments(b2, b4)
    for name, x, m in (("rho", rho, m1), ("rho^2", rho2, m2)):
        xb = np.asarray(x).reshape(25, -1).mean(axis=1)
        err = xb.std(ddof=1) / np.sqrt(len(xb))
        z = (xb.mean() - m) / err
        print("%s %.5f +- %.5f exact %.5f z %+.2f (acceptance %.3f)"
              % (name, xb.mean(), err, m, z, np.mean(acc)), flush=True)
        assert abs(z) < 5


def test_training_raises_acceptance(lceq):
    print("mlp training", flush=True)
    shape = (16, 16, 16)
    coup = lceq.couplings_from_trace_convention(9.0, 0.34, -0.9, 0.1)
    G = lceq.Lattice(*shape)
    G.set_couplings(*coup)
    G.hotstart(21, 0.5)
    G.rng_philox(21, 0)
    d = G.sweep(0, 100, n_or=4, delta=0.4, measflag=0)
    M = lceq.Mlp(seed=7)
    for c in range(16):
        d = G.sweep(0, 5, n_or=4, delta=d, measflag=0)
        M.collect(G, nsteps=8, stride=1)
    assert M.records() == 16 * 16 ** 3
    loss = M.train(400, batch=4096, lr=2e-3, seed=3)
    print("loss first %.4f last %.4f" % (loss[:20].mean(), loss[-20:].mean()), flush=True)
    assert loss[-20:].mean() < loss[:20].mean() - 0.1
    acc_mlp = np.mean([G.higgs_mlp(M) for _ in range(20)])
    acc_plain = np.mean([G.ahiggs1(d, 1)[1] for _ in range(20)])
    print("acceptance learned %.3f plain %.3f" % (acc_mlp, acc_plain), flush=True)
    assert acc_mlp > acc_plain
    G.sweep_mlp(M, 3, n_or=4, measflag=1)
    h = G.history(3)
    assert np.isfinite(h).all() and (h[:, 2] > 0).all()
