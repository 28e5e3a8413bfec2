"""The replica batch at the full size of the scan configuration against the
CPU restatement (oracle/lceq_oracle.c), and the batched learned proposal.

- 96 replicas cycling 16^3, 24^3, 32^3, 40^3, 48^3 and 24x24x96 on the
  beta_A x beta_2 grid (seed 12345 + r), hot start, swept by the batch's
  launches: one replica of every shape follows the restatement's sweep
  (f/agauge.f, link overrelaxation, f/ahiggs1.f) with the same seed and
  couplings, and every replica keeps finite, unitary fields;
- the learned Higgs proposal over a batch of uneven replicas reproduces, bit
  for bit, the single-lattice learned chain of every replica, with and
  without graph replay; with a bias-only network it follows the
  restatement's learned update (oc_higgs_mlp);
- training records collected from the batch leave every replica where the
  same collect on a single lattice leaves it.

Tolerances after k chained sweeps are those of the restatement's own
sensitivity: two restatement chains started 1 ulp apart differ by about
1.6e-13 after one production sweep and 1.7e-12 after two at 16^3
(tests/test_oracle_chaos.py), so GPU == restatement within 1e-12 after one
sweep and 1e-11 after two, with no accept/reject flip."""
import numpy as np
import pytest

from helpers import assert_close, rand_conf

pytestmark = pytest.mark.gpu


def _oracle_sweep(O, d, measflag=1):
    O.agauge(0)
    for k in range(4):
        O.link_or(measflag if k == 3 else 0)
    d, _ = O.ahiggs1(d, measflag)
    return d


def _flips(got, exp):
    d = np.abs(np.asarray(got) - np.asarray(exp)).max(axis=-1)
    return int((d > 1e-9).sum()), float(d.max())


def test_c4_batch_against_oracle(oracle_lib):
    import lceqcd_amd
    from lceqcd_amd.workloads import c4_replicas
    shapes, coup, seeds, grid = c4_replicas(96)
    print("C4 batch: 96 replicas, %d sites" % sum(np.prod(s) for s in shapes), flush=True)
    B = lceqcd_amd.Batch(shapes)
    B.set_couplings(coup)
    B.hotstart(seeds, 0.5)
    B.rng_philox(seeds)
    # one replica of every shape, spread over the grid
    picks = [i + 6 * ((3 * i) % 16) for i in range(6)]
    assert sorted(set(shapes[r] for r in picks)) == sorted(set(shapes))
    orc = {}
    for r in picks:
        O = oracle_lib.OracleLattice(*shapes[r])
        O.set_couplings(*coup[r])
        O.hotstart(seeds[r], 0.5)
        O.philox(seeds[r], 0)
        O.set_renorm_on_write(1)
        orc[r] = [O, 0.4]
        assert_close(B.get_links(r), O.get_links(), 1e-14, "hot start links %d" % r)
        assert_close(B.get_higgs(r), O.get_higgs(), 1e-14, "hot start higgs %d" % r)
    for sweep, tol in ((1, 1e-12), (2, 1e-11)):
        B.sweep(1, n_or=4, measflag=1, graph=False)
        h = B.history(1)[0]
        for r in picks:
            O, d = orc[r]
            orc[r][1] = d = _oracle_sweep(O, d)
            ob = O.observables()
            fl_u, du = _flips(B.get_links(r), O.get_links())
            fl_a, da = _flips(B.get_higgs(r), O.get_higgs())
            print("sweep %d replica %2d %-12s beta_A %.3f beta_2 %.3f: links %.2e higgs %.2e "
                  "flips %d %d" % (sweep, r, "x".join(map(str, shapes[r])), grid[r][0],
                                   grid[r][1], du, da, fl_u, fl_a), flush=True)
            assert fl_u == 0 and fl_a == 0
            assert_close(B.get_links(r), O.get_links(), tol, "links %d" % r)
            assert_close(B.get_higgs(r), O.get_higgs(), tol, "higgs %d" % r)
            assert_close(h[r, 0], ob["plaq0"], tol, "plaq0 %d" % r)
            assert_close(h[r, 2], ob["a02"], tol, "a02 %d" % r)
            assert_close(h[r, 11], d, 1e-15, "delta %d" % r)
    # a further graph-replayed stretch, then every replica finite and unitary
    B.sweep(8, n_or=4, measflag=1, graph=True)
    worst = 0.0
    for r in range(96):
        u, a = B.get_links(r), B.get_higgs(r)
        assert np.isfinite(u).all() and np.isfinite(a).all(), r
        worst = max(worst, float(np.abs(np.linalg.norm(u, axis=-1) - 1.0).max()))
    print("largest |U| - 1 over 96 replicas: %.2e" % worst, flush=True)
    assert worst < 1e-14
    h = B.history(8)
    assert np.isfinite(h).all()


SHAPES = [(4, 4, 4), (8, 6, 4), (16, 16, 16), (6, 6, 24), (10, 8, 12), (2, 2, 2)]
PROD = (9.0, 0.175, -0.45, 0.025)


def _coup(r):
    return (9.0, 0.15 + 0.01 * r, -0.5 + 0.03 * r, 0.025)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_learned_matches_single_lattices(graph):
    """One network over all replicas: replica r's chain is the single
    lattice's lceq_sweep_mlp chain with the same seed and couplings."""
    import lceqcd_amd
    print("batched learned vs single, graph", graph, flush=True)
    M = lceqcd_amd.Mlp(seed=31)
    B = lceqcd_amd.Batch(SHAPES)
    seeds = [4000 + r for r in range(len(SHAPES))]
    B.set_couplings([_coup(r) for r in range(len(SHAPES))])
    B.hotstart(seeds, 0.6)
    B.rng_philox(seeds)
    B.sweep(2, n_or=4, delta=0.4, measflag=0, graph=False)
    B.sweep_mlp(M, 9, n_or=4, measflag=1, graph=graph)
    hist = B.history(9)
    for r, shp in enumerate(SHAPES):
        L = lceqcd_amd.Lattice(*shp)
        L.set_couplings(*_coup(r))
        L.hotstart(seeds[r], 0.6)
        L.rng_philox(seeds[r])
        L.sweep(0, 2, n_or=4, delta=0.4, measflag=0)
        L.sweep_mlp(M, 9, n_or=4, measflag=1)
        h = L.history(9)
        assert np.array_equal(B.get_links(r), L.get_links()), r
        assert np.array_equal(B.get_higgs(r), L.get_higgs()), r
        np.testing.assert_array_equal(hist[:, r, :12], h[:, :12])
        L.close()


def _bias_only_state(M):
    st = M.get_state()
    st[:5254] = 0.0
    st[5248:5254] = np.array([0.05, -0.02, 0.0, np.log(0.3), np.log(0.25), np.log(0.25)],
                             np.float32)
    st[5254:5264] = 0.0
    st[5264:5274] = 1.0
    return st


def test_batch_learned_bias_only_equals_oracle(oracle_lib):
    """With a bias-only network (a proposal independent of a) every replica
    follows the restatement's sweep with its learned update, oc_higgs_mlp."""
    import lceqcd_amd
    shapes = [(4, 4, 4), (8, 6, 4), (10, 6, 8), (8, 8, 8)]
    M = lceqcd_amd.Mlp(seed=2)
    st = _bias_only_state(M)
    M.set_state(st)
    B = lceqcd_amd.Batch(shapes)
    coup = [_coup(r) for r in range(len(shapes))]
    B.set_couplings(coup)
    orc = []
    for r, shp in enumerate(shapes):
        u, a = rand_conf(shp, 50 + r, 0.4)
        B.put_links(r, u)
        B.put_higgs(r, a)
        O = oracle_lib.OracleLattice(*shp)
        O.set_couplings(*coup[r])
        O.put_links(u)
        O.put_higgs(a)
        O.philox(700 + r, 0)
        O.set_renorm_on_write(1)
        orc.append(O)
    B.rng_philox([700 + r for r in range(len(shapes))])
    for sweep, tol in ((1, 1e-12), (2, 1e-11)):
        B.sweep_mlp(M, 1, n_or=4, measflag=1, graph=False)
        h = B.history(1)[0]
        for r, O in enumerate(orc):
            O.agauge(0)
            for k in range(4):
                O.link_or(1 if k == 3 else 0)
            acc = O.higgs_mlp(st)
            fl, dev = _flips(B.get_higgs(r), O.get_higgs())
            print("sweep %d replica %d: higgs %.2e flips %d, acceptance %.4f / %.4f"
                  % (sweep, r, dev, fl, h[r, 8], acc), flush=True)
            assert fl == 0
            assert h[r, 8] == acc
            assert_close(B.get_links(r), O.get_links(), tol, "links %d" % r)
            assert_close(B.get_higgs(r), O.get_higgs(), tol, "higgs %d" % r)
            assert_close(h[r, 2], O.observables()["a02"], tol, "a02 %d" % r)


def test_batch_collect_matches_single_lattices():
    """The batch's collect keeps its Metropolis steps in the chains exactly as
    lceq_mlp_collect does on a single lattice (same passes, same steps), and
    counts the same records."""
    import lceqcd_amd
    print("batched collect vs single", flush=True)
    seeds = [600 + r for r in range(len(SHAPES))]
    B = lceqcd_amd.Batch(SHAPES)
    B.set_couplings([_coup(r) for r in range(len(SHAPES))])
    B.hotstart(seeds, 0.5)
    B.rng_philox(seeds)
    M = lceqcd_amd.Mlp(seed=3)
    B.sweep(1, n_or=2, delta=0.4, measflag=0, graph=False)
    B.mlp_collect(M, nsteps=3, stride=2)
    B.sweep(1, n_or=2, measflag=1, graph=False)
    nrec = sum(2 * ((s[0] * s[1] * s[2] // 2 + 1) // 2) for s in SHAPES)
    assert M.records() == nrec
    for r, shp in enumerate(SHAPES):
        L = lceqcd_amd.Lattice(*shp)
        L.set_couplings(*_coup(r))
        L.hotstart(seeds[r], 0.5)
        L.rng_philox(seeds[r])
        Ms = lceqcd_amd.Mlp(seed=3)
        d = L.sweep(0, 1, n_or=2, delta=0.4, measflag=0)
        Ms.collect(L, nsteps=3, stride=2)
        L.sweep(0, 1, n_or=2, delta=d, measflag=1)
        assert np.array_equal(B.get_links(r), L.get_links()), r
        assert np.array_equal(B.get_higgs(r), L.get_higgs()), r
        L.close()
    loss = M.train(50, batch=256, lr=2e-3, seed=1)
    assert np.isfinite(loss).all()
