"""Batched replicas (lceq_batch_*): every replica of a batch of uneven
lattices with different couplings reproduces, to the bit, the chain of a
single lattice with the same seed and couplings (same per-site code, same
Philox counters, same reduction order), with and without hipGraph replay.

The batch's launches always use the schedule of large cache-resident
lattices (the operands at the site itself issued first, a fenced uniform
phase), while these small single lattices use the other schedule, so the
tests also pin the two schedules to the same bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4), (8, 6, 4), (16, 16, 16), (6, 6, 24), (2, 2, 2), (10, 8, 12)]


def _coup(r):
    return (9.0, 0.14 + 0.01 * r, -0.6 + 0.05 * r, 0.025)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_lattices(graph):
    import lceqcd_amd
    print("batch vs single, graph", graph, flush=True)
    B = lceqcd_amd.Batch(SHAPES)
    seeds = [1000 + r for r in range(len(SHAPES))]
    B.set_couplings([_coup(r) for r in range(len(SHAPES))])
    B.hotstart(seeds, 0.5)
    B.rng_philox(seeds)
    B.sweep(3, n_or=4, delta=0.4, measflag=1, graph=graph)
    hist = B.history(3)
    for r, shp in enumerate(SHAPES):
        L = lceqcd_amd.Lattice(*shp)
        L.set_couplings(*_coup(r))
        L.hotstart(seeds[r], 0.5)
        L.rng_philox(seeds[r])
        L.sweep(0, 3, n_or=4, delta=0.4, measflag=1)
        h = L.history(3)
        assert np.array_equal(B.get_links(r), L.get_links()), r
        assert np.array_equal(B.get_higgs(r), L.get_higgs()), r
        np.testing.assert_array_equal(hist[:, r, :12], h[:, :12])


def test_batch_graph_replay_equals_direct_launches():
    import lceqcd_amd
    out = []
    for graph in (False, True):
        B = lceqcd_amd.Batch(SHAPES)
        seeds = [7 + r for r in range(len(SHAPES))]
        B.set_couplings([_coup(r) for r in range(len(SHAPES))])
        B.hotstart(seeds, 0.5)
        B.rng_philox(seeds)
        B.sweep(2, n_or=4, delta=0.4, measflag=0, graph=graph)
        B.sweep(4, n_or=4, measflag=1, graph=graph)
        out.append((B.history(4), [B.get_links(r) for r in range(len(SHAPES))]))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for x, y in zip(out[0][1], out[1][1]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("shape", [(64, False), (128, True), (512, True)])
def test_launch_shapes_keep_the_fields(shape):
    """Every launch shape gives the single-lattice fields bit for bit; the
    observables (sums in a shape-dependent order) agree to rounding."""
    import lceqcd_amd
    print("batch shape", shape, flush=True)
    B = lceqcd_amd.Batch(SHAPES)
    B.set_shape(*shape)
    assert B.get_shape() == shape
    seeds = [500 + r for r in range(len(SHAPES))]
    B.set_couplings([_coup(r) for r in range(len(SHAPES))])
    B.hotstart(seeds, 0.5)
    B.rng_philox(seeds)
    B.sweep(9, n_or=4, delta=0.4, measflag=1, graph=True)
    hist = B.history(9)
    for r, shp in enumerate(SHAPES):
        L = lceqcd_amd.Lattice(*shp)
        L.set_couplings(*_coup(r))
        L.hotstart(seeds[r], 0.5)
        L.rng_philox(seeds[r])
        L.sweep(0, 9, n_or=4, delta=0.4, measflag=1)
        h = L.history(9)
        assert np.array_equal(B.get_links(r), L.get_links()), r
        assert np.array_equal(B.get_higgs(r), L.get_higgs()), r
        np.testing.assert_allclose(hist[:, r, :12], h[:, :12], rtol=1e-13, atol=1e-15)


def test_autotune_restores_the_chains():
    """Tuning times every shape on the live batch but leaves its fields,
    observables and sweep counter as they were: sweeping after tuning gives
    the same chain as sweeping without tuning."""
    import lceqcd_amd
    out = []
    for tune in (False, True):
        B = lceqcd_amd.Batch(SHAPES)
        seeds = [900 + r for r in range(len(SHAPES))]
        B.set_couplings([_coup(r) for r in range(len(SHAPES))])
        B.hotstart(seeds, 0.5)
        B.rng_philox(seeds)
        B.sweep(2, n_or=4, delta=0.4, measflag=0)
        if tune:
            ms = B.autotune(n_or=4, trials=2)
            print("autotune", {k: round(v, 3) for k, v in ms.items()}, "->", B.get_shape(),
                  flush=True)
            assert len(ms) == 8 and min(ms.values()) > 0
            assert ms[B.get_shape()] == min(ms.values())
        B.sweep(3, n_or=4, measflag=1)
        out.append([B.get_links(r) for r in range(len(SHAPES))])
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)
