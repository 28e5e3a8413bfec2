"""Batched replicas (lceq_batch_*): every replica of a batch of uneven
lattices with different couplings reproduces, to the bit, the chain of a
single lattice with the same seed and couplings (same per-site code, same
Philox counters, same reduction order), with and without hipGraph replay."""
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
