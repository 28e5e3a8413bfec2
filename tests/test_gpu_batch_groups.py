"""Replica groups of the batch (lceq_batch_set_group_bytes): a sweep runs its
launches over one run of consecutive replicas at a time.  The split follows
the requested bytes, and every grouping gives the one-group chains bit for
bit (fields and history rows), for the plain and the learned Higgs update,
with and without graph replay, and through the launch-shape autotuner."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(4, 4, 4), (8, 6, 4), (16, 16, 16), (6, 6, 24), (2, 2, 2), (10, 8, 12), (12, 12, 12)]


def _coup(r):
    return (9.0, 0.14 + 0.01 * r, -0.6 + 0.05 * r, 0.025)


def _field_bytes(shape):
    n = shape[0] * shape[1] * shape[2] // 2
    return 30 * 8 * ((n + 63) & ~63)


def _split(gb):
    """The split the library documents: ceil(total / gb) runs of about equal
    bytes (greedy, closing a run once it reaches its share)."""
    by = [_field_bytes(s) for s in SHAPES]
    total, n = sum(by), len(by)
    ng = max(1, min(n, -(-total // gb) if gb > 0 else 1))
    first, cum = [0], 0
    for r in range(n - 1):
        cum += by[r]
        g = len(first) - 1
        left, todo = n - 1 - r, ng - 1 - g
        if g < ng - 1 and left >= todo and (cum * ng >= total * (g + 1) or left == todo):
            first.append(r + 1)
    return first


def _batch(seed0):
    import lceqcd_amd
    B = lceqcd_amd.Batch(SHAPES)
    seeds = [seed0 + r for r in range(len(SHAPES))]
    B.set_couplings([_coup(r) for r in range(len(SHAPES))])
    B.hotstart(seeds, 0.5)
    B.rng_philox(seeds)
    return B


def test_group_split_follows_the_bytes():
    total = sum(_field_bytes(s) for s in SHAPES)
    B = _batch(10)
    assert B.groups() == (288 << 20, [0])  # default: the batch fits one group
    for gb in (0, 1, total // 2, total // 3 + 1, _field_bytes((16, 16, 16)), total, 10 * total):
        B.set_group_bytes(gb)
        got_gb, first = B.groups()
        print("group bytes", gb, "first replicas", first, flush=True)
        assert got_gb == gb
        assert first == _split(gb)
    assert B.groups()[1] == [0]
    B.set_group_bytes(1)
    assert B.groups()[1] == list(range(len(SHAPES)))
    import lceqcd_amd
    with pytest.raises(lceqcd_amd.LceqError):
        B.set_group_bytes(-1)
    B.close()


def _run(gb, graph, learned, tune=False):
    import lceqcd_amd
    B = _batch(300)
    B.set_group_bytes(gb)
    B.sweep(2, n_or=4, delta=0.4, measflag=0, graph=False)
    if tune:
        B.autotune(n_or=4, trials=2)
    if learned:
        M = lceqcd_amd.Mlp(seed=5)
        B.sweep_mlp(M, 9, n_or=4, measflag=1, graph=graph)
    else:
        B.sweep(9, n_or=4, measflag=1, graph=graph)
    out = (B.history(9), [B.get_links(r) for r in range(len(SHAPES))],
           [B.get_higgs(r) for r in range(len(SHAPES))], B.groups()[1])
    B.close()
    return out


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("learned", [False, True])
def test_groupings_give_the_same_chains(graph, learned):
    total = sum(_field_bytes(s) for s in SHAPES)
    h0, u0, a0, g0 = _run(0, graph, learned)
    assert g0 == [0]
    for gb in (1, total // 3 + 1):
        h1, u1, a1, g1 = _run(gb, graph, learned)
        print("graph", graph, "learned", learned, "groups", g1, flush=True)
        assert len(g1) > 1
        np.testing.assert_array_equal(h1, h0)
        for r in range(len(SHAPES)):
            assert np.array_equal(u1[r], u0[r]), r
            assert np.array_equal(a1[r], a0[r]), r


def test_autotune_keeps_the_groups_and_the_chains():
    total = sum(_field_bytes(s) for s in SHAPES)
    h0, u0, a0, _ = _run(total // 2, True, False)
    h1, u1, a1, g1 = _run(total // 2, True, False, tune=True)
    assert len(g1) == 2
    # the tuned shape may sum the observables in another order
    np.testing.assert_allclose(h1[:, :, :12], h0[:, :, :12], rtol=1e-13, atol=1e-15)
    for r in range(len(SHAPES)):
        assert np.array_equal(u1[r], u0[r]), r
        assert np.array_equal(a1[r], a0[r]), r
