"""Batched single-matrix commits (one launch per LDE pass / Merkle level over
all jobs, optionally replayed as a hipGraph): every root and LDE equals the
one-at-a-time path and the oracle."""
import numpy as np
import pytest

import pico_amd
from pico_amd import BABYBEAR, KOALABEAR

pytestmark = pytest.mark.gpu

JOBS = [(3, 8, 1), (6, 13, 2), (10, 64, 1), (12, 8, 3), (13, 100, 1), (10, 5, 2), (12, 33, 1),
        (0, 4, 2), (11, 16, 3), (14, 9, 1), (13, 8, 2)]


@pytest.mark.parametrize("f", [KOALABEAR, BABYBEAR])
@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_sequential_and_oracle(ctx, oracle, f, use_graph):
    hosts = [oracle.random_matrix(f, 4000 + i, 1 << L, w) for i, (L, w, _) in enumerate(JOBS)]
    mats = [ctx.upload(a, field=f) for a in hosts]
    batch = pico_amd.CommitBatch(ctx, f, [(m, lb) for m, (_, _, lb) in zip(mats, JOBS)], use_graph=use_graph)
    roots = batch.run()
    again = batch.run()  # replay: same outputs
    assert np.array_equal(roots, again)
    g = pico_amd.GENERATORS[f]
    for j, (a, (L, w, lb)) in enumerate(zip(hosts, JOBS)):
        pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
        r_seq, pd = pcs.commit([(pcs.natural_domain_for_degree(1 << L), mats[j])])
        pd.free()
        assert np.array_equal(roots[j], r_seq), j
        if L <= 12:
            lde = oracle.coset_lde(f, a, lb, g)
            assert np.array_equal(roots[j], oracle.MerkleTree(f, [lde]).root), j
    # prover data handed out: LDE and openings
    for j in (2, 4):
        L, w, lb = JOBS[j]
        pd = batch.take(j)
        mm = pico_amd.MerkleTreeMmcs(ctx, f)
        lde = mm.get_matrices(pd)[0].to_host(f)
        assert np.array_equal(lde, oracle.coset_lde(f, hosts[j], lb, g))
        rows, path = mm.open_batch(5, pd)
        assert mm.verify_batch(roots[j], [(w, 1 << (L + lb))], 5, rows, path)
        pd.free()
    with pytest.raises(pico_amd.PgError):
        batch.run()  # trees were handed out
    batch.free()


def test_batch_launch_count_is_small(ctx, oracle):
    f = KOALABEAR
    mats = [ctx.random_matrix(f, 50 + i, 1 << (10 + i % 5), 8 + 3 * i) for i in range(32)]
    batch = pico_amd.CommitBatch(ctx, f, [(m, 1 + i % 3) for i, m in enumerate(mats)], use_graph=True)
    # 5 distinct heights: 2 single-pass + 3 x 3 passes, 1 leaf, <= 17 levels, 1 gather
    assert batch.num_launches <= 2 + 9 + 1 + 17 + 1
    batch.run()
    batch.free()
