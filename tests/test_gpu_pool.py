"""Device memory pool behind every allocation of a context: blocks of any
size are carved from shared segments, freed blocks merge and are handed out
again.  No live block may overlap another (each keeps its seeded contents
through a long random sequence of allocations and frees), the live-byte count
returns to its starting value, and commits through a small-budget context
(forced segment release and regrowth) still produce the oracle's roots."""
import numpy as np
import pytest

import pico_amd
from pico_amd import KOALABEAR

pytestmark = pytest.mark.gpu


def test_pool_blocks_never_overlap(oracle):
    f = KOALABEAR
    rng = np.random.default_rng(7)
    with pico_amd.Context(0) as ctx:
        base = ctx.bytes_in_use
        live = {}
        seed = 0
        for step in range(400):
            if live and (rng.random() < 0.45 or len(live) > 40):
                k = list(live)[rng.integers(len(live))]
                live.pop(k)[0].free()
                continue
            # sizes from 512 B to ~24 MiB, many of them not multiples of the alignment
            rows = int(rng.integers(1, 1 << int(rng.integers(1, 17))))
            cols = int(rng.integers(1, 97))
            seed += 1
            live[seed] = (ctx.random_matrix(f, seed, rows, cols), rows, cols)
            if step % 50 == 49:
                for s, (m, r, c) in live.items():
                    assert np.array_equal(m.to_host(field=f), oracle.random_matrix(f, s, r, c)), (step, s)
        for s, (m, r, c) in live.items():
            assert np.array_equal(m.to_host(field=f), oracle.random_matrix(f, s, r, c)), s
        for m, _, _ in live.values():
            m.free()
        assert ctx.bytes_in_use == base


def test_pool_large_blocks_reused_across_shapes(oracle):
    """Big buffers of changing shapes in sequence: later, smaller requests are
    carved from the freed space and every commit still matches the oracle."""
    f = KOALABEAR
    with pico_amd.Context(0) as ctx:
        for i, (L, w, lb) in enumerate([(14, 200, 3), (15, 64, 2), (13, 150, 3), (12, 333, 1), (14, 8, 2)]):
            host = oracle.random_matrix(f, 50 + i, 1 << L, w)
            pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
            root, pd = pcs.commit([(pcs.natural_domain_for_degree(1 << L), ctx.upload(host, field=f))])
            lde = oracle.coset_lde(f, host, lb, pico_amd.GENERATORS[f])
            assert np.array_equal(root, oracle.MerkleTree(f, [lde]).root), (L, w, lb)
            pd.free()


def test_pool_small_budget_context(oracle):
    """A budget below two of the jobs' footprints: allocation beyond it fails
    with the admission error, and work within it keeps succeeding."""
    f = KOALABEAR
    budget = 96 << 20
    with pico_amd.Context(0, hbm_budget=budget) as ctx:
        host = oracle.random_matrix(f, 9, 1 << 12, 64)
        pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
        for _ in range(3):
            root, pd = pcs.commit([(pcs.natural_domain_for_degree(1 << 12), ctx.upload(host, field=f))])
            pd.free()
        lde = oracle.coset_lde(f, host, 1, pico_amd.GENERATORS[f])
        assert np.array_equal(root, oracle.MerkleTree(f, [lde]).root)
        with pytest.raises(pico_amd.AdmissionError):
            ctx.matrix(1 << 20, 64)  # 256 MiB > budget


def test_pinned_host_upload(oracle):
    """A trace written into page-locked memory from pg_host_alloc uploads and
    commits to the same root as the same trace from pageable memory."""
    f = KOALABEAR
    with pico_amd.Context(0) as ctx:
        host = oracle.random_matrix(f, 77, 1 << 12, 37)
        pinned = ctx.host_matrix(*host.shape)
        pinned.array[:] = host
        pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
        dom = pcs.natural_domain_for_degree(1 << 12)
        r_pin, pd1 = pcs.commit([(dom, ctx.upload(pinned.array, field=f))])
        r_page, pd2 = pcs.commit([(dom, ctx.upload(host, field=f))])
        lde = oracle.coset_lde(f, host, 1, pico_amd.GENERATORS[f])
        assert np.array_equal(r_pin, r_page)
        assert np.array_equal(r_pin, oracle.MerkleTree(f, [lde]).root)
        pd1.free()
        pd2.free()
        pinned.free()
        empty = ctx.host_matrix(0, 5)
        assert empty.array.shape == (0, 5)
        empty.free()
