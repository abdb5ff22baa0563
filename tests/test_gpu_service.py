"""Commit service: jobs submitted together overlap on the device, each result
equals the same job committed alone, openings come from the resident prover
data, and HBM admission serialises or refuses jobs by footprint."""
import numpy as np
import pytest

import pico_amd
from pico_amd import BABYBEAR, KOALABEAR

pytestmark = pytest.mark.gpu

JOBS = [(KOALABEAR, 11, 16, 1), (BABYBEAR, 12, 40, 3), (KOALABEAR, 13, 24, 1), (BABYBEAR, 10, 8, 1),
        (KOALABEAR, 12, 100, 3), (KOALABEAR, 14, 16, 1), (BABYBEAR, 13, 20, 1), (KOALABEAR, 9, 300, 2)]


def alone(ctx, f, seed, L, w, lb):
    m = ctx.random_matrix(f, seed, 1 << L, w)
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
    root, pd = pcs.commit([(pcs.natural_domain_for_degree(1 << L), m)])
    pd.free()
    return root


def test_service_mix_equals_alone_and_opens(ctx, oracle):
    svc = pico_amd.CommitService(ctx, budget=0, workers=3)
    ids = [svc.submit_synthetic(f, 7000 + i, 1 << L, w, lb) for i, (f, L, w, lb) in enumerate(JOBS)]
    roots = [svc.wait(jid, JOBS[i][0])[0] for i, jid in enumerate(ids)]
    for i, (f, L, w, lb) in enumerate(JOBS):
        assert np.array_equal(roots[i], alone(ctx, f, 7000 + i, L, w, lb)), i
    # openings from resident data verify against the root
    rng = np.random.default_rng(3)
    for i, (f, L, w, lb) in enumerate(JOBS):
        n = 1 << (L + lb)
        q = rng.integers(0, n, 50)
        rows, paths = svc.open(ids[i], f, q, w, L + lb)
        mm = pico_amd.MerkleTreeMmcs(ctx, f)
        for k in range(0, 50, 7):
            assert mm.verify_batch(roots[i], [(w, n)], int(q[k]), [rows[k]], paths[k])
        svc.release(ids[i])
    assert svc.stats()["reserved_bytes"] == 0
    svc.close()


def test_service_host_traces_equal_device_jobs(ctx, oracle):
    """Jobs submitted as host arrays (canonical values and Montgomery words)
    interleaved with device jobs: every root equals the job committed alone
    (and the oracle's up to 2^13 rows), and openings from the resident data
    verify against that root on the oracle."""
    rng = np.random.default_rng(17)
    svc = pico_amd.CommitService(ctx, budget=0, workers=3)
    specs = [(KOALABEAR, 12, 40, 1, True), (BABYBEAR, 11, 24, 3, False), (KOALABEAR, 10, 300, 2, True),
             (BABYBEAR, 13, 9, 1, True), (KOALABEAR, 16, 128, 1, False)]
    try:
        ids, want, held = [], [], []
        for i, (f, L, w, lb, canon) in enumerate(specs):
            p = pico_amd.PRIMES[f]
            vals = rng.integers(0, p, (1 << L, w)).astype(np.uint32)
            arr = vals if canon else pico_amd.to_monty(f, vals).astype(np.uint32)
            held.append(vals)
            ids.append(svc.submit_host(f, arr, lb, canonical=canon))
            del arr  # the service keeps its own reference until wait()
            # a device job between host jobs
            ids.append(svc.submit_synthetic(f, 4100 + i, 1 << 10, 16, 1))
            want.append(oracle.MerkleTree(f, [oracle.coset_lde(f, vals, lb, pico_amd.GENERATORS[f])]).root
                        if L <= 13 else None)
        for i, (f, L, w, lb, canon) in enumerate(specs):
            root, _ = svc.wait(ids[2 * i], f)
            m = ctx.upload(held[i], field=f)
            pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
            alone_root, pd = pcs.commit([(pcs.natural_domain_for_degree(1 << L), m)])
            pd.free()
            m.free()
            assert np.array_equal(root, alone_root), i
            if want[i] is not None:
                assert [int(x) for x in root] == [int(x) for x in want[i]], i
            n = 1 << (L + lb)
            q = rng.integers(0, n, 12)
            rows, paths = svc.open(ids[2 * i], f, q, w, L + lb)
            for k in range(12):
                assert oracle.mmcs_verify(f, [n], [w], [int(x) for x in root], int(q[k]), [rows[k]], paths[k])
            assert np.array_equal(svc.wait(ids[2 * i + 1], f)[0], alone(ctx, f, 4100 + i, 10, 16, 1))
        for jid in ids:
            svc.release(jid)
        assert svc.stats()["reserved_bytes"] == 0
        assert not svc._held
    finally:
        svc.close()


def test_service_host_trace_rejects_bad_shapes(ctx):
    svc = pico_amd.CommitService(ctx, budget=0, workers=1)
    try:
        with pytest.raises(pico_amd.PgError):  # not a power of two
            svc.submit_host(KOALABEAR, np.zeros((12, 4), np.uint32), 1)
        with pytest.raises(pico_amd.PgError):  # LDE beyond the two-adicity
            svc.submit_host(BABYBEAR, np.zeros((1 << 10, 1), np.uint32), 20)
    finally:
        svc.close()


def test_service_small_budget_serialises_and_refuses(ctx):
    f = KOALABEAR
    # one 2^12 x 16 job at blowup 2 needs ~2.4 MiB; budget fits one at a time
    svc = pico_amd.CommitService(ctx, budget=3 << 20, workers=3)
    ids = [svc.submit_synthetic(f, 900 + i, 1 << 12, 16, 1) for i in range(4)]
    big = svc.submit_synthetic(f, 999, 1 << 16, 64, 1)
    for i, jid in enumerate(ids):
        root, (ts, t0, t1) = svc.wait(jid, f)
        assert np.array_equal(root, alone(ctx, f, 900 + i, 12, 16, 1))
        svc.release(jid)  # frees the reservation for the next job
    assert svc.stats()["peak_running"] == 1
    with pytest.raises(pico_amd.AdmissionError):
        svc.wait(big, f)
    svc.release(big)
    svc.close()


def test_service_overlaps_jobs(ctx):
    svc = pico_amd.CommitService(ctx, budget=0, workers=3)
    ids = [svc.submit_synthetic(KOALABEAR, 50 + i, 1 << 16, 64, 1) for i in range(9)]
    for jid in ids:
        svc.wait(jid, KOALABEAR)
    assert svc.stats()["peak_running"] >= 2
    for jid in ids:
        svc.release(jid)
    svc.close()
