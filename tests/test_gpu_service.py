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
