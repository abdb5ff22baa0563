"""The multi-job workloads at full size against the CPU oracle's golden roots
(tests/golden/make_golden.py) and, where no root is small enough to
recompute, against a verifier written from the reference's FRI circuit:

  - the served stream: KoalaBear and BabyBear jobs of 2^16..2^20 rows and
    widths up to 512 at log_blowup 1 and 3, submitted together to the commit
    service (vm/src/instances/machine/riscv.rs:93-129 hands shards to the
    prover the same way); every root equals the oracle's and the job's root
    alone, and 50 openings per job verify against the oracle's root; the
    same jobs again as host arrays, their upload part of each job;
  - 64 batched single-matrix jobs (log_h 10..18, widths 8..512, blowup 2..8):
    graph-replayed batch and one-by-one commits both equal the oracle;
  - one shard: 14 traces of 2^14..2^22 rows in one mixed-height commit and
    the 4 quotient chunks (strided views) on their coset domains, both roots
    equal to the oracle's; then the whole Pcs::open (both rounds at zeta and
    the next-row point, FRI down to the blowup, PoW 16, 100 queries) checked
    by tests/verifier.py.
"""
import json
import os

import numpy as np
import pytest

import pico_amd
from pico_amd import KOALABEAR, workloads
from tests.verifier import verify_open

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _golden(name):
    with open(os.path.join(HERE, "golden", name)) as fh:
        return json.load(fh)


def _alone(ctx, f, seed, L, w, lb):
    m = ctx.random_matrix(f, seed, 1 << L, w)
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
    root, pd = pcs.commit([(pcs.natural_domain_for_degree(1 << L), m)])
    pd.free()
    m.free()
    return root


def test_served_mix_matches_golden(ctx, oracle):
    g = _golden("served_mix.json")
    jobs = [tuple(j) for j in g["jobs_field_log_h_width_log_blowup"]]
    print(f"served mix: {len(jobs)} jobs {jobs}", flush=True)
    svc = pico_amd.CommitService(ctx, budget=0, workers=3)
    try:
        ids = [svc.submit_synthetic(f, (5 << 32) | (20000 + k), 1 << L, w, b) for k, (f, L, w, b) in enumerate(jobs)]
        rng = np.random.default_rng(5)
        for k, (f, L, w, b) in enumerate(jobs):
            root, _ = svc.wait(ids[k], f)
            assert [int(x) for x in root] == g["roots"][k], k
            n = 1 << (L + b)
            q = rng.integers(0, n, 50)
            rows, paths = svc.open(ids[k], f, q, w, L + b)
            for j in range(50):
                assert oracle.mmcs_verify(f, [n], [w], g["roots"][k], int(q[j]), [rows[j]], paths[j]), (k, j)
            svc.release(ids[k])
        stats = svc.stats()
    finally:
        svc.close()
    assert stats["peak_running"] >= 2  # jobs overlapped on the device
    assert stats["reserved_bytes"] == 0
    for k, (f, L, w, b) in enumerate(jobs):  # the same result alone
        assert [int(x) for x in _alone(ctx, f, (5 << 32) | (20000 + k), L, w, b)] == g["roots"][k], k


def test_served_mix_from_host_matches_golden(ctx, oracle):
    """The same served jobs arriving as host arrays (the emulator's records
    reach the prover in host memory): the upload is part of each job, and
    every root still equals the oracle's."""
    g = _golden("served_mix.json")
    jobs = [tuple(j) for j in g["jobs_field_log_h_width_log_blowup"]]
    host = []
    for k, (f, L, w, _) in enumerate(jobs):
        m = ctx.random_matrix(f, (5 << 32) | (20000 + k), 1 << L, w)
        host.append(m.to_host())  # Montgomery words
        m.free()
    svc = pico_amd.CommitService(ctx, budget=0, workers=3)
    try:
        ids = [svc.submit_host(f, host[k], b, canonical=False) for k, (f, L, w, b) in enumerate(jobs)]
        for k, (f, L, w, b) in enumerate(jobs):
            root, _ = svc.wait(ids[k], f)
            assert [int(x) for x in root] == g["roots"][k], k
            q = np.random.default_rng(k).integers(0, 1 << (L + b), 8)
            rows, paths = svc.open(ids[k], f, q, w, L + b)
            for j in range(8):
                assert oracle.mmcs_verify(f, [1 << (L + b)], [w], g["roots"][k], int(q[j]), [rows[j]], paths[j])
            svc.release(ids[k])
        assert svc.stats()["peak_running"] >= 2
    finally:
        svc.close()


def test_batched_jobs_match_golden(ctx):
    g = _golden("batched_kb_64jobs.json")
    shapes = [tuple(s) for s in g["jobs_log_h_width_log_blowup"]]
    assert shapes == workloads.batched_shapes()
    f = KOALABEAR
    mats = [ctx.random_matrix(f, (4 << 32) | (1000 + j), 1 << L, w) for j, (L, w, _) in enumerate(shapes)]
    try:
        batch = pico_amd.CommitBatch(ctx, f, [(m, lb) for m, (_, _, lb) in zip(mats, shapes)], use_graph=True)
        for rep in range(2):  # replays of the captured graph
            roots = batch.run()
            for j in range(len(shapes)):
                assert [int(x) for x in roots[j]] == g["roots"][j], (rep, j)
        batch.free()
        for j, (m, (L, w, lb)) in enumerate(zip(mats, shapes)):
            pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
            root, pd = pcs.commit([(pcs.natural_domain_for_degree(1 << L), m)])
            pd.free()
            assert [int(x) for x in root] == g["roots"][j], j
    finally:
        for m in mats:
            m.free()


def test_shard_commit_and_open_full_size(ctx, oracle):
    g = _golden("shard_kb_14.json")
    shapes = [tuple(s) for s in g["main_log_h_width"]]
    assert shapes == workloads.shard_shapes()
    f, lb = KOALABEAR, g["log_blowup"]
    p = pico_amd.PRIMES[f]
    gen = pico_amd.GENERATORS[f]
    print(f"shard: {len(shapes)} matrices {shapes}", flush=True)
    mats = [ctx.random_matrix(f, (3 << 32) | (100 + i), 1 << L, w) for i, (L, w) in enumerate(shapes)]
    qd = g["quotient"]
    q = ctx.random_matrix(f, qd["seed"], qd["rows"], qd["cols"])
    main = quot = None
    try:
        pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
        root_main, main = pcs.commit([(pcs.natural_domain_for_degree(m.rows), m) for m in mats])
        assert [int(x) for x in root_main] == g["main_root"]
        qdom = pico_amd.TwoAdicMultiplicativeCoset(qd["rows"].bit_length() - 1, gen, f)
        chunks = qdom.split_evals(qd["chunks"], q)
        root_q, quot = pcs.commit(list(zip(qdom.split_domains(qd["chunks"]), chunks)))
        assert [int(x) for x in root_q] == g["quotient_root"]
        for c in chunks:
            c.free()
        # Pcs::open of both rounds: traces at zeta and zeta * w_h (next row),
        # quotient chunks at zeta
        ch = pico_amd.DuplexChallenger(f)
        ch.observe(root_main)
        ch.observe(root_q)
        zeta = np.asarray(ch.sample_ext(), np.uint64)
        main_pts = [np.stack([zeta, zeta * pico_amd.two_adic_generator(f, L) % p]).astype(np.uint32)
                    for L, _ in shapes]
        quot_pts = [zeta.astype(np.uint32)[None, :]] * qd["chunks"]
        rounds = [(main, main_pts), (quot, quot_pts)]
        nq, pow_bits = 100, 16
        opened, proof = pcs.open(rounds, ch, num_queries=nq, pow_bits=pow_bits)
        vch = oracle.Challenger(f)
        vch.observe(g["main_root"])
        vch.observe(g["quotient_root"])
        assert [int(x) for x in vch.sample_ext()] == [int(x) for x in zeta]
        mm = pico_amd.MerkleTreeMmcs(ctx, f)
        vr = [(np.asarray(g["main_root"], np.uint32), [(m.cols, m.rows) for m in mm.get_matrices(main)], main_pts),
              (np.asarray(g["quotient_root"], np.uint32), [(m.cols, m.rows) for m in mm.get_matrices(quot)],
               quot_pts)]
        idx = verify_open(oracle, f, vch, vr, opened, proof, lb, nq, pow_bits)
        assert len(set(idx)) > 90
    finally:
        for x in [main, quot, q] + mats:
            if x is not None:
                x.free()
