"""Pcs::open on the device against the oracle: opened values at out-of-domain
points, reduced openings per LDE height, and the whole open flow (values ->
challenger -> alpha -> reduced openings -> FRI -> PoW -> queries) checked the
way the verifier does (circuit/fri.rs:84-241): the input openings of every
query verify against their round commitments and reproduce the reduced
opening that FRI folded at that index."""
import numpy as np
import pytest

import pico_amd
from pico_amd import BABYBEAR, KOALABEAR
from tests import pyref
from tests.verifier import verify_open

pytestmark = pytest.mark.gpu
FIELDS = [KOALABEAR, BABYBEAR]


def _commit(ctx, oracle, f, shapes, lb, seed0, shift=1):
    p = pico_amd.PRIMES[f]
    hosts = [oracle.random_matrix(f, seed0 + i, 1 << L, w) for i, (L, w) in enumerate(shapes)]
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
    doms = [pico_amd.TwoAdicMultiplicativeCoset(L, shift) for L, _ in shapes]
    root, pd = pcs.commit([(d, ctx.upload(a, field=f)) for d, a in zip(doms, hosts)])
    lde_shift = pico_amd.GENERATORS[f] * pow(shift, p - 2, p) % p
    ldes = [oracle.coset_lde(f, a, lb, lde_shift) for a in hosts]
    return pcs, root, pd, ldes


@pytest.mark.parametrize("f", FIELDS)
@pytest.mark.parametrize("lb", [1, 3])
def test_open_values_match_oracle(ctx, oracle, f, lb):
    p = pico_amd.PRIMES[f]
    shapes = [(0, 3), (1, 5), (6, 8), (9, 33), (12, 4)]
    pcs, _, pd, ldes = _commit(ctx, oracle, f, shapes, lb, 700)
    rng = np.random.default_rng(9)
    pts = rng.integers(0, p, (2, 4)).astype(np.uint32)
    for k, lde in enumerate(ldes):
        got = pcs.open_values(pd, k, pts)
        for q in range(2):
            assert np.array_equal(got[q], oracle.pcs_open_values(f, lde, lb, pts[q])), (k, q)
    pd.free()


@pytest.mark.parametrize("f", FIELDS)
def test_reduce_openings_match_oracle(ctx, oracle, f):
    """Two rounds (a mixed-height main batch opened at zeta and g zeta, a
    quotient-like batch on a shifted domain opened at zeta), widths that are
    and are not multiples of 4."""
    p = pico_amd.PRIMES[f]
    lb = 1
    pcs, _, main, main_ldes = _commit(ctx, oracle, f, [(10, 12), (8, 7), (10, 1)], lb, 800)
    qshift = pico_amd.GENERATORS[f] * pico_amd.two_adic_generator(f, 11) % p
    _, _, quot, quot_ldes = _commit(ctx, oracle, f, [(10, 8)], lb, 900, shift=qshift)
    rng = np.random.default_rng(11)
    zeta = rng.integers(0, p, 4).astype(np.uint32)
    gz = (zeta.astype(np.uint64) * pico_amd.two_adic_generator(f, 10) % p).astype(np.uint32)
    main_pts = [np.stack([zeta, gz])] * 3
    quot_pts = [zeta[None, :]]
    main_vals = [pcs.open_values(main, k, main_pts[k]) for k in range(3)]
    quot_vals = [pcs.open_values(quot, 0, quot_pts[0])]
    alpha = rng.integers(0, p, 4).astype(np.uint32)
    ro = pcs.reduce_openings([(main, main_pts), (quot, quot_pts)], [main_vals, quot_vals], alpha)
    mats = main_ldes + quot_ldes
    npts = [2, 2, 2, 1]
    pts = np.concatenate(main_pts + quot_pts)
    vals = [v[q] for v in main_vals for q in range(2)] + [quot_vals[0][0]]
    ref = oracle.pcs_reduce(f, mats, npts, pts, vals, alpha)
    assert sorted(ro) == sorted(ref) == [9, 11]
    for lh in ref:
        assert np.array_equal(ro[lh].to_host(f), ref[lh]), lh
    main.free()
    quot.free()


def _ext_inv(f, a):
    p = pico_amd.PRIMES[f]
    r, e, b = [1, 0, 0, 0], p ** 4 - 2, list(a)
    while e:
        if e & 1:
            r = pyref.ext_mul(f, r, b)
        b = pyref.ext_mul(f, b, b)
        e >>= 1
    assert pyref.ext_inv_by_search_check(f, a, r)
    return r


def _two_rounds(ctx, oracle, f):
    p = pico_amd.PRIMES[f]
    pcs, root_main, main, _ = _commit(ctx, oracle, f, [(9, 16), (8, 6), (9, 3)], 1, 1000)
    qshift = pico_amd.GENERATORS[f] * pico_amd.two_adic_generator(f, 10) % p
    _, root_q, quot, _ = _commit(ctx, oracle, f, [(9, 4)], 1, 1100, shift=qshift)
    return pcs, (root_main, main), (root_q, quot)


def _points(f, ch, log_trace):
    p = pico_amd.PRIMES[f]
    zeta = ch.sample_ext()
    gz = (np.asarray(zeta, np.uint64) * pico_amd.two_adic_generator(f, log_trace) % p).astype(np.uint32)
    return zeta, [np.stack([zeta, gz])] * 3, [np.asarray(zeta, np.uint32)[None, :]]


@pytest.mark.parametrize("f", FIELDS)
@pytest.mark.parametrize("log_final", [None, 3])
def test_open_verifies(ctx, oracle, f, log_final):
    """pg_pcs_open on a two-round commitment (a mixed-height main batch at
    zeta and g zeta, a quotient-like batch on a shifted domain at zeta),
    checked by a verifier written from circuit/fri.rs:36-436: transcript
    replay (values observed, alpha, one beta per committed round, the final
    value, the PoW witness, the query indices), every input opening against
    its round root, each height's reduced opening recomputed from the opened
    rows, and every query's fold chain through the commit-phase openings to
    the final value (tests/verifier.py)."""
    p = pico_amd.PRIMES[f]
    pcs, (root_main, main), (root_q, quot) = _two_rounds(ctx, oracle, f)
    ch = pico_amd.DuplexChallenger(f)
    ch.observe(root_main)
    ch.observe(root_q)
    zeta, main_pts, quot_pts = _points(f, ch, 9)
    rounds = [(main, main_pts), (quot, quot_pts)]
    nq, pow_bits = 12, 6
    opened, proof = pcs.open(rounds, ch, num_queries=nq, pow_bits=pow_bits, log_final=log_final)
    vch = oracle.Challenger(f)
    vch.observe(root_main)
    vch.observe(root_q)
    assert [int(x) for x in vch.sample_ext()] == [int(x) for x in zeta]
    mm = pico_amd.MerkleTreeMmcs(ctx, f)
    vr = [(root_main, [(m.cols, m.rows) for m in mm.get_matrices(main)], main_pts),
          (root_q, [(m.cols, m.rows) for m in mm.get_matrices(quot)], quot_pts)]
    verify_open(oracle, f, vch, vr, opened, proof, 1, nq, pow_bits, log_final=log_final)
    if log_final is None:
        # the FriProof's bincode (vm/src/machine/proof.rs:86-101 stores it
        # in every BaseProof) equals an independent encoding of the verified
        # parts, and decodes to them
        queries = [([(rows, path) for rows, path in proof["input_openings"][q]],
                    [(proof["fri_siblings"][q][r], proof["fri_paths"][q][r]) for r in range(len(proof["commits"]))])
                   for q in range(nq)]
        want = pyref.encode_fri_proof(f, proof["commits"], queries, proof["final_poly"][0], proof["pow_witness"],
                                      canonical=False)
        assert proof["bincode"] == want
        d = pico_amd.decode_fri_proof(f, proof["bincode"])
        assert d["reencoded"] == want and d["pow_witness"] == proof["pow_witness"]
        assert np.array_equal(d["commits"], proof["commits"])
    else:
        assert proof["bincode"] is None  # a non-constant final vector has no FriProof form
    # a planted mistake in one opened value is rejected
    bad = [[np.array(v, copy=True) for v in per] for per in opened]
    bad[0][1][0, 0, 0] = (int(bad[0][1][0, 0, 0]) + 1) % p
    vch = oracle.Challenger(f)
    vch.observe(root_main)
    vch.observe(root_q)
    vch.sample_ext()
    with pytest.raises(AssertionError):
        verify_open(oracle, f, vch, vr, bad, proof, 1, nq, pow_bits, log_final=log_final)
    main.free()
    quot.free()


def _open_stepwise(pcs, rounds, ch, num_queries, pow_bits, log_final=None):
    """The same open composed from the single-purpose entry points."""
    ctx, f = pcs.ctx, pcs.field
    opened = [[pcs.open_values(d, k, p) for k, p in enumerate(mp)] for d, mp in rounds]
    for per in opened:
        for v in per:
            ch.observe(np.asarray(v).ravel())
    alpha = ch.sample_ext()
    ro = pcs.reduce_openings(rounds, opened, alpha)
    log_stop = pcs.log_blowup if log_final is None else log_final
    fri = pico_amd.fri_commit_phase(ctx, f, ch, ro, log_stop, final_mode=0 if log_final is None else 1)
    for b in ro.values():
        b.free()
    w = pico_amd.fri_grind(ctx, f, ch, pow_bits)
    log_max = max(ro)
    idx = [ch.sample_bits(log_max) for _ in range(num_queries)]
    mm = pico_amd.MerkleTreeMmcs(ctx, f)
    inputs = [[mm.open_batch(i >> (log_max - d.log_height), d) for d, _ in rounds] for i in idx]
    sib, paths = pico_amd.fri_answer_queries(fri, idx, log_max)
    out = (opened, alpha, fri.commits, fri.betas, fri.final_poly, w, idx, inputs, sib, paths)
    fri.free()
    return out


@pytest.mark.parametrize("f", FIELDS)
@pytest.mark.parametrize("log_final", [None, 3])
def test_open_matches_stepwise(ctx, oracle, f, log_final):
    pcs, (root_main, main), (root_q, quot) = _two_rounds(ctx, oracle, f)
    chs = []
    for _ in range(2):
        ch = pico_amd.DuplexChallenger(f)
        ch.observe(root_main)
        ch.observe(root_q)
        chs.append(ch)
    _, main_pts, quot_pts = _points(f, chs[0], 9)
    _points(f, chs[1], 9)
    rounds = [(main, main_pts), (quot, quot_pts)]
    opened, proof = pcs.open(rounds, chs[0], num_queries=9, pow_bits=5, log_final=log_final)
    s_open, alpha, commits, betas, final, w, idx, inputs, sib, paths = _open_stepwise(pcs, rounds, chs[1], 9, 5,
                                                                                      log_final)
    for a, b in zip(opened, s_open):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    assert np.array_equal(proof["alpha"], alpha)
    assert np.array_equal(proof["commits"], commits) and np.array_equal(proof["betas"], betas)
    assert np.array_equal(proof["final_poly"], final) and proof["pow_witness"] == w
    assert proof["query_indices"] == idx
    for q in range(len(idx)):
        for r in range(2):
            rows_a, path_a = proof["input_openings"][q][r]
            rows_b, path_b = inputs[q][r]
            assert all(np.array_equal(x, y) for x, y in zip(rows_a, rows_b))
            assert np.array_equal(path_a, path_b)
        for r in range(len(commits)):
            assert np.array_equal(proof["fri_paths"][q][r], paths[q][r])
    assert np.array_equal(proof["fri_siblings"], sib)
    # both challengers end in the same state
    assert chs[0].sample_bits(20) == chs[1].sample_bits(20)
    main.free()
    quot.free()
