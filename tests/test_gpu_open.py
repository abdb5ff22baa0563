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


def test_open_flow_verifies(ctx, oracle):
    """The whole open on a two-round commitment, then the verifier's checks at
    every query: input openings verify, and the reduced opening recomputed
    from the opened rows equals the folded vector's entry."""
    f = KOALABEAR
    p = pico_amd.PRIMES[f]
    lb = 1
    pcs, root_main, main, _ = _commit(ctx, oracle, f, [(9, 16), (8, 6), (9, 3)], lb, 1000)
    qshift = pico_amd.GENERATORS[f] * pico_amd.two_adic_generator(f, 10) % p
    _, root_q, quot, _ = _commit(ctx, oracle, f, [(9, 4)], lb, 1100, shift=qshift)
    ch = pico_amd.DuplexChallenger(f)
    ch.observe(root_main)
    ch.observe(root_q)
    zeta = ch.sample_ext()
    gz = (np.asarray(zeta, np.uint64) * pico_amd.two_adic_generator(f, 9) % p).astype(np.uint32)
    rounds = [(main, [np.stack([zeta, gz])] * 3), (quot, [np.asarray(zeta, np.uint32)[None, :]])]
    opened, proof = pcs.open(rounds, ch, num_queries=12, pow_bits=6)
    ro = proof["reduced_openings"]
    log_max = max(ro)
    ro_host = {lh: v.to_host(f) for lh, v in ro.items()}
    alpha = [int(a) for a in proof["alpha"]]
    mm = pico_amd.MerkleTreeMmcs(ctx, f)
    for qi, idx in enumerate(proof["query_indices"]):
        # recompute every height's reduced opening at this query from the opened rows
        acc = {}
        pow_ctr = {}
        for r, (d, mat_pts) in enumerate(rounds):
            rows, path = proof["input_openings"][qi][r]
            dims = [(m.cols, m.rows) for m in mm.get_matrices(d)]
            ridx = idx >> (log_max - d.log_height)
            assert mm.verify_batch(d.root(), dims, ridx, rows, path)
            for k, (w, n) in enumerate(dims):
                lh = n.bit_length() - 1
                i = idx >> (log_max - lh)
                x = pyref.GEN[f] * pow(pyref.root_of_unity(f, lh), pyref.rev(i, lh), p) % p
                for q, z in enumerate(np.asarray(mat_pts[k]).reshape(-1, 4)):
                    num = [0, 0, 0, 0]
                    for c in range(w):
                        ap = [1, 0, 0, 0]
                        for _ in range(pow_ctr.get(lh, 0)):
                            ap = pyref.ext_mul(f, ap, alpha)
                        v = [int(t) for t in opened[r][k][q, c]]
                        diff = [(v[0] - int(rows[k][c])) % p] + v[1:]
                        num = [(a + b) % p for a, b in zip(num, pyref.ext_mul(f, ap, diff))]
                        pow_ctr[lh] = pow_ctr.get(lh, 0) + 1
                    den = [(int(z[0]) - x) % p, int(z[1]), int(z[2]), int(z[3])]
                    acc.setdefault(lh, []).append((num, den, i))
        for lh, terms in acc.items():
            i = terms[0][2]
            got = [int(t) for t in ro_host[lh][i]]
            # sum num_t / den_t == got  <=>  got * prod den == sum num_t prod_{u != t} den_u
            lhs = got
            for _, den, _ in terms:
                lhs = pyref.ext_mul(f, lhs, den)
            rhs = [0, 0, 0, 0]
            for t, (num, _, _) in enumerate(terms):
                term = num
                for u, (_, den, _) in enumerate(terms):
                    if u != t:
                        term = pyref.ext_mul(f, term, den)
                rhs = [(a + b) % p for a, b in zip(rhs, term)]
            assert lhs == rhs, (qi, lh)
    # the reference's final vector is constant after folding to 2^log_blowup
    assert np.all(proof["final_poly"] == proof["final_poly"][0])
    proof["fri_data"].free()
    main.free()
    quot.free()
