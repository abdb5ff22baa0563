"""One small shard end to end on the device, checked the way the verifier
would (vm/src/compiler/recursion/circuit/fri.rs:75-436): mixed-height main
commit, quotient commit on a shifted domain, FRI commit phase, proof of work,
query openings; every opening verifies and every query's fold chain lands on
the next round's committed value and finally on the final vector."""
import numpy as np
import pytest

import pico_amd
from pico_amd import KOALABEAR
from tests import pyref

pytestmark = pytest.mark.gpu


def test_small_shard_commit_fri_queries(ctx, oracle):
    f = KOALABEAR
    p = pico_amd.PRIMES[f]
    shapes = [(9, 12), (7, 8), (9, 3), (6, 30), (8, 1)]
    hosts = [oracle.random_matrix(f, 300 + i, 1 << L, w) for i, (L, w) in enumerate(shapes)]
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
    root, main = pcs.commit([(pcs.natural_domain_for_degree(1 << L), ctx.upload(a, field=f))
                             for (L, _), a in zip(shapes, hosts)])
    # oracle: same LDEs, same mixed-height tree
    ldes = [oracle.coset_lde(f, a, 1, pico_amd.GENERATORS[f]) for a in hosts]
    assert np.array_equal(root, oracle.MerkleTree(f, ldes).root)
    # quotient chunk on a disjoint coset, blowup 8
    qh = oracle.random_matrix(f, 400, 1 << 7, 4)
    qshift = pico_amd.GENERATORS[f] * pico_amd.two_adic_generator(f, 10) % p
    qroot, quot = pico_amd.TwoAdicFriPcs(ctx, f, 3).commit(
        [(pico_amd.TwoAdicMultiplicativeCoset(7, qshift), ctx.upload(qh, field=f))])
    qlde = oracle.coset_lde(f, qh, 3, pico_amd.GENERATORS[f] * pow(qshift, p - 2, p) % p)
    assert np.array_equal(qroot, oracle.MerkleTree(f, [qlde]).root)

    log_n, log_stop = 10, 2
    v = oracle.random_matrix(f, 500, 1 << log_n, 4)
    ch = pico_amd.DuplexChallenger(f)
    for c in (root, qroot):
        ch.observe(c)
    fri = pico_amd.fri_commit_phase(ctx, f, ch, {log_n: ctx.upload(v, field=f)}, log_stop, final_mode=1)
    w = pico_amd.fri_grind(ctx, f, ch, 10)
    idx = [ch.sample_bits(log_n) for _ in range(30)]
    sib, paths = pico_amd.fri_answer_queries(fri, idx, log_n)
    mm = pico_amd.MerkleTreeMmcs(ctx, f)

    # transcript replay on the oracle side
    och = oracle.Challenger(f)
    for c in (root, qroot):
        och.observe(c)
    roots, layers, final, betas = oracle.fri_commit_phase(och, v, log_stop)
    och.observe(final.ravel())
    assert och.check_witness(10, w)
    assert [och.sample_bits(log_n) for _ in range(30)] == idx
    assert np.array_equal(fri.commits, roots) and np.array_equal(fri.betas, betas)

    dims = [(wd, 1 << (L + 1)) for L, wd in shapes]
    for q, i in enumerate(idx):
        # input openings (the main tree is 2^10 tall here, like the FRI vector)
        rows, path = mm.open_batch(i, main)
        assert oracle.mmcs_verify(f, [d[1] for d in dims], [d[0] for d in dims], root, i, rows, path)
        # fold chain: value at i, sibling, beta -> value at i >> 1 of the next round
        cur = [int(x) for x in layers[0][i]]
        for r in range(log_n - log_stop):
            ir = i >> r
            s = [int(x) for x in sib[q, r]]
            pair = [cur, s] if ir % 2 == 0 else [s, cur]
            leaf = np.array(pair[0] + pair[1], np.uint32)
            assert oracle.mmcs_verify(f, [1 << (log_n - r - 1)], [8], fri.commits[r], ir >> 1, [leaf],
                                      paths[q][r])
            x = pow(pyref.root_of_unity(f, log_n - r), pyref.rev(ir >> 1, log_n - r - 1), p)
            cur = pyref.fold_pair(f, pair[0], pair[1], [int(b) for b in betas[r]], x)
            nxt = layers[r + 1][ir >> 1] if r + 1 < len(layers) else final[ir >> 1]
            assert cur == [int(t) for t in nxt], (q, r)
    main.free()
    quot.free()
    fri.free()
