"""A verifier of TwoAdicFriPcs::open proofs for the GPU tests, written from
the reference's recursion circuit (vm/src/compiler/recursion/circuit/fri.rs):
transcript order (:36-73, :84-99), reduced openings recomputed from the
opened rows (:118-241; the row of matrix k at query q sits at
x = GEN w^rev(q >> (log_max - log_h)), :151-171), the fold line (:288-387)
and verify_batch (:389-436).  The challenger and the Merkle checks are the
CPU oracle's; the sums over columns are numpy-vectorised so proofs of
shard-sized commitments (thousands of columns, 100 queries) verify in
seconds.  All values canonical.
"""
import numpy as np

from tests import pyref


def ext_mul_vec(f, a, b):
    """Row-wise ext4 products of two (n, 4) arrays of canonical values."""
    p, w = np.uint64(pyref.P[f]), np.uint64(pyref.W[f])
    a = np.asarray(a, np.uint64)
    b = np.asarray(b, np.uint64)
    c = np.zeros(a.shape, np.uint64)
    for i in range(4):
        for j in range(4):
            t = a[:, i] * b[:, j] % p
            if i + j < 4:
                c[:, i + j] += t
            else:
                c[:, i + j - 4] += w * t % p
    return c % p


def ext_inv(f, a):
    p = pyref.P[f]
    r, e, b = [1, 0, 0, 0], p ** 4 - 2, [int(x) for x in a]
    while e:
        if e & 1:
            r = pyref.ext_mul(f, r, b)
        b = pyref.ext_mul(f, b, b)
        e >>= 1
    assert pyref.ext_mul(f, [int(x) for x in a], r) == [1, 0, 0, 0]
    return r


def alpha_powers(f, alpha, n):
    out = np.zeros((max(n, 1), 4), np.uint64)
    cur = [1, 0, 0, 0]
    for k in range(n):
        out[k] = cur
        cur = pyref.ext_mul(f, cur, alpha)
    return out


def verify_open(oracle, f, ch, rounds, opened, proof, log_blowup, num_queries, pow_bits, log_final=None):
    """rounds: [(root[8], dims [(width, lde_height)] in commit order,
    points [npts x 4 per matrix])]; opened: per round, per matrix, an
    npts x width x 4 array; proof: the dict TwoAdicFriPcs.open returns; ch:
    an oracle.Challenger in the state the prover's challenger had when
    Pcs::open began.  Asserts every check; returns the query indices."""
    p = pyref.P[f]
    # transcript: values, alpha, one beta per committed round, final, PoW, indices
    for per_round in opened:
        for v in per_round:
            ch.observe(np.asarray(v, np.uint32).ravel())
    alpha = [int(a) for a in ch.sample_ext()]
    assert alpha == [int(a) for a in proof["alpha"]]
    R = len(proof["commits"])
    for r in range(R):
        ch.observe(np.asarray(proof["commits"][r], np.uint32))
        assert [int(b) for b in ch.sample_ext()] == [int(b) for b in proof["betas"][r]], r
    final = np.asarray(proof["final_poly"]).reshape(-1, 4)
    if log_final is None:
        assert np.all(final == final[0])  # the reference stops at 2^log_blowup with a constant vector
        ch.observe(final[0])
    else:
        ch.observe(final.ravel())
    assert ch.check_witness(pow_bits, proof["pow_witness"])
    log_max = max((h.bit_length() - 1) for _, dims, _ in rounds for _, h in dims)
    stop = log_blowup if log_final is None else log_final
    assert R == log_max - stop
    idx_all = [ch.sample_bits(log_max) for _ in range(num_queries)]
    assert idx_all == [int(i) for i in proof["query_indices"]]
    # alpha powers: one counter per LDE log-height, advanced per column
    need = {}
    for _, dims, pts in rounds:
        for (w, h), zs in zip(dims, pts):
            lh = h.bit_length() - 1
            need[lh] = need.get(lh, 0) + w * len(np.asarray(zs).reshape(-1, 4))
    ap = alpha_powers(f, alpha, max(need.values()))
    for qi, idx in enumerate(idx_all):
        ro, ctr, inv_cache = {}, {}, {}
        for r, (root, dims, pts) in enumerate(rounds):
            rows, path = proof["input_openings"][qi][r]
            lr = max(h for _, h in dims).bit_length() - 1
            assert oracle.mmcs_verify(f, [h for _, h in dims], [w for w, _ in dims], root, idx >> (log_max - lr),
                                      rows, path), (qi, r)
            for k, (w, h) in enumerate(dims):
                lh = h.bit_length() - 1
                i = idx >> (log_max - lh)
                x = pyref.GEN[f] * pow(pyref.root_of_unity(f, lh), pyref.rev(i, lh), p) % p
                row = np.asarray(rows[k], np.uint64)
                for q, z in enumerate(np.asarray(pts[k]).reshape(-1, 4)):
                    c0 = ctr.get(lh, 0)
                    v = np.asarray(opened[r][k][q], np.uint64).reshape(w, 4).copy()
                    v[:, 0] = (v[:, 0] + np.uint64(p) - row) % np.uint64(p)
                    num = ext_mul_vec(f, ap[c0:c0 + w], v).sum(axis=0) % np.uint64(p)
                    ctr[lh] = c0 + w
                    key = (lh, tuple(int(t) for t in z))
                    if key not in inv_cache:
                        inv_cache[key] = ext_inv(f, [(int(z[0]) - x) % p, int(z[1]), int(z[2]), int(z[3])])
                    term = pyref.ext_mul(f, [int(t) for t in num], inv_cache[key])
                    ro[lh] = [(a + b) % p for a, b in zip(ro.get(lh, [0, 0, 0, 0]), term)]
        # fold chain through the commit-phase openings down to the final value
        cur = ro[log_max]
        for r in range(R):
            L = log_max - r
            ir = idx >> r
            s = [int(t) for t in proof["fri_siblings"][qi][r]]
            pair = [cur, s] if ir % 2 == 0 else [s, cur]
            leaf = np.array(pair[0] + pair[1], np.uint32)
            assert oracle.mmcs_verify(f, [1 << (L - 1)], [8], proof["commits"][r], ir >> 1, [leaf],
                                      proof["fri_paths"][qi][r]), (qi, r)
            x = pow(pyref.root_of_unity(f, L), pyref.rev(ir >> 1, L - 1), p)
            cur = pyref.fold_pair(f, pair[0], pair[1], [int(b) for b in proof["betas"][r]], x)
            if L - 1 in ro:
                cur = [(a + b) % p for a, b in zip(cur, ro[L - 1])]
        fi = (idx >> R) if log_final is not None else 0
        assert cur == [int(t) for t in final[fi]], qi
    return idx_all
