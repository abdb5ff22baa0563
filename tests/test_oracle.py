"""Pins the CPU oracle before anything is compared against it.

Each check is independent of the oracle's own code: the reference-held
Poseidon2 KATs, Python big-int arithmetic, the polynomial definition of the
LDE, the verify_batch walk, the fold line formula, closed forms (shift = 1
LDE reproduces its input; folding a low-degree vector stays low degree) and
planted mistakes that must be caught.
"""
import json
import os

import numpy as np
import pytest

from tests import pyref

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = [0, 1]


def kats():
    with open(os.path.join(HERE, "golden", "poseidon2_kat.json")) as f:
        return json.load(f)


# ----------------------------------------------------------------- field --
@pytest.mark.parametrize("field,name", [(0, "koalabear"), (1, "babybear")])
def test_poseidon2_kat(oracle, field, name):
    k = kats()[name]
    assert list(oracle.poseidon2(field, k["input"])) == k["output"]


@pytest.mark.parametrize("f", FIELDS)
def test_field_ops_vs_bigint(oracle, f):
    p = pyref.P[f]
    rng = np.random.default_rng(1)
    L = oracle.lib()
    for a, b in rng.integers(0, p, size=(200, 2)):
        a, b = int(a), int(b)
        assert L.or_add(f, a, b) == (a + b) % p
        assert L.or_sub(f, a, b) == (a - b) % p
        assert L.or_mul(f, a, b) == a * b % p
        if a:
            assert L.or_inv(f, a) * a % p == 1
        assert L.or_to_monty(f, a) == (a << 32) % p
        assert L.or_from_monty(f, (a << 32) % p) == a


@pytest.mark.parametrize("f", FIELDS)
def test_generators(oracle, f):
    p = pyref.P[f]
    g = oracle.generator(f)
    # primitive root: g^((p-1)/q) != 1 for every prime q | p - 1
    n, q, primes = p - 1, 2, set()
    while q * q <= n:
        while n % q == 0:
            primes.add(q)
            n //= q
        q += 1
    if n > 1:
        primes.add(n)
    assert all(pow(g, (p - 1) // q, p) != 1 for q in primes)
    adic = oracle.lib().or_two_adicity(f)
    assert (p - 1) % (1 << adic) == 0 and ((p - 1) >> adic) % 2 == 1
    for k in range(adic + 1):
        w = oracle.two_adic_generator(f, k)
        assert pow(w, 1 << k, p) == 1
        if k:
            assert pow(w, 1 << (k - 1), p) == p - 1  # order exactly 2^k
            assert w * w % p == oracle.two_adic_generator(f, k - 1)


def test_koalabear_generator_chain_head(oracle):
    # the p3 KoalaBear table starts 0x1, 0x7f000000, 0x7e010002, 0x6832fe4a and
    # ends 0x6ac49f88 = 3^127 (generator chain of two_adic_generator)
    assert [oracle.two_adic_generator(0, k) for k in range(4)] == [1, 0x7F000000, 0x7E010002, 0x6832FE4A]
    assert oracle.two_adic_generator(0, 24) == 0x6AC49F88 == pow(3, 127, pyref.P[0])
    assert oracle.two_adic_generator(1, 27) == 0x1A427A41 == pow(31, 15, pyref.P[1])


@pytest.mark.parametrize("f", FIELDS)
def test_ext4_mul_inv(oracle, f):
    p = pyref.P[f]
    rng = np.random.default_rng(2)
    for _ in range(50):
        a = [int(x) for x in rng.integers(0, p, 4)]
        b = [int(x) for x in rng.integers(0, p, 4)]
        assert list(oracle.ext_mul(f, a, b)) == pyref.ext_mul(f, a, b)
    a = [int(x) for x in rng.integers(1, p, 4)]
    assert pyref.ext_inv_by_search_check(f, a, [int(x) for x in oracle.ext_inv(f, a)])
    # x^4 = W
    x = [0, 1, 0, 0]
    x2 = pyref.ext_mul(f, x, x)
    assert pyref.ext_mul(f, x2, x2) == [pyref.W[f], 0, 0, 0]


# --------------------------------------------------------- sponge/compress --
@pytest.mark.parametrize("f", FIELDS)
def test_sponge_semantics(oracle, f):
    p = pyref.P[f]
    rng = np.random.default_rng(3)
    assert list(oracle.sponge(f, [])) == [0] * 8
    for n in [1, 7, 8, 9, 16, 17, 300]:
        v = [int(x) for x in rng.integers(0, p, n)]
        st = [0] * 16
        for off in range(0, n, 8):  # overwrite mode, not add
            chunk = v[off:off + 8]
            st[:len(chunk)] = chunk
            st = [int(x) for x in oracle.poseidon2(f, st)]
        assert list(oracle.sponge(f, v)) == st[:8]
    l, r = rng.integers(0, p, 8), rng.integers(0, p, 8)
    assert list(oracle.compress(f, l, r)) == list(oracle.poseidon2(f, np.concatenate([l, r]))[:8])


# --------------------------------------------------------------------- LDE --
@pytest.mark.parametrize("f", FIELDS)
@pytest.mark.parametrize("log_h,w,b", [(0, 2, 1), (1, 3, 1), (2, 1, 2), (3, 3, 3), (4, 2, 1), (5, 1, 2)])
def test_lde_vs_polynomial_definition(oracle, f, log_h, w, b):
    p = pyref.P[f]
    h = 1 << log_h
    m = oracle.random_matrix(f, 77 + log_h, h, w)
    for shift in [oracle.generator(f), 5, pow(oracle.two_adic_generator(f, 7), 3, p)]:
        for br in (True, False):
            got = oracle.coset_lde(f, m, b, shift, bit_reverse=br)
            naive = oracle.coset_lde(f, m, b, shift, bit_reverse=br, naive=True)
            assert np.array_equal(got, naive)
            for c in range(w):
                ref = pyref.coset_lde_column(f, [int(x) for x in m[:, c]], b, shift, br)
                assert [int(x) for x in got[:, c]] == ref


@pytest.mark.parametrize("f", FIELDS)
def test_lde_closed_forms(oracle, f):
    # shift = 1: natural row j * 2^b is the input row j (the subgroup <w_h> sits
    # inside <w_N>); a constant column stays constant; linearity.
    m = oracle.random_matrix(f, 5, 256, 3)
    out = oracle.coset_lde(f, m, 2, 1, bit_reverse=False)
    assert np.array_equal(out[::4], m)
    const = np.full((64, 2), 12345, np.uint32)
    assert np.all(oracle.coset_lde(f, const, 3, oracle.generator(f)) == 12345)
    a, b2 = oracle.random_matrix(f, 6, 64, 2), oracle.random_matrix(f, 7, 64, 2)
    p = pyref.P[f]
    s = ((a.astype(np.uint64) + b2) % p).astype(np.uint32)
    g = oracle.generator(f)
    la, lb, ls = (oracle.coset_lde(f, x, 1, g) for x in (a, b2, s))
    assert np.array_equal(((la.astype(np.uint64) + lb) % p).astype(np.uint32), ls)


def test_lde_planted_mistake_is_caught(oracle):
    # a wrong shift (off by one power of the root) must change the LDE
    f = 0
    m = oracle.random_matrix(f, 9, 16, 1)
    g = oracle.generator(f)
    good = oracle.coset_lde(f, m, 1, g)
    bad = oracle.coset_lde(f, m, 1, g * oracle.two_adic_generator(f, 5) % pyref.P[f])
    assert not np.array_equal(good, bad)


# -------------------------------------------------------------------- MMCS --
SHAPES = [[(8, 1), (8, 7), (4, 8), (2, 9)], [(16, 300)], [(1, 5)], [(4, 3), (8, 0), (8, 2), (1, 4)],
          [(8, 8), (4, 8), (4, 1), (2, 3), (1, 2)]]


@pytest.mark.parametrize("f", FIELDS)
@pytest.mark.parametrize("shape", SHAPES)
def test_mmcs_vs_definition(oracle, f, shape):
    mats = [oracle.random_matrix(f, 100 + i, h, w) for i, (h, w) in enumerate(shape)]
    tree = oracle.MerkleTree(f, mats)
    rows = [[[int(x) for x in r] for r in m] for m in mats]
    root, layers = pyref.merkle_root(f, rows, oracle.sponge, oracle.compress)
    assert [int(x) for x in tree.root] == [int(x) for x in root]
    heights = [h for h, _ in shape]
    widths = [w for _, w in shape]
    for idx in range(tree.max_height):
        opened, path = tree.open(idx)
        for k, m in enumerate(mats):
            assert np.array_equal(opened[k], m[idx >> (tree.log_max - (m.shape[0].bit_length() - 1))])
        assert pyref.verify_batch(f, heights, widths, tree.root, idx, opened, path, oracle.sponge,
                                  oracle.compress)
        assert oracle.mmcs_verify(f, heights, widths, tree.root, idx, opened, path)
        # planted mistakes: a flipped opened value or sibling must not verify
        bad = [o.copy() for o in opened]
        nz = [k for k, o in enumerate(bad) if o.size]
        bad[nz[0]][0] ^= 1
        assert not oracle.mmcs_verify(f, heights, widths, tree.root, idx, bad, path)
        if len(path):
            bp = path.copy()
            bp[0, 0] ^= 1
            assert not oracle.mmcs_verify(f, heights, widths, tree.root, idx, opened, bp)


SEEDED_SHAPES = [
    # (log_h, width, domain shift exponent of GEN*w^e or None for 1)
    [(5, 13, None), (5, 3, None), (4, 8, None), (2, 9, None), (4, 1, None)],
    [(6, 300, None)],
    [(3, 17, 1), (3, 5, 2), (1, 4, None)],
    [(0, 7, None), (0, 2, None)],
]


@pytest.mark.parametrize("f", FIELDS)
@pytest.mark.parametrize("shape", SEEDED_SHAPES)
@pytest.mark.parametrize("b", [1, 2, 3])
@pytest.mark.parametrize("block_cols", [1, 3, 8, 64])
def test_pcs_commit_seeded_vs_whole_ldes(oracle, f, shape, b, block_cols):
    """The memory-lean commit (LDE in column blocks, sponge states carried
    across blocks and matrices) equals the commitment of whole LDEs built
    from the same seeded traces."""
    p = oracle.PRIMES[f]
    g = oracle.generator(f)
    seeds = [(9 << 32) | (50 + i) for i in range(len(shape))]
    shifts = []
    for L, _, e in shape:
        shifts.append(1 if e is None else g * pow(oracle.two_adic_generator(f, L + 2), e, p) % p)
    ldes = []
    for s, (L, w, _), sh in zip(seeds, shape, shifts):
        a = oracle.random_matrix(f, s, 1 << L, w)
        ldes.append(oracle.coset_lde(f, a, b, g * pow(sh, p - 2, p) % p))
    want = oracle.MerkleTree(f, ldes).root
    got = oracle.pcs_commit_seeded(f, seeds, [1 << L for L, _, _ in shape], [w for _, w, _ in shape], b,
                                   domain_shifts=shifts, block_cols=block_cols)
    assert np.array_equal(got, want)
    # a planted mistake (a different seed for one matrix) changes the root
    bad = list(seeds)
    bad[-1] ^= 1
    assert not np.array_equal(
        oracle.pcs_commit_seeded(f, bad, [1 << L for L, _, _ in shape], [w for _, w, _ in shape], b,
                                 domain_shifts=shifts, block_cols=block_cols), want)


# --------------------------------------------------------------------- FRI --
@pytest.mark.parametrize("f", FIELDS)
def test_fri_fold_vs_line(oracle, f):
    p = pyref.P[f]
    n = 32
    v = oracle.random_matrix(f, 11, n, 4)
    beta = [int(x) for x in oracle.random_matrix(f, 12, 1, 4)[0]]
    out = oracle.fri_fold(f, v, beta)
    wn = pyref.root_of_unity(f, 5)
    for i in range(n // 2):
        x = pow(wn, pyref.rev(i, 4), p)
        exp = pyref.fold_pair(f, [int(t) for t in v[2 * i]], [int(t) for t in v[2 * i + 1]], beta, x)
        assert [int(t) for t in out[i]] == exp


@pytest.mark.parametrize("f", FIELDS)
def test_fri_fold_keeps_low_degree(oracle, f):
    # v = bit-reversed evaluations on <w_n> of an ext4 polynomial of degree < n/4
    # (each coordinate a base LDE with shift 1).  Folding halves the degree
    # bound; after log(n/4)+1 folds... every fold of a degree < d polynomial is
    # degree < d/2, so its iDFT's top half vanishes, and folding to length
    # 2^log_b leaves a constant when d = n / 2^log_b.
    n, lb = 256, 2
    coeff_cols = [oracle.random_matrix(f, 20 + k, n >> lb, 1) for k in range(4)]
    cols = [oracle.coset_lde(f, c, lb, 1, bit_reverse=True)[:, 0] for c in coeff_cols]
    v = np.stack(cols, axis=1)
    betas = oracle.random_matrix(f, 30, 8, 4)
    cur, r = v, 0
    while cur.shape[0] > (1 << lb):
        cur = oracle.fri_fold(f, cur, betas[r])
        r += 1
        # natural-order evaluations of the folded vector: degree < len / 2^lb
        nat = np.empty_like(cur)
        ln = cur.shape[0].bit_length() - 1
        for i in range(cur.shape[0]):
            nat[i] = cur[pyref.rev(i, ln)]
        for k in range(4):
            co = pyref.interpolate_coeffs(f, [int(x) for x in nat[:, k]])
            assert all(c == 0 for c in co[cur.shape[0] >> lb:])
    assert np.all(cur == cur[0])


@pytest.mark.parametrize("f", FIELDS)
def test_challenger_vs_duplex_rules(oracle, f):
    ch = oracle.Challenger(f)
    py = pyref.PyChallenger(f, oracle.poseidon2)
    rng = np.random.default_rng(4)
    p = pyref.P[f]
    for step in range(60):
        op = rng.integers(0, 4)
        if op == 0:
            v = int(rng.integers(0, p))
            ch.observe(v)
            py.observe(v)
        elif op == 1:
            assert ch.sample() == py.sample()
        elif op == 2:
            assert list(ch.sample_ext()) == py.sample_ext()
        else:
            assert ch.sample_bits(10) == py.sample_bits(10)
    # PoW check_witness = observe then sample_bits == 0
    c2 = ch.clone()
    ok = c2.check_witness(4, 17)
    py.observe(17)
    assert ok == (py.sample_bits(4) == 0)


@pytest.mark.parametrize("f", FIELDS)
def test_fri_commit_phase_vs_loop(oracle, f):
    log_n, log_stop = 9, 2
    v = oracle.random_matrix(f, 40, 1 << log_n, 4)
    extra = {6: oracle.random_matrix(f, 41, 1 << 6, 4)}
    ch = oracle.Challenger(f)
    ch.observe([1, 2, 3])
    py = ch.clone()
    roots, layers, final, betas = oracle.fri_commit_phase(ch, v, log_stop, inputs=extra)
    cur = v
    p = pyref.P[f]
    for r in range(log_n - log_stop):
        assert np.array_equal(layers[r], cur)
        leaves = cur.reshape(-1, 8)
        t = oracle.MerkleTree(f, [leaves])
        assert np.array_equal(t.root, roots[r])
        py.observe(t.root)
        beta = py.sample_ext()
        assert np.array_equal(beta, betas[r])
        cur = oracle.fri_fold(f, cur, beta)
        if cur.shape[0].bit_length() - 1 in extra:
            cur = ((cur.astype(np.uint64) + extra[cur.shape[0].bit_length() - 1]) % p).astype(np.uint32)
    assert np.array_equal(cur, final)


# ------------------------------------------------------------------ inputs --
@pytest.mark.parametrize("f", FIELDS)
def test_seeded_inputs(oracle, f):
    m = oracle.random_matrix(f, 0xC0FFEE, 4, 16).ravel()
    assert [int(x) for x in m] == [pyref.splitmix_element(f, 0xC0FFEE, i) for i in range(64)]
    big = oracle.random_matrix(f, 1, 1 << 12, 16)
    assert big.max() < pyref.P[f]


# ------------------------------------------------------------------ pcs open --
def _ext_pow(f, a, e):
    r = [1, 0, 0, 0]
    for _ in range(e):
        r = pyref.ext_mul(f, r, a)
    return r


@pytest.mark.parametrize("f", [0, 1])
def test_oracle_open_values_monomials(oracle, f):
    """The value at z of the column holding x^k on the trace domain is z^k.
    Evaluation is linear in the column, and the monomials k < h span every
    column, so this pins or_pcs_open_values completely; linearity is checked
    on a random combination."""
    p = pyref.P[f]
    rng = np.random.default_rng(5 + f)
    for lh, b in [(0, 1), (1, 2), (3, 1), (3, 3), (5, 2)]:
        h = 1 << lh
        w = pyref.root_of_unity(f, lh)
        trace = np.array([[pow(w, j * k, p) for k in range(h)] for j in range(h)], np.uint32)
        lde = oracle.coset_lde(f, trace, b, pyref.GEN[f])
        z = [int(x) for x in rng.integers(0, p, 4)]
        vals = oracle.pcs_open_values(f, lde, b, z)
        for k in range(h):
            assert [int(x) for x in vals[k]] == _ext_pow(f, z, k), (lh, b, k)
    # linearity: open(3 A + 5 B) = 3 open(A) + 5 open(B)
    a = oracle.random_matrix(f, 1, 16, 2)
    bm = oracle.random_matrix(f, 2, 16, 2)
    c = ((3 * a.astype(np.uint64) + 5 * bm.astype(np.uint64)) % p).astype(np.uint32)
    z = [7, 11, 13, 17]
    va, vb, vc = (oracle.pcs_open_values(f, oracle.coset_lde(f, m, 1, pyref.GEN[f]), 1, z) for m in (a, bm, c))
    assert np.array_equal(vc, (3 * va.astype(np.uint64) + 5 * vb.astype(np.uint64)) % p)


def _coset_coeffs(f, col, shift):
    """Coefficients (natural order) of the polynomial whose values on
    shift * <w_n> in bit-reversed order are col."""
    p = pyref.P[f]
    n = len(col)
    ln = n.bit_length() - 1
    nat = [0] * n
    for i in range(n):
        nat[pyref.rev(i, ln)] = int(col[i])
    g = pyref.interpolate_coeffs(f, nat)  # g(X) = poly(shift X)
    sinv = pow(shift, p - 2, p)
    return [g[k] * pow(sinv, k, p) % p for k in range(n)]


def _reduce_setup(oracle, f, corrupt=False):
    p = pyref.P[f]
    rng = np.random.default_rng(40 + f)
    shapes = [(4, 3, 2), (3, 2, 2), (4, 1, 1)]  # (log h, width, points) ; blowup 2 -> heights 32, 16, 32
    mats, npts, pts, vals = [], [], [], []
    zeta = [int(x) for x in rng.integers(0, p, 4)]
    g_zeta = [zeta[i] * pyref.root_of_unity(f, 4) % p for i in range(4)]  # a second, shifted point
    for i, (lh, w, k) in enumerate(shapes):
        lde = oracle.coset_lde(f, oracle.random_matrix(f, 60 + i, 1 << lh, w), 1, pyref.GEN[f])
        mats.append(lde)
        npts.append(k)
        for z in [zeta, g_zeta][:k]:
            pts.append(z)
            v = oracle.pcs_open_values(f, lde, 1, z)
            if corrupt and i == 0:
                v = v.copy()
                v[0, 0] = (int(v[0, 0]) + 1) % p
            vals.append(v)
    alpha = [int(x) for x in rng.integers(0, p, 4)]
    return mats, npts, np.array(pts, np.uint32), vals, alpha


@pytest.mark.parametrize("f", [0, 1])
def test_oracle_reduce_low_degree(oracle, f):
    """With correct opened values every (p(z) - p(x)) / (z - x) is a
    polynomial of degree < h - 1, so each height's reduced opening is a
    low-degree vector on GEN <w_N>; a corrupted value breaks that."""
    for corrupt in (False, True):
        mats, npts, pts, vals, alpha = _reduce_setup(oracle, f, corrupt)
        ro = oracle.pcs_reduce(f, mats, npts, pts, vals, alpha)
        assert sorted(ro) == [4, 5]
        for lh, h_trace in ((5, 16), (4, 8)):
            for c in range(4):
                coeffs = _coset_coeffs(f, ro[lh][:, c], pyref.GEN[f])
                top_zero = all(x == 0 for x in coeffs[h_trace - 1:])
                assert top_zero != (corrupt and lh == 5), (corrupt, lh, c)


@pytest.mark.parametrize("f", [0, 1])
def test_oracle_reduce_matches_verifier_loop(oracle, f):
    """At sampled rows, the verifier's loop (circuit/fri.rs:118-241: alpha
    powers counted per height across matrices, points and columns) gives
    the oracle's value; checked cross-multiplied, without inverses."""
    p = pyref.P[f]
    mats, npts, pts, vals, alpha = _reduce_setup(oracle, f)
    ro = oracle.pcs_reduce(f, mats, npts, pts, vals, alpha)
    for lh in (5, 4):
        n = 1 << lh
        wn = pyref.root_of_unity(f, lh)
        for i in (0, 1, 7, n - 1):
            x = pyref.GEN[f] * pow(wn, pyref.rev(i, lh), p) % p
            pow_ctr, acc_by_point, pi, vi = 0, {}, 0, 0
            for m, k in zip(mats, npts):
                for q in range(k):
                    z = tuple(int(t) for t in pts[pi])
                    v = vals[vi]
                    pi += 1
                    vi += 1
                    if m.shape[0] != n:
                        continue
                    acc = [0, 0, 0, 0]
                    for col in range(m.shape[1]):
                        ap = _ext_pow(f, alpha, pow_ctr)
                        diff = [(int(v[col, 0]) - int(m[i, col])) % p] + [int(t) for t in v[col, 1:]]
                        acc = [(a + b) % p for a, b in zip(acc, pyref.ext_mul(f, ap, diff))]
                        pow_ctr += 1
                    prev = acc_by_point.get(z, [0, 0, 0, 0])
                    acc_by_point[z] = [(a + b) % p for a, b in zip(prev, acc)]
            # ro * prod (z - x) == sum_z acc_z * prod_{z' != z} (z' - x)
            zs = list(acc_by_point)
            dens = {z: [(z[0] - x) % p, z[1], z[2], z[3]] for z in zs}
            lhs = [int(t) for t in ro[lh][i]]
            for z in zs:
                lhs = pyref.ext_mul(f, lhs, dens[z])
            rhs = [0, 0, 0, 0]
            for z in zs:
                term = acc_by_point[z]
                for z2 in zs:
                    if z2 != z:
                        term = pyref.ext_mul(f, term, dens[z2])
                rhs = [(a + b) % p for a, b in zip(rhs, term)]
            assert lhs == rhs, (lh, i)


@pytest.mark.parametrize("isa", [1, 2], ids=["avx2", "avx512"])
@pytest.mark.parametrize("f", [0, 1])
def test_cpu_baseline_matches_oracle(oracle, f, isa):
    """The CPU baseline bench.py times (oracle/cpu_baseline.c: Montgomery form,
    AVX2 and AVX-512, radix-4 passes, blocked layers) commits exactly what the
    checker commits: the whole LDE and the root, over heights 1..2^13
    (whole-matrix, paired and blocked DIF layers), widths that leave a
    partial sponge chunk and partial row groups of either lane count,
    blowups 1..8."""
    if isa == 2 and oracle.cpu_isa() < 2:
        pytest.skip("this host has no AVX-512")
    for (L, w, lb) in [(0, 1, 0), (0, 3, 1), (1, 5, 1), (2, 9, 3), (3, 8, 2), (9, 13, 1), (10, 16, 3),
                       (11, 24, 1), (12, 7, 2), (13, 9, 1), (4, 17, 0), (5, 33, 2)]:
        a = oracle.random_matrix(f, 100 + L, 1 << L, w)
        root, lde = oracle.cpu_pcs_commit(f, a, lb, want_lde=True, isa=isa)
        ref = oracle.coset_lde(f, a, lb, oracle.generator(f))
        assert np.array_equal(lde, ref), (L, w, lb)
        assert np.array_equal(root, oracle.MerkleTree(f, [ref]).root), (L, w, lb)


def _ext_scale(f, z, c):
    p = oracle_mod().PRIMES[f]
    return [int(x) * c % p for x in z]


def oracle_mod():
    from oracle import oracle as o
    return o


def _ext_from_coords(oracle, f, vals):
    """sum_e x^e v_e for the four coordinate values v_e (ext4 each)"""
    p = oracle.PRIMES[f]
    acc = [0, 0, 0, 0]
    for e, v in enumerate(vals):
        mono = [0, 0, 0, 0]
        mono[e] = 1
        acc = [(a + b) % p for a, b in zip(acc, oracle.ext_mul(f, mono, v))]
    return acc


@pytest.mark.parametrize("f", [0, 1])
@pytest.mark.parametrize("name", ["toy", "fibonacci", "cube", "gated"])
def test_quotient_oracle_meets_the_verifier_identity(oracle, f, name):
    """The oracle's compute_quotient_values, pinned by the verifier's
    out-of-domain check (vm/src/machine/verifier.rs:307-356): for a valid
    trace, the quotient interpolated from its values on Q and evaluated at a
    random ext4 zeta equals the constraints folded at zeta (from the trace
    polynomials at zeta and zeta w_n) times 1 / Z_H(zeta).  A trace with one
    wrong entry breaks the identity."""
    from tests import airs
    rng = np.random.default_rng(40 + f)
    log_n = 6
    n = 1 << log_n
    builder, qdb = getattr(airs, name + "_air")(f)
    nodes, asserts = builder.node_array(), builder.asserts
    if name == "gated":
        trace, pre = airs.gated_trace(f, n, rng)
        pv = []
    else:
        trace, pv = getattr(airs, name + "_trace")(f, n, rng)
        pre = None
    g = oracle.generator(f)
    p = oracle.PRIMES[f]
    wn = oracle.two_adic_generator(f, log_n)
    zeta = [int(x) for x in rng.integers(0, p, 4)]
    alpha = [int(x) for x in rng.integers(0, p, 4)]

    def check(trace):
        main_q = oracle.coset_lde(f, trace, qdb, g, bit_reverse=False)
        pre_q = oracle.coset_lde(f, pre, qdb, g, bit_reverse=False) if pre is not None else None
        quot = oracle.quotient_values(f, nodes, asserts, main_q, log_n, qdb, alpha, pv, pre_q)
        # q(zeta): the values on Q (natural order) as a committed LDE at blowup 1
        qn = quot.shape[0]
        rev = [int(format(i, f"0{log_n + qdb}b")[::-1], 2) if log_n + qdb else 0 for i in range(qn)]
        coords = oracle.pcs_open_values(f, quot[rev], 0, zeta)
        q_at = _ext_from_coords(oracle, f, coords)
        lde = oracle.coset_lde(f, trace, 1, g)
        ml, mn = oracle.pcs_open_values(f, lde, 1, zeta), oracle.pcs_open_values(f, lde, 1, _ext_scale(f, zeta, wn))
        pl = pn = None
        if pre is not None:
            plde = oracle.coset_lde(f, pre, 1, g)
            pl, pn = oracle.pcs_open_values(f, plde, 1, zeta), oracle.pcs_open_values(f, plde, 1,
                                                                                     _ext_scale(f, zeta, wn))
        folded, inv_zh = oracle.constraints_at_point(f, nodes, asserts, ml, mn, zeta, log_n, alpha, pv, pl, pn)
        return [int(x) for x in oracle.ext_mul(f, folded, inv_zh)] == q_at

    assert check(trace)
    bad = trace.copy()
    bad[n // 2, 0] = (int(bad[n // 2, 0]) + 1) % p
    assert not check(bad)


@pytest.mark.parametrize("f", [0, 1])
def test_quotient_oracle_lookup_meets_the_verifier_identity(oracle, f):
    """The permutation (LogUp) constraints of a valid two-interaction batch
    (tests/airs.py lookup_air): the oracle's quotient on Q, interpolated and
    evaluated at a random zeta, equals the constraints folded at zeta from
    the main and permutation traces' values at zeta and zeta w_n times
    1 / Z_H(zeta) (vm/src/machine/verifier.rs:307-356, the permutation
    columns recombined from their base coordinates).  A wrong running-sum
    entry breaks it, and so does a wrong regional sum."""
    from tests import airs
    rng = np.random.default_rng(60 + f)
    log_n = 5
    n = 1 << log_n
    builder, qdb = airs.lookup_air(f)
    nodes, asserts = builder.node_array(), builder.asserts
    main, perm, chal, rsum, gsum = airs.lookup_trace(f, n, rng, oracle)
    g = oracle.generator(f)
    p = oracle.PRIMES[f]
    wn = oracle.two_adic_generator(f, log_n)
    zeta = [int(x) for x in rng.integers(0, p, 4)]
    alpha = [int(x) for x in rng.integers(0, p, 4)]

    def check(perm, rsum):
        main_q = oracle.coset_lde(f, main, qdb, g, bit_reverse=False)
        perm_q = oracle.coset_lde(f, perm, qdb, g, bit_reverse=False)
        quot = oracle.quotient_values(f, nodes, asserts, main_q, log_n, qdb, alpha, (), None, perm_q, chal, rsum,
                                      gsum)
        qn = quot.shape[0]
        rev = [int(format(i, f"0{log_n + qdb}b")[::-1], 2) for i in range(qn)]
        q_at = _ext_from_coords(oracle, f, oracle.pcs_open_values(f, quot[rev], 0, zeta))
        zn = _ext_scale(f, zeta, wn)
        lde, plde = oracle.coset_lde(f, main, 1, g), oracle.coset_lde(f, perm, 1, g)
        folded, inv_zh = oracle.constraints_at_point(
            f, nodes, asserts, oracle.pcs_open_values(f, lde, 1, zeta), oracle.pcs_open_values(f, lde, 1, zn), zeta,
            log_n, alpha, (), None, None, oracle.pcs_open_values(f, plde, 1, zeta),
            oracle.pcs_open_values(f, plde, 1, zn), chal, rsum, gsum)
        return [int(x) for x in oracle.ext_mul(f, folded, inv_zh)] == q_at

    assert check(perm, rsum)
    bad = perm.copy()
    bad[n // 3, 5] = (int(bad[n // 3, 5]) + 1) % p  # a running-sum coordinate
    assert not check(bad, rsum)
    assert not check(perm, [(rsum[0] + 1) % p] + list(rsum[1:]))


def test_air_compile_reuses_registers():
    """pg_air_compile: a chain of 200 asserts over shared subexpressions
    needs few registers (each freed after its last reader; leaves are read
    in place and take none), keeps every assert, stages each column window
    once when all fit, and rejects malformed DAGs."""
    import pico_amd
    from pico_amd.air import AirBuilder
    b = AirBuilder(0)
    acc = b.main(0)
    for k in range(200):
        acc = acc * b.main(k % 2) + k
        b.assert_zero(acc - b.main(1 - k % 2, next=True))
    prog = b.compile()
    assert prog.nregs <= 3
    # one instruction per inner node and per assert, plus the staging of the
    # two column windows read (main columns 0 and 1 at the local row and at
    # the next row), which the window slots hold for the whole program
    kinds = [kind for kind, _, _ in b.nodes]
    assert prog.ninstr == len(b.asserts) + sum(1 for k in kinds if k >= 7) + 2
    bad = np.array([[7, 1, 0]], np.uint32)  # add of nodes not yet defined
    with pytest.raises(pico_amd.PgError):
        pico_amd.air.AirProgram(0, bad, np.zeros(1, np.uint32))
    with pytest.raises(pico_amd.PgError):
        pico_amd.air.AirProgram(0, np.array([[0, 2, 0]], np.uint32), np.zeros(1, np.uint32))  # offset 2
    with pytest.raises(pico_amd.PgError):
        pico_amd.air.AirProgram(0, np.array([[0, 0, 0]], np.uint32), np.array([3], np.uint32))  # assert range
    with pytest.raises(pico_amd.PgError):  # permutation row offset 2
        pico_amd.air.AirProgram(0, np.array([[11, 2, 0]], np.uint32), np.zeros(1, np.uint32))
    with pytest.raises(pico_amd.PgError):  # global-sum coordinate beyond 13
        pico_amd.air.AirProgram(0, np.array([[14, 0, 14]], np.uint32), np.zeros(1, np.uint32))
    with pytest.raises(pico_amd.PgError):  # unknown kind
        pico_amd.air.AirProgram(0, np.array([[15, 0, 0]], np.uint32), np.zeros(1, np.uint32))


def test_air_compile_lowers_extension_arithmetic():
    """pg_air_compile: an extension-valued assert becomes four base asserts
    (one per coordinate), a base-by-extension product four products, an
    extension product 16 products and the W terms; a base assert stays one.
    Counted through the program length of single-assert AIRs whose leaves all
    sit in one column window."""
    import pico_amd
    from pico_amd.air import AirBuilder

    def ninstr(make):
        b = AirBuilder(0)
        b.assert_zero(make(b))
        return b.compile().ninstr

    stages = ninstr(lambda b: b.main(0) * b.main(1)) - 2  # one mul + one assert
    assert stages == 1
    # ext4 column: four coordinate asserts (+ the window of base columns 0..3)
    assert ninstr(lambda b: b.permutation(0)) == 1 + 4
    # base x ext: four products and four asserts (+ the main column's window)
    assert ninstr(lambda b: b.main(0) * b.perm_challenge(0)) == 1 + 4 + 4
    # ext x ext: 16 products, 12 sums (3 per coordinate: c_0 = a0b0 +
    # W(a1b3 + a2b2 + a3b1), ..., c_3 = a0b3 + a1b2 + a2b1 + a3b0), 3 products
    # by W, 4 asserts
    assert ninstr(lambda b: b.permutation(0) * b.perm_challenge(1)) == 1 + 16 + 12 + 3 + 4
    # a base constant added to an ext value changes coordinate 0 only
    assert ninstr(lambda b: b.regional_sum() + 5) == 1 + 4
