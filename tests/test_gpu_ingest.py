"""Trace ingest on the device: strided views (pg_buf_view) and the quotient
chunk split that Pcs::commit receives in the reference
(vm/src/machine/prover.rs:513-523: `RowMajorMatrix::new_col(values)
.flatten_to_base()`, `domain.split_evals(d, ..)`, `domain.split_domains(d)`).

The split is checked two ways: the commitment of the d views equals the
oracle's commitment of the host-split chunks, and the chunk polynomials
opened at an out-of-domain point recombine to the original quotient there
with the verifier's selector formula (each zp_i is 1 on chunk domain i and 0
on the others), which fixes the chunk order and the chunk shifts without the
oracle."""
import numpy as np
import pytest

import pico_amd
from pico_amd import BABYBEAR, KOALABEAR
from pico_amd._lib import PgError, lib
from tests import pyref

pytestmark = pytest.mark.gpu
FIELDS = [KOALABEAR, BABYBEAR]


def test_views_download_and_bounds(ctx, oracle):
    f = KOALABEAR
    host = oracle.random_matrix(f, 31, 64, 12)
    m = ctx.upload(host, field=f)
    v = m.view(row0=3, row_step=5, col0=2, cols=7)
    assert (v.rows, v.cols, v.ld) == (13, 7, 60)
    assert np.array_equal(v.to_host(f), host[3::5, 2:9])
    assert np.array_equal(v.to_host(f, row0=4, nrows=6), host[3::5, 2:9][4:10])
    raw = m.to_host()
    assert np.array_equal(v.to_host(), raw[3::5, 2:9])
    vv = v.view(row0=1, row_step=2, col0=3, cols=2)  # a view of a view
    assert np.array_equal(vv.to_host(f), host[3::5, 2:9][1::2, 3:5])
    one = m.view(row0=63, row_step=7, rows=1)
    assert np.array_equal(one.to_host(f), host[63:64])
    empty = m.view(row0=0, rows=0, cols=3)
    assert empty.to_host(f).shape == (0, 3)
    for bad in [dict(row0=64, rows=1), dict(row0=3, row_step=5, rows=14), dict(col0=10, cols=3),
                dict(row_step=0, rows=2), dict(cols=0)]:
        with pytest.raises(PgError):
            m.view(**bad)
    # kernels that read rows * cols contiguous words turn strided views away
    with pytest.raises(PgError):
        pico_amd.MerkleTreeMmcs(ctx, f).commit([v])
    assert lib().pg_fill_random(ctx.handle, v.handle, f, 5) != 0
    assert lib().pg_upload(ctx.handle, v.handle, pico_amd.core._ptr(np.zeros((13, 7), np.uint32)), -1, 0) != 0
    # a view with unit step over whole rows is dense and allowed everywhere
    d = m.view(row0=8, rows=16)
    assert d.ld == d.cols == 12
    assert np.array_equal(d.to_host(f), host[8:24])
    for x in (vv, v, one, empty, d):
        x.free()
    m.free()


@pytest.mark.parametrize("f", FIELDS)
@pytest.mark.parametrize("log_n,d,lb", [(9, 2, 1), (8, 4, 3), (3, 8, 1), (12, 2, 2)])
def test_quotient_split_commit_matches_oracle(ctx, oracle, f, log_n, d, lb):
    p = pico_amd.PRIMES[f]
    g = pico_amd.GENERATORS[f]
    nd = (1 << log_n) * d
    flat = oracle.random_matrix(f, 4000 + log_n * 10 + d, nd, 4)  # (n d) x 4, ext4 rows flattened
    dev = ctx.upload(flat, field=f)
    qdom = pico_amd.TwoAdicMultiplicativeCoset(nd.bit_length() - 1, g, f)
    doms = qdom.split_domains(d)
    chunks = qdom.split_evals(d, dev)
    assert all(c.rows == 1 << log_n and c.ld == 4 * d for c in chunks)
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
    root, pd = pcs.commit(list(zip(doms, chunks)))
    w = pico_amd.two_adic_generator(f, nd.bit_length() - 1)
    for i, dm in enumerate(doms):
        assert dm.log_n == log_n and dm.shift == g * pow(w, i, p) % p
    ldes = [oracle.coset_lde(f, flat[i::d], lb, g * pow(dm.shift, p - 2, p) % p) for i, dm in enumerate(doms)]
    assert np.array_equal(root, oracle.MerkleTree(f, ldes).root)
    # same root from chunks split on the host and uploaded one by one
    dense = [ctx.upload(np.ascontiguousarray(flat[i::d]), field=f) for i in range(d)]
    root2, pd2 = pcs.commit(list(zip(doms, dense)))
    assert np.array_equal(root, root2)
    for k in range(d):
        gdom = pico_amd.TwoAdicMultiplicativeCoset(doms[k].log_n, g)
        assert np.array_equal(pcs.get_evaluations_on_domain(pd, k, gdom), pcs.get_evaluations_on_domain(pd2, k, gdom))
    # as in p3 TwoAdicFriPcs, a domain whose shift is not the generator is refused
    if d > 1:
        with pytest.raises(PgError):
            pcs.get_evaluations_on_domain(pd, 1, doms[1])
    pd.free()
    pd2.free()
    dev.free()


def _ext_pow(f, a, e):
    r = [1, 0, 0, 0]
    while e:
        if e & 1:
            r = pyref.ext_mul(f, r, a)
        a = pyref.ext_mul(f, a, a)
        e >>= 1
    return r


@pytest.mark.parametrize("f", FIELDS)
def test_quotient_chunks_recombine_at_zeta(ctx, f):
    """q(X) of degree < n d on D = g<w_{nd}>; chunk i interpolates q on
    D_i = g w^i <w_n>.  sum_i zp_i(zeta) q_i(zeta) = q(zeta) with
    zp_i = prod_{j != i} Z_{D_j}(zeta) / Z_{D_j}(g w^i), Z_{D_j}(X) = X^n - s_j^n."""
    p = pico_amd.PRIMES[f]
    g = pico_amd.GENERATORS[f]
    log_n, d = 7, 4
    n, nd = 1 << log_n, (1 << log_n) * d
    rng = np.random.default_rng(21 + f)
    coeffs = rng.integers(0, p, (nd, 4)).astype(np.uint64)  # 4 base polynomials of degree < n d
    w = pico_amd.two_adic_generator(f, nd.bit_length() - 1)
    xs = np.array([g * pow(w, j, p) % p for j in range(nd)], np.uint64)
    vals = np.zeros((nd, 4), np.uint64)
    pw = np.ones(nd, np.uint64)
    for k in range(nd):
        vals = (vals + coeffs[k][None, :] * pw[:, None]) % p
        pw = pw * xs % p
    dev = ctx.upload(vals.astype(np.uint32), field=f)
    qdom = pico_amd.TwoAdicMultiplicativeCoset(nd.bit_length() - 1, g, f)
    doms = qdom.split_domains(d)
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
    _, pd = pcs.commit(list(zip(doms, qdom.split_evals(d, dev))))
    zeta = [int(t) for t in rng.integers(0, p, 4)]
    opened = [pcs.open_values(pd, i, np.array([zeta], np.uint32))[0] for i in range(d)]  # d x 4 cols x ext4
    zn = _ext_pow(f, zeta, n)
    sn = [pow(dm.shift, n, p) for dm in doms]
    want = []
    for c in range(4):
        acc = [0, 0, 0, 0]
        for k in range(nd - 1, -1, -1):  # Horner: q_c(zeta)
            acc = pyref.ext_mul(f, acc, zeta)
            acc[0] = (acc[0] + int(coeffs[k, c])) % p
        want.append(acc)
    for c in range(4):
        got = [0, 0, 0, 0]
        for i in range(d):
            zp = [1, 0, 0, 0]
            for j in range(d):
                if j != i:
                    num = [(zn[0] - sn[j]) % p] + zn[1:]
                    den = pow((sn[i] - sn[j]) % p, p - 2, p)
                    zp = pyref.ext_mul(f, zp, [t * den % p for t in num])
            term = pyref.ext_mul(f, zp, [int(t) for t in opened[i][c]])
            got = [(a + b) % p for a, b in zip(got, term)]
        assert got == want[c], c
    pd.free()
    dev.free()


def _to_monty(f, a):
    p = pico_amd.PRIMES[f]
    return ((a.astype(np.uint64) << np.uint64(32)) % np.uint64(p)).astype(np.uint32)


@pytest.mark.parametrize("f", FIELDS)
@pytest.mark.parametrize("lb", [1, 2])
def test_commit_host_matches_upload_then_commit(ctx, oracle, f, lb):
    """Mixed heights in an order that is not sorted (the tree sorts, uploads
    follow the tree's order), widths that are and are not multiples of 8."""
    p = pico_amd.PRIMES[f]
    shapes = [(9, 16), (6, 5), (9, 3), (12, 8), (3, 40), (12, 1), (0, 7)]
    hosts = [oracle.random_matrix(f, 5000 + i, 1 << L, w) for i, (L, w) in enumerate(shapes)]
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
    doms = [pcs.natural_domain_for_degree(1 << L) for L, _ in shapes]
    # one matrix on a shifted (quotient-like) domain
    doms[1] = pico_amd.TwoAdicMultiplicativeCoset(6, pico_amd.GENERATORS[f] * pico_amd.two_adic_generator(f, 8) % p, f)
    root_h, pd_h = pcs.commit_host(list(zip(doms, hosts)))
    root_d, pd_d = pcs.commit([(d, ctx.upload(a, field=f)) for d, a in zip(doms, hosts)])
    ldes = [oracle.coset_lde(f, a, lb, pico_amd.GENERATORS[f] * pow(d.shift, p - 2, p) % p) for d, a in zip(doms, hosts)]
    assert np.array_equal(root_h, oracle.MerkleTree(f, ldes).root)
    assert np.array_equal(root_h, root_d)
    for k, d in enumerate(doms):
        gdom = pico_amd.TwoAdicMultiplicativeCoset(d.log_n, pico_amd.GENERATORS[f])
        assert np.array_equal(pcs.get_evaluations_on_domain(pd_h, k, gdom),
                              pcs.get_evaluations_on_domain(pd_d, k, gdom)), k
    # Montgomery words on the host, from page-locked memory
    pinned = []
    for a in hosts:
        pm = ctx.host_matrix(*a.shape)
        pm.array[:] = _to_monty(f, a)
        pinned.append(pm)
    root_m, pd_m = pcs.commit_host([(d, pm.array) for d, pm in zip(doms, pinned)], canonical=False)
    assert np.array_equal(root_m, root_h)
    for x in (pd_h, pd_d, pd_m):
        x.free()
    for pm in pinned:
        pm.free()


def test_commit_host_rejects_bad_shapes_without_leaking(ctx):
    f = KOALABEAR
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
    base = ctx.bytes_in_use
    good = np.ones((64, 4), np.uint32)
    for bad in [np.ones((48, 4), np.uint32), np.ones((64, 0), np.uint32)]:
        with pytest.raises(PgError):
            pcs.commit_host([(pcs.natural_domain_for_degree(64), good), (pcs.natural_domain_for_degree(64), bad)])
    too_big = pico_amd.TwoAdicFriPcs(ctx, f, 30)
    with pytest.raises(PgError):
        too_big.commit_host([(pcs.natural_domain_for_degree(64), good)])
    assert ctx.bytes_in_use == base


def test_commit_host_column_blocks_match_oracle(ctx, oracle):
    """Tallest matrices cut into column blocks (uploaded through the stage,
    extended and absorbed block by block), with the block size forced down
    to 64 MiB / 8 columns so small matrices split: widths 300 (blocks 128,
    128, 44) and 200 (128, 72, starting mid-chunk at input 300 = 37 * 8 + 4),
    an injected shorter matrix; Montgomery words in page-locked and in
    pageable memory."""
    f = KOALABEAR
    shapes = [(17, 300), (15, 9), (17, 200)]
    hosts = [oracle.random_matrix(f, 6000 + i, 1 << L, w) for i, (L, w) in enumerate(shapes)]
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
    doms = [pcs.natural_domain_for_degree(1 << L) for L, _ in shapes]
    pinned = []
    for a in hosts:
        pm = ctx.host_matrix(*a.shape)
        pm.array[:] = _to_monty(f, a)
        pinned.append(pm)
    ctx.debug_host_blocks(64 << 20, 8)
    try:
        root_h, pd_h = pcs.commit_host([(d, pm.array) for d, pm in zip(doms, pinned)], canonical=False)
        root_p, pd_p = pcs.commit_host([(d, pm.array.copy()) for d, pm in zip(doms, pinned)], canonical=False)
    finally:
        ctx.debug_host_blocks()
    root_d, pd_d = pcs.commit([(d, ctx.upload(a, field=f)) for d, a in zip(doms, hosts)])
    assert np.array_equal(root_h, root_d)
    assert np.array_equal(root_p, root_d)
    ldes = [oracle.coset_lde(f, a, 1, pico_amd.GENERATORS[f]) for a in hosts]
    assert np.array_equal(root_h, oracle.MerkleTree(f, ldes).root)
    for pd in (pd_h, pd_p):
        mats = pico_amd.MerkleTreeMmcs(ctx, f).get_matrices(pd)
        for k in (0, 2):
            assert np.array_equal(mats[k].to_host(f, row0=1000, nrows=64), ldes[k][1000:1064]), k
    for x in (pd_h, pd_d, pd_p):
        x.free()
    for pm in pinned:
        pm.free()


def test_commit_host_default_blocks(ctx):
    """Default block sizing (256 MiB, at least 64 columns): a 2^20 x 72
    matrix is cut into blocks of 64 and 8 columns; an injected shorter
    matrix rides along.  Roots equal upload-then-commit (itself pinned to the
    oracle by the parity tests), from page-locked and pageable memory."""
    f = KOALABEAR
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
    shapes = [(20, 72), (18, 20)]
    devs = [ctx.random_matrix(f, 6100 + i, 1 << L, w) for i, (L, w) in enumerate(shapes)]
    doms = [pcs.natural_domain_for_degree(1 << L) for L, _ in shapes]
    root_d, pd_d = pcs.commit(list(zip(doms, devs)))
    pinned = [ctx.host_matrix(1 << L, w) for L, w in shapes]
    for pm, d in zip(pinned, devs):
        pm.array[:] = d.to_host()
    root_h, pd_h = pcs.commit_host([(d, pm.array) for d, pm in zip(doms, pinned)], canonical=False)
    root_p, pd_p = pcs.commit_host([(d, pm.array.copy()) for d, pm in zip(doms, pinned)], canonical=False)
    assert np.array_equal(root_h, root_d)
    assert np.array_equal(root_p, root_d)
    for x in (pd_h, pd_d, pd_p):
        x.free()
    for pm in pinned:
        pm.free()


@pytest.mark.parametrize("f", FIELDS)
def test_commit_host_split_tail_canonical(ctx, oracle, f):
    """Several matrices of the tallest height: the last one's columns are
    uploaded, converted from canonical form and absorbed in narrower pieces
    (24 columns -> 8 + 16); roots equal the oracle's, pageable and
    page-locked."""
    shapes = [(12, 40), (10, 5), (12, 24)]
    hosts = [oracle.random_matrix(f, 6200 + i, 1 << L, w) for i, (L, w) in enumerate(shapes)]
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
    doms = [pcs.natural_domain_for_degree(1 << L) for L, _ in shapes]
    ldes = [oracle.coset_lde(f, a, 1, pico_amd.GENERATORS[f]) for a in hosts]
    want = oracle.MerkleTree(f, ldes).root
    root, pd = pcs.commit_host(list(zip(doms, hosts)))
    assert np.array_equal(root, want)
    pd.free()
    pinned = []
    for a in hosts:
        pm = ctx.host_matrix(*a.shape)
        pm.array[:] = a
        pinned.append(pm)
    root, pd = pcs.commit_host([(d, pm.array) for d, pm in zip(doms, pinned)])
    assert np.array_equal(root, want)
    pd.free()
    for pm in pinned:
        pm.free()


@pytest.mark.parametrize("f", FIELDS)
def test_commit_host_single_tall_matrix_with_shorter_runs(ctx, oracle, f):
    """One matrix of the tallest height (hashed by the tree's own leaf pass)
    and two shorter runs, one of two matrices: the shorter runs are uploaded,
    extended and row-hashed first and their layers only compress."""
    shapes = [(9, 9), (11, 20), (9, 7), (4, 3)]
    hosts = [oracle.random_matrix(f, 6300 + i, 1 << L, w) for i, (L, w) in enumerate(shapes)]
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, 2)
    doms = [pcs.natural_domain_for_degree(1 << L) for L, _ in shapes]
    ldes = [oracle.coset_lde(f, a, 2, pico_amd.GENERATORS[f]) for a in hosts]
    root, pd = pcs.commit_host(list(zip(doms, hosts)))
    assert np.array_equal(root, oracle.MerkleTree(f, ldes).root)
    mats = pico_amd.MerkleTreeMmcs(ctx, f).get_matrices(pd)
    for k in range(len(shapes)):
        assert np.array_equal(mats[k].to_host(f), ldes[k]), k
    pd.free()


def test_commit_host_concurrent_threads_share_the_staging_pool(ctx):
    """Two threads commit pageable host matrices at the same time (their
    staged copies share the process's host copy threads) and a third
    downloads through the stage meanwhile: every root equals its solo
    commit, the download equals its solo download."""
    import threading

    f = KOALABEAR
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
    jobs = [[(16, 72), (14, 9)], [(17, 40), (15, 16)]]
    hosts = []
    for j, shapes in enumerate(jobs):
        mats = [ctx.random_matrix(f, 6400 + 10 * j + i, 1 << L, w) for i, (L, w) in enumerate(shapes)]
        hosts.append([m.to_host() for m in mats])
        for m in mats:
            m.free()
    doms = [[pcs.natural_domain_for_degree(1 << L) for L, _ in shapes] for shapes in jobs]
    solo = []
    for d, h in zip(doms, hosts):
        r, pd = pcs.commit_host(list(zip(d, h)), canonical=False)
        solo.append(r)
        pd.free()
    big = ctx.random_matrix(f, 6500, 1 << 18, 64)
    big_solo = big.to_host()
    got, errors = [None] * 3, []

    def run(j):
        try:
            for _ in range(3):
                if j < 2:
                    r, pd = pcs.commit_host(list(zip(doms[j], hosts[j])), canonical=False)
                    pd.free()
                    if not np.array_equal(r, solo[j]):
                        got[j] = r
                        return
                    got[j] = r
                else:
                    got[j] = big.to_host()
                    if not np.array_equal(got[j], big_solo):
                        return
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=run, args=(j,)) for j in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert np.array_equal(got[0], solo[0]) and np.array_equal(got[1], solo[1])
    assert np.array_equal(got[2], big_solo)
    big.free()


def test_commit_host_default_blocks_split_tail_matches_oracle(ctx, oracle):
    """Default block sizing on a matrix whose last column block is split: at
    2^19 rows a 256 MiB block is 128 columns, so 200 columns go as 128 + 72,
    and the 72-column tail is uploaded and absorbed in narrower pieces.  The
    root and the whole LDE of the split matrix equal the oracle's; a shorter
    matrix is injected beside it."""
    f = KOALABEAR
    shapes = [(19, 200), (17, 12)]
    hosts = [oracle.random_matrix(f, 6400 + i, 1 << L, w) for i, (L, w) in enumerate(shapes)]
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
    doms = [pcs.natural_domain_for_degree(1 << L) for L, _ in shapes]
    print("default-block commit_host 2^19 x 200 + 2^17 x 12", flush=True)
    root, pd = pcs.commit_host([(d, pico_amd.to_monty(f, a)) for d, a in zip(doms, hosts)], canonical=False)
    try:
        ldes = [oracle.coset_lde(f, a, 1, pico_amd.GENERATORS[f]) for a in hosts]
        assert np.array_equal(root, oracle.MerkleTree(f, ldes).root)
        got = pico_amd.MerkleTreeMmcs(ctx, f).get_matrices(pd)
        assert np.array_equal(got[0].to_host(f), ldes[0])
        assert np.array_equal(got[1].to_host(f), ldes[1])
    finally:
        pd.free()
