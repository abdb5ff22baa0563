"""Quotient evaluation on the device (pg_air_compile + pg_quotient_values)
against the oracle's compute_quotient_values restatement
(vm/src/machine/utils.rs:201-356), and the whole device-resident chain the
reference runs per chip (vm/src/machine/prover.rs:441-540): main trace
committed, quotient evaluated from the committed LDE (no readback), split
into chunks and committed, everything opened at zeta and checked with the
verifier's equation (vm/src/machine/verifier.rs:307-356)."""
import numpy as np
import pytest

import pico_amd
from pico_amd import BABYBEAR, KOALABEAR, air
from tests import airs

pytestmark = pytest.mark.gpu
FIELDS = [KOALABEAR, BABYBEAR]


def _trace(name, f, n, rng):
    if name == "gated":
        main, pre = airs.gated_trace(f, n, rng)
        return main, pre, []
    main, pv = getattr(airs, name + "_trace")(f, n, rng)
    return main, None, pv


def _commit(ctx, f, lb, mat):
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
    m = ctx.upload(mat, field=f)
    root, pd = pcs.commit([(pcs.natural_domain_for_degree(mat.shape[0]), m)])
    m.free()
    return root, pd


@pytest.mark.parametrize("f", FIELDS)
@pytest.mark.parametrize("name", ["toy", "fibonacci", "cube", "gated"])
def test_quotient_matches_oracle(ctx, oracle, f, name):
    """Every AIR, trace heights 2^0..2^12, blowup 2 and 4, valid and random
    traces (random ones give non-polynomial quotients: every selector and
    constraint term shows)."""
    rng = np.random.default_rng(7 + f)
    p = pico_amd.PRIMES[f]
    builder, qdb = getattr(airs, name + "_air")(f)
    prog = builder.compile()
    nodes, asserts = builder.node_array(), builder.asserts
    g = pico_amd.GENERATORS[f]
    for log_n in (0, 1, 3, 8, 12):
        for lb in (max(1, qdb), 2):
            for valid in (True, False):
                n = 1 << log_n
                main, pre, pv = _trace(name, f, n, rng)
                if not valid:
                    main = rng.integers(0, p, main.shape).astype(np.uint32)
                    pv = [int(x) for x in rng.integers(0, p, len(pv))]
                print(f"quotient {name} field {f} 2^{log_n} b={lb} valid={valid}", flush=True)
                alpha = [int(x) for x in rng.integers(0, p, 4)]
                _, md = _commit(ctx, f, lb, main)
                pd = _commit(ctx, f, lb, pre)[1] if pre is not None else None
                q = air.quotient_values(ctx, prog, md, 0, lb, qdb, alpha, pv, pd, 0)
                got = q.to_host(f)
                q.free()
                md.free()
                if pd is not None:
                    pd.free()
                main_q = oracle.coset_lde(f, main, qdb, g, bit_reverse=False)
                pre_q = oracle.coset_lde(f, pre, qdb, g, bit_reverse=False) if pre is not None else None
                want = oracle.quotient_values(f, nodes, asserts, main_q, log_n, qdb, alpha, pv, pre_q)
                assert np.array_equal(got, want), (name, log_n, lb, valid)


@pytest.mark.parametrize("f", FIELDS)
def test_quotient_wide_random_air_matches_oracle(ctx, oracle, f):
    """Random constraints over 70 main and 21 preprocessed columns: the
    program stages many column windows (more than its slots hold, so
    windows are replaced and staged again), windows that run past the trace
    width, local and next-row windows of both traces, constants and public
    values.  Heights 2^0..2^9 (row blocks smaller than a workgroup up to
    several blocks), quotient degree 2^0..2^2."""
    rng = np.random.default_rng(31 + f)
    p = pico_amd.PRIMES[f]
    g = pico_amd.GENERATORS[f]
    builder, _ = airs.random_air(f, rng, 70, 21, 120)
    prog = builder.compile()
    nodes, asserts = builder.node_array(), builder.asserts
    for log_n in (0, 2, 5, 6, 9):
        for qdb, lb in ((0, 1), (1, 1), (2, 2)):
            n = 1 << log_n
            print(f"wide quotient field {f} 2^{log_n} qdb={qdb} b={lb}", flush=True)
            main = rng.integers(0, p, (n, 70)).astype(np.uint32)
            pre = rng.integers(0, p, (n, 21)).astype(np.uint32)
            pv = [int(x) for x in rng.integers(0, p, 3)]
            alpha = [int(x) for x in rng.integers(0, p, 4)]
            _, md = _commit(ctx, f, lb, main)
            _, pd = _commit(ctx, f, lb, pre)
            q = air.quotient_values(ctx, prog, md, 0, lb, qdb, alpha, pv, pd, 0)
            got = q.to_host(f)
            q.free()
            md.free()
            pd.free()
            want = oracle.quotient_values(f, nodes, asserts, oracle.coset_lde(f, main, qdb, g, bit_reverse=False),
                                          log_n, qdb, alpha, pv,
                                          oracle.coset_lde(f, pre, qdb, g, bit_reverse=False))
            assert np.array_equal(got, want), (log_n, qdb, lb)


@pytest.mark.parametrize("f", FIELDS)
def test_quotient_lookup_matches_oracle(ctx, oracle, f):
    """The permutation (LogUp) constraints of a two-interaction batch
    (tests/airs.py lookup_air: extension-valued columns, challenges, the
    regional sum, the global-sum check) over the committed main and
    permutation traces, heights 2^0..2^10, blowup 2 and 4, valid and random
    traces, against the oracle's direct extension arithmetic."""
    rng = np.random.default_rng(71 + f)
    p = pico_amd.PRIMES[f]
    g = pico_amd.GENERATORS[f]
    builder, qdb = airs.lookup_air(f)
    prog = builder.compile()
    nodes, asserts = builder.node_array(), builder.asserts
    for log_n in (0, 1, 4, 7, 10):
        for lb in (1, 2):
            for valid in (True, False):
                n = 1 << log_n
                print(f"lookup quotient field {f} 2^{log_n} b={lb} valid={valid}", flush=True)
                main, perm, chal, rsum, gsum = airs.lookup_trace(f, n, rng, oracle)
                if not valid:
                    perm = rng.integers(0, p, perm.shape).astype(np.uint32)
                    rsum = [int(x) for x in rng.integers(0, p, 4)]
                alpha = [int(x) for x in rng.integers(0, p, 4)]
                _, md = _commit(ctx, f, lb, main)
                _, qd = _commit(ctx, f, lb, perm)
                q = air.quotient_values(ctx, prog, md, 0, lb, qdb, alpha, perm_data=qd, challenges=chal,
                                        regional_sum=rsum, global_sum=gsum)
                got = q.to_host(f)
                q.free()
                md.free()
                qd.free()
                want = oracle.quotient_values(f, nodes, asserts, oracle.coset_lde(f, main, qdb, g, bit_reverse=False),
                                              log_n, qdb, alpha, (), None,
                                              oracle.coset_lde(f, perm, qdb, g, bit_reverse=False), chal, rsum, gsum)
                assert np.array_equal(got, want), (log_n, lb, valid)


@pytest.mark.parametrize("f", FIELDS)
def test_quotient_random_extension_air_matches_oracle(ctx, oracle, f):
    """Random extension-field constraints (every leaf kind, extension sums,
    differences, products and negations, base factors, filters) over 20 main
    and 6 permutation ext4 columns (24 base columns: windows past the width),
    heights 2^0..2^8, quotient degree 2^0..2^2."""
    rng = np.random.default_rng(83 + f)
    p = pico_amd.PRIMES[f]
    g = pico_amd.GENERATORS[f]
    builder, _ = airs.random_ext_air(f, rng, 20, 6, 60)
    prog = builder.compile()
    nodes, asserts = builder.node_array(), builder.asserts
    for log_n in (0, 3, 6, 8):
        for qdb, lb in ((0, 1), (1, 1), (2, 2)):
            n = 1 << log_n
            print(f"random ext quotient field {f} 2^{log_n} qdb={qdb} b={lb}", flush=True)
            main = rng.integers(0, p, (n, 20)).astype(np.uint32)
            perm = rng.integers(0, p, (n, 24)).astype(np.uint32)
            chal = [[int(x) for x in rng.integers(0, p, 4)] for _ in range(2)]
            rsum = [int(x) for x in rng.integers(0, p, 4)]
            gsum = [int(x) for x in rng.integers(0, p, 14)]
            alpha = [int(x) for x in rng.integers(0, p, 4)]
            _, md = _commit(ctx, f, lb, main)
            _, qd = _commit(ctx, f, lb, perm)
            q = air.quotient_values(ctx, prog, md, 0, lb, qdb, alpha, perm_data=qd, challenges=chal,
                                    regional_sum=rsum, global_sum=gsum)
            got = q.to_host(f)
            q.free()
            md.free()
            qd.free()
            want = oracle.quotient_values(f, nodes, asserts, oracle.coset_lde(f, main, qdb, g, bit_reverse=False),
                                          log_n, qdb, alpha, (), None,
                                          oracle.coset_lde(f, perm, qdb, g, bit_reverse=False), chal, rsum, gsum)
            assert np.array_equal(got, want), (log_n, qdb, lb)


def test_quotient_rejects_bad_requests(ctx):
    f = KOALABEAR
    builder, _ = airs.cube_air(f)
    prog = builder.compile()
    rng = np.random.default_rng(1)
    main, pv = airs.cube_trace(f, 16, rng)
    _, md = _commit(ctx, f, 1, main)
    try:
        with pytest.raises(pico_amd.PgError):  # Q would leave the committed LDE
            air.quotient_values(ctx, prog, md, 0, 1, 2, [1, 2, 3, 4], pv)
        with pytest.raises(pico_amd.PgError):  # missing public values
            air.quotient_values(ctx, prog, md, 0, 1, 1, [1, 2, 3, 4], pv[:2])
        wide, _ = airs.toy_air(f)
        wide.assert_zero(wide.main(7))
        with pytest.raises(pico_amd.PgError):  # column beyond the trace
            air.quotient_values(ctx, wide.compile(), md, 0, 1, 0, [1, 2, 3, 4])
        gated, _ = airs.gated_air(f)
        with pytest.raises(pico_amd.PgError):  # preprocessed trace missing
            air.quotient_values(ctx, gated.compile(), md, 0, 1, 0, [1, 2, 3, 4])
        look, _ = airs.lookup_air(f)
        lp = look.compile()
        ch = [[1, 2, 3, 4], [5, 6, 7, 8]]
        gs = list(range(14))
        with pytest.raises(pico_amd.PgError):  # permutation trace missing
            air.quotient_values(ctx, lp, md, 0, 1, 1, [1, 2, 3, 4], challenges=ch, regional_sum=[1, 0, 0, 0],
                                global_sum=gs)
        with pytest.raises(pico_amd.PgError):  # permutation trace of another height
            air.quotient_values(ctx, lp, md, 0, 1, 1, [1, 2, 3, 4], perm_data=_commit(
                ctx, f, 1, np.zeros((8, 8), np.uint32))[1], challenges=ch, regional_sum=[1, 0, 0, 0], global_sum=gs)
        qd = _commit(ctx, f, 1, np.zeros((16, 8), np.uint32))[1]
        with pytest.raises(pico_amd.PgError):  # one challenge of two
            air.quotient_values(ctx, lp, md, 0, 1, 1, [1, 2, 3, 4], perm_data=qd, challenges=ch[:1],
                                regional_sum=[1, 0, 0, 0], global_sum=gs)
        with pytest.raises(pico_amd.PgError):  # regional sum missing
            air.quotient_values(ctx, lp, md, 0, 1, 1, [1, 2, 3, 4], perm_data=qd, challenges=ch, global_sum=gs)
        with pytest.raises(pico_amd.PgError):  # global sum missing
            air.quotient_values(ctx, lp, md, 0, 1, 1, [1, 2, 3, 4], perm_data=qd, challenges=ch,
                                regional_sum=[1, 0, 0, 0])
        qd.free()
    finally:
        md.free()


@pytest.mark.parametrize("f", FIELDS)
def test_quotient_device_resident_chain_verifies(ctx, oracle, f):
    """The cube AIR (degree 3: two quotient chunks) on a 2^14-row trace:
    commit, quotient on the device from the committed LDE, chunks split as
    strided views and committed on their domains, Pcs::open of both rounds
    at zeta (main also at zeta w_n), then the verifier's equation: the
    chunks' opened values recombined with the zps factors equal the
    constraints folded at zeta times 1 / Z_H(zeta)."""
    p = pico_amd.PRIMES[f]
    g = pico_amd.GENERATORS[f]
    log_n, lb = 14, 1
    n = 1 << log_n
    rng = np.random.default_rng(11 + f)
    builder, qdb = airs.cube_air(f)
    prog = builder.compile()
    main, pv = airs.cube_trace(f, n, rng)
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
    m = ctx.upload(main, field=f)
    root_main, md = pcs.commit([(pcs.natural_domain_for_degree(n), m)])
    ch = pico_amd.DuplexChallenger(f)
    ch.observe(root_main)
    alpha = ch.sample_ext()
    q = air.quotient_values(ctx, prog, md, 0, lb, qdb, alpha, pv)
    assert np.array_equal(q.to_host(f), oracle.quotient_values(
        f, builder.node_array(), builder.asserts, oracle.coset_lde(f, main, qdb, g, bit_reverse=False), log_n, qdb,
        [int(x) for x in alpha], pv))
    d = 1 << qdb
    qdom = pico_amd.TwoAdicMultiplicativeCoset(log_n + qdb, g, f)
    chunks = qdom.split_evals(d, q)
    doms = qdom.split_domains(d)
    root_q, qd = pcs.commit(list(zip(doms, chunks)))
    for c in chunks:
        c.free()
    ch.observe(root_q)
    zeta = ch.sample_ext()
    wn = pico_amd.two_adic_generator(f, log_n)
    zn = (np.asarray(zeta, np.uint64) * wn % p).astype(np.uint32)
    rounds = [(md, [np.stack([zeta, zn])]), (qd, [np.asarray(zeta, np.uint32)[None, :]] * d)]
    opened, _ = pcs.open(rounds, ch, num_queries=8, pow_bits=1)
    ml, mn = opened[0][0][0], opened[0][0][1]
    folded, inv_zh = oracle.constraints_at_point(f, builder.node_array(), builder.asserts, ml, mn,
                                                 [int(x) for x in zeta], log_n, [int(x) for x in alpha], pv)
    want = [int(x) for x in oracle.ext_mul(f, folded, inv_zh)]

    def zp(shift, x):  # (x / shift)^n - 1, x ext4
        y = [int(v) * pow(shift, p - 2, p) % p for v in x]
        for _ in range(log_n):
            y = [int(v) for v in oracle.ext_mul(f, y, y)]
        y[0] = (y[0] - 1) % p
        return y

    shifts = [dm.shift for dm in doms]
    total = [0, 0, 0, 0]
    for i in range(d):
        z = [1, 0, 0, 0]
        for j in range(d):
            if j != i:
                num = zp(shifts[j], zeta)
                den = zp(shifts[j], [shifts[i], 0, 0, 0])
                z = [int(v) for v in oracle.ext_mul(f, z, oracle.ext_mul(f, num, oracle.ext_inv(f, den)))]
        chunk = [0, 0, 0, 0]
        for e in range(4):
            mono = [0, 0, 0, 0]
            mono[e] = 1
            v = oracle.ext_mul(f, mono, opened[1][i][0, e])
            chunk = [(a + int(b)) % p for a, b in zip(chunk, v)]
        total = [(a + int(b)) % p for a, b in zip(total, oracle.ext_mul(f, z, chunk))]
    assert total == want
    q.free()
    md.free()
    qd.free()
