"""Every alternative LDE tile shape gives the oracle's LDE (every pass kind,
transform sizes 2^7..2^11, both fields, narrow and wide widths), batches
built under the alternative shapes give the oracle's roots, and the
autotuner runs, records its decisions and leaves results unchanged."""
import numpy as np
import pytest

import pico_amd
from pico_amd import BABYBEAR, KOALABEAR

pytestmark = pytest.mark.gpu


@pytest.fixture
def table():
    saved = pico_amd.tune_export()
    yield saved
    pico_amd.tune_import(saved)


def _all_alternative(text):
    return "".join(" ".join(line.split()[:4]) + " 1\n" for line in text.splitlines() if line)


# (log_h, width, log_blowup): single-pass sizes 2^7..2^11 and four-step
# sizes whose strided and block transforms are 2^7..2^11
CASES = [(7, 9, 1), (8, 3, 2), (9, 40, 1), (10, 5, 3), (11, 17, 1), (14, 9, 1), (16, 33, 2), (18, 4, 1),
         (20, 2, 1), (20, 37, 1), (21, 3, 1), (22, 2, 1)]


@pytest.mark.parametrize("f", [KOALABEAR, BABYBEAR])
def test_alternative_shapes_match_oracle(ctx, oracle, table, f):
    pico_amd.tune_import(_all_alternative(table))
    dft = pico_amd.Radix2DitParallel(ctx, f)
    for L, w, lb in CASES:
        a = oracle.random_matrix(f, 9000 + 7 * L + w, 1 << L, w)
        out = dft.coset_lde_batch(ctx.upload(a, field=f), lb, pico_amd.GENERATORS[f])
        assert np.array_equal(out.to_host(f), oracle.coset_lde(f, a, lb, pico_amd.GENERATORS[f])), (L, w, lb)
        out.free()


def test_batch_under_alternative_shapes(ctx, oracle, table):
    f = KOALABEAR
    pico_amd.tune_import(_all_alternative(table))
    shapes = [(14, 12, 1), (16, 40, 2), (14, 3, 3), (9, 64, 1), (16, 8, 1)]
    mats = [ctx.random_matrix(f, 9500 + i, 1 << L, w) for i, (L, w, _) in enumerate(shapes)]
    batch = pico_amd.CommitBatch(ctx, f, [(m, lb) for m, (_, _, lb) in zip(mats, shapes)], use_graph=False)
    roots = batch.run()
    for i, (m, (L, w, lb)) in enumerate(zip(mats, shapes)):
        host = m.to_host(f)
        lde = oracle.coset_lde(f, host, lb, pico_amd.GENERATORS[f])
        assert np.array_equal(roots[i], oracle.MerkleTree(f, [lde]).root), i
    batch.free()


def test_autotune_records_decisions(ctx, oracle, table):
    f = KOALABEAR
    report = pico_amd.autotune(ctx, f, 10, 14, [8, 64], log_blowup=1, reps=1)
    lines = [l for l in report.splitlines() if l and not l.startswith("#")]
    assert lines, report
    for l in lines:
        parts = l.split()
        choice = int(parts[-1])
        ms0, ms1 = float(parts[4]), float(parts[5])
        if ms1 != ms0:  # the report rounds; equal-looking times may go either way
            assert choice == (1 if ms1 < ms0 else 0), l
        kind = pico_amd.LDE_PASSES[parts[1].replace("lde_pass_", "").replace("lde_", "")]
        width = {0: 8, 1: 32, 2: 128, 3: 256}[int(parts[3])]
        assert pico_amd.tune_get(f, kind, int(parts[2]), width) == choice
    # results are the same whichever shapes were chosen
    a = oracle.random_matrix(f, 9900, 1 << 14, 64)
    out = pico_amd.Radix2DitParallel(ctx, f).coset_lde_batch(ctx.upload(a, field=f), 1, pico_amd.GENERATORS[f])
    assert np.array_equal(out.to_host(f), oracle.coset_lde(f, a, 1, pico_amd.GENERATORS[f]))
    out.free()
    # the table survives an export / import round trip
    text = pico_amd.tune_export()
    pico_amd.tune_import(table)
    pico_amd.tune_import(text)
    assert pico_amd.tune_export() == text
