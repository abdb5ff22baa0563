"""pg_pcs_open_values_noalloc, the allocation-free form of
pg_pcs_open_values (values of a committed matrix at out-of-domain points; the
device part of TwoAdicFriPcs::open, vm/src/machine/prover.rs:588-589).
Called through the C ABI and compared with pg_pcs_open_values on calls in a
row, each smaller than the one before, with the next device allocation armed
to fail: the form must not allocate, and it must not depend on what an
earlier call left in its scratch array."""
import numpy as np
import pytest

import pico_amd
from pico_amd import BABYBEAR, KOALABEAR
from pico_amd._lib import PG_E_INVALID, check, lib
from pico_amd.core import _ptr, to_monty

pytestmark = pytest.mark.gpu


def _points(field, seed, n):
    rng = np.random.default_rng(seed)
    return to_monty(field, rng.integers(0, pico_amd.PRIMES[field], (n, 4), dtype=np.uint64))


def _commit(ctx, field, seed, log_h, w, lb):
    m = ctx.random_matrix(field, seed, 1 << log_h, w)
    pcs = pico_amd.TwoAdicFriPcs(ctx, field, lb)
    _, pd = pcs.commit([(pcs.natural_domain_for_degree(1 << log_h), m)])
    m.free()
    return pd


def _open(fn, ctx, pd, lb, pts, w):
    out = np.full((len(pts), w, 4), 0xFFFFFFFF, np.uint32)
    rc = fn(ctx.handle, pd.handle, 0, lb, _ptr(pts), len(pts), _ptr(out))
    return rc, out


@pytest.mark.parametrize("field", [KOALABEAR, BABYBEAR])
def test_noalloc_open_values_equal_allocating_form(ctx, field):
    L = lib()
    # (log trace height, width, log_blowup, points): the largest first, then
    # smaller ones, so stale scratch contents would show
    cases = [(20, 256, 1, 2), (14, 100, 3, 3), (9, 7, 2, 1)]
    for i, (log_h, w, lb, npts) in enumerate(cases):
        print(f"noalloc open values: field {field} 2^{log_h} x {w} b={lb} points={npts}", flush=True)
        pd = _commit(ctx, field, (7 << 32) | (10 * field + i), log_h, w, lb)
        try:
            pts = np.ascontiguousarray(_points(field, 100 + i, npts))
            in_use = L.pg_bytes_in_use(ctx.handle)
            check(L.pg_debug_fail_alloc(ctx.handle, 0))  # any device allocation now fails
            try:
                rc, got = _open(L.pg_pcs_open_values_noalloc, ctx, pd, lb, pts, w)
            finally:
                check(L.pg_debug_fail_alloc(ctx.handle, -1))
            check(rc)
            assert L.pg_bytes_in_use(ctx.handle) == in_use
            rc, ref = _open(L.pg_pcs_open_values, ctx, pd, lb, pts, w)
            check(rc)
            assert np.array_equal(got, ref), (field, log_h, w, lb, npts)
        finally:
            pd.free()


def test_noalloc_open_values_rejects_oversize(ctx):
    # trace height 2^21 at two points: 8 MiB of twiddles + 64 MiB of weights
    pd = _commit(ctx, KOALABEAR, (7 << 32) | 99, 21, 8, 1)
    try:
        pts = np.ascontiguousarray(_points(KOALABEAR, 5, 2))
        rc, _ = _open(lib().pg_pcs_open_values_noalloc, ctx, pd, 1, pts, 8)
        assert rc == PG_E_INVALID
        assert b"scratch" in lib().pg_last_error()
        rc, got = _open(lib().pg_pcs_open_values_noalloc, ctx, pd, 1, pts[:1], 8)  # one point fits
        check(rc)
        rc, ref = _open(lib().pg_pcs_open_values, ctx, pd, 1, pts[:1], 8)
        check(rc)
        assert np.array_equal(got, ref)
    finally:
        pd.free()
