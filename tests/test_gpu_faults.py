"""Error paths under injected allocation failures (pg_debug_fail_alloc).

For each entry point the k-th device allocation of the call fails, for
k = 0, 1, ... until the call needs fewer allocations than k and succeeds.
Every failing call must return PG_E_OOM, leave the context's live device
bytes where they were (nothing leaked, nothing freed twice), and the call
that finally succeeds must give the undisturbed result.  The reference has
no such path (it panics on allocation failure); these are the library's own
guarantees.  Also: a queued service job can be cancelled."""
import numpy as np
import pytest

import pico_amd
from pico_amd import BABYBEAR, KOALABEAR
from pico_amd._lib import PG_E_CANCELLED, PG_E_INVALID, PG_E_OOM, PgError

pytestmark = pytest.mark.gpu


@pytest.fixture
def fctx():
    c = pico_amd.Context(0)
    yield c
    c.debug_fail_alloc(-1)
    c.close()


def sweep(ctx, op, limit=64):
    """Run op once undisturbed (also warms the table caches), then with the
    k-th allocation failing; returns how many allocations the call makes."""
    ref = op(ctx)
    base = ctx.bytes_in_use
    for k in range(limit):
        ctx.debug_fail_alloc(k)
        try:
            got = op(ctx)
        except PgError as e:
            ctx.debug_fail_alloc(-1)
            assert e.code == PG_E_OOM, (k, str(e))
            assert ctx.bytes_in_use == base, (k, ctx.bytes_in_use - base)
            continue
        ctx.debug_fail_alloc(-1)
        assert len(got) == len(ref)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), k
        assert ctx.bytes_in_use == base
        return k
    pytest.fail(f"the call still fails after {limit} injected failures")


def _free(*objs):
    for o in objs:
        if o is not None:
            o.free()


def op_coset_lde(ctx):
    f = KOALABEAR
    m = out = None
    try:
        m = ctx.random_matrix(f, 11, 1 << 13, 20)
        out = pico_amd.Radix2DitParallel(ctx, f).coset_lde_batch(m, 2, pico_amd.GENERATORS[f])
        return [out.to_host(f)]
    finally:
        _free(out, m)


def op_pcs_commit(ctx):
    f = BABYBEAR
    p = pico_amd.PRIMES[f]
    mats, pd = [], None
    try:
        for i, (L, w) in enumerate([(12, 10), (10, 9), (12, 3)]):
            mats.append(ctx.random_matrix(f, 20 + i, 1 << L, w))
        pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
        doms = [pcs.natural_domain_for_degree(m.rows) for m in mats]
        doms[1] = pico_amd.TwoAdicMultiplicativeCoset(10, pico_amd.GENERATORS[f] * pico_amd.two_adic_generator(f, 12) % p, f)
        root, pd = pcs.commit(list(zip(doms, mats)))
        return [root, pcs.get_evaluations_on_domain(pd, 1, pico_amd.TwoAdicMultiplicativeCoset(10, pico_amd.GENERATORS[f]))]
    finally:
        _free(pd, *mats)


def op_commit_host(ctx):
    """Page-locked Montgomery input wide enough to be cut into column blocks
    (sponge state buffer, pitched uploads) plus an injected shorter matrix."""
    f = KOALABEAR
    pinned, pd = [], None
    try:
        for i, (L, w) in enumerate([(17, 160), (15, 9)]):
            m = ctx.random_matrix(f, 30 + i, 1 << L, w)
            pm = ctx.host_matrix(1 << L, w)
            pinned.append(pm)
            pm.array[:] = m.to_host()
            m.free()
        pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
        root, pd = pcs.commit_host([(pcs.natural_domain_for_degree(pm.array.shape[0]), pm.array) for pm in pinned],
                                   canonical=False)
        return [root]
    finally:
        _free(pd, *pinned)


def op_mmcs_open(ctx):
    f = KOALABEAR
    mats, pd = [], None
    try:
        for i, (L, w) in enumerate([(9, 8), (8, 5), (9, 13)]):
            mats.append(ctx.random_matrix(f, 40 + i, 1 << L, w))
        mm = pico_amd.MerkleTreeMmcs(ctx, f)
        root, pd = mm.commit(mats)
        rows, path = mm.open_batch(77, pd)
        return [root, np.concatenate([np.asarray(r).ravel() for r in rows]), np.asarray(path)]
    finally:
        _free(pd, *mats)


def op_open(ctx):
    """Values at two points, then reduced openings over two heights."""
    f = KOALABEAR
    p = pico_amd.PRIMES[f]
    mats, pd, ro = [], None, {}
    try:
        for i, (L, w) in enumerate([(10, 12), (8, 7)]):
            mats.append(ctx.random_matrix(f, 50 + i, 1 << L, w))
        pcs = pico_amd.TwoAdicFriPcs(ctx, f, 1)
        _, pd = pcs.commit([(pcs.natural_domain_for_degree(m.rows), m) for m in mats])
        zeta = np.array([5, 6, 7, 8], np.uint64)
        pts = np.stack([zeta, zeta * pico_amd.two_adic_generator(f, 10) % p])
        vals = [pcs.open_values(pd, k, pts) for k in range(2)]
        ro = pcs.reduce_openings([(pd, [pts, pts])], [vals], np.array([9, 8, 7, 6], np.uint64))
        return vals + [ro[lh].to_host(f) for lh in sorted(ro)]
    finally:
        _free(*ro.values())
        _free(pd, *mats)


def op_fri(ctx):
    f = KOALABEAR
    v = fri = None
    try:
        v = ctx.random_matrix(f, 60, 1 << 12, 4)
        ch = pico_amd.DuplexChallenger(f)
        fri = pico_amd.core.fri_commit_phase(ctx, f, ch, {12: v}, 3, final_mode=1)
        sib, _ = pico_amd.core.fri_answer_queries(fri, [5, 4000, 123], 12)
        w = pico_amd.core.fri_grind(ctx, f, ch, 8)
        return [fri.commits, fri.betas, sib, np.array([w])]
    finally:
        _free(fri, v)


def op_batch(ctx):
    f = BABYBEAR
    mats, b = [], None
    try:
        for i, (L, w, lb) in enumerate([(12, 16, 1), (9, 40, 3), (13, 8, 2)]):
            mats.append(ctx.random_matrix(f, 70 + i, 1 << L, w))
        b = pico_amd.CommitBatch(ctx, f, [(m, lb) for m, (_, _, lb) in zip(mats, [(12, 16, 1), (9, 40, 3), (13, 8, 2)])],
                                 use_graph=False)
        return [b.run()]
    finally:
        _free(b, *mats)


def op_view_download(ctx):
    f = KOALABEAR
    m = v = None
    try:
        m = ctx.random_matrix(f, 80, 256, 12)
        v = m.view(row0=1, row_step=3, col0=2, cols=5)
        return [v.to_host(f)]  # the canonical download stages through a device buffer
    finally:
        _free(v, m)


OPS = {"coset_lde": op_coset_lde, "pcs_commit": op_pcs_commit, "commit_host": op_commit_host,
       "mmcs_open": op_mmcs_open, "open": op_open, "fri": op_fri, "batch": op_batch,
       "view_download": op_view_download}


@pytest.mark.parametrize("name", sorted(OPS))
def test_allocation_failures_leave_nothing_behind(fctx, name):
    n = sweep(fctx, OPS[name])
    assert n >= 1  # every one of these calls allocates


def test_service_job_failure_and_cancel(fctx):
    f = KOALABEAR
    svc = pico_amd.CommitService(fctx, budget=0, workers=1)
    try:
        # a job whose allocation fails reports PG_E_OOM from wait and can be released
        fctx.debug_fail_alloc(0)
        bad = svc.submit_synthetic(f, 90, 1 << 12, 16, 1)
        with pytest.raises(PgError) as ei:
            svc.wait(bad, f)
        assert ei.value.code == PG_E_OOM
        fctx.debug_fail_alloc(-1)
        svc.release(bad)
        # one worker: the last of three queued jobs is still waiting when cancelled
        a = svc.submit_synthetic(f, 91, 1 << 20, 64, 1)
        b = svc.submit_synthetic(f, 92, 1 << 12, 16, 1)
        c = svc.submit_synthetic(f, 93, 1 << 12, 16, 1)
        svc.cancel(c)
        with pytest.raises(PgError) as ei:
            svc.wait(c, f)
        assert ei.value.code == PG_E_CANCELLED
        ra, _ = svc.wait(a, f)
        rb, _ = svc.wait(b, f)
        with pytest.raises(PgError) as ei:
            svc.cancel(a)  # finished
        assert ei.value.code == PG_E_INVALID
        for j in (a, b, c):
            svc.release(j)
        assert svc.stats()["reserved_bytes"] == 0 and svc.stats()["pending"] == 0
        for seed, L, w, r in ((91, 20, 64, ra), (92, 12, 16, rb)):
            m = fctx.random_matrix(f, seed, 1 << L, w)
            pcs = pico_amd.TwoAdicFriPcs(fctx, f, 1)
            root, pd = pcs.commit([(pcs.natural_domain_for_degree(1 << L), m)])
            assert np.array_equal(root, r)
            pd.free()
            m.free()
    finally:
        svc.close()
