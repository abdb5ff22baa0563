"""Job streams for the shard-commit workloads, drawn deterministically.

Shapes come from SplitMix64 seeded with (workload_id << 32) | 0; traces are
the device-side seeded inputs of pg_fill_random.  Each workload is run on the
device through the C ABI and reports wall time from the host clock.

  shard_commit_fri   one shard: 14 matrices with log-heights in 14..22 and
                     widths in {8,16,32,64,128,256,300}, one mixed-height
                     commit at blowup 2; a quotient vector (2^21 ext4) split
                     into 4 chunks on their coset domains (strided views, no
                     copy) and committed at blowup 8;
                     the open's values (zeta, g zeta) and reduced openings;
                     FRI on a 2^22 ext4 vector down to 2^6 with a commit per
                     round, 16-bit proof of work and 100 query openings
  shard_ingest       the same 14 traces in host memory (page-locked and
                     pageable): upload then commit vs pg_pcs_commit_host
  chip_tables        one shard's 40 chip tables (log-heights 10..22, mostly
                     small; widths 8..300) in one Pcs::commit vs an LDE
                     call per table then one MMCS commit
  quotient_eval      compute_quotient_values on the device for a 256-column
                     chip (2^20 rows, degree 3) against the readback it
                     replaces; and the LogUp side for 32 interactions in
                     batches of 2 (extension-field constraints over a
                     17-column permutation trace)
  batched_jobs       64 single-matrix commits, log-heights 10..18, widths
                     8..512, blowup 2/4/8: one launch sequence (hipGraph) vs
                     the same jobs committed one by one
  served_stream      Poisson arrivals of commit jobs (log-heights 16..22,
                     widths 16..512, KoalaBear/BabyBear, blowup 2 or 8), 50
                     openings per job, HBM-budgeted service vs sequential
"""
import math
import threading
import time

import numpy as np

from . import core

GOLDEN = 0x9E3779B97F4A7C15
MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s + GOLDEN) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform_int(self, lo, hi):  # inclusive
        return lo + self.next() % (hi - lo + 1)

    def uniform01(self):
        return (self.next() >> 11) / float(1 << 53)


def percentile(xs, q):
    s = sorted(xs)
    if not s:
        return None
    k = (len(s) - 1) * q / 100.0
    lo, hi = math.floor(k), math.ceil(k)
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


# ------------------------------------------------------------------ shard --
def shard_shapes(seed_base=3 << 32, n=14, min_log=14, max_log=22):
    r = SplitMix64(seed_base)
    widths = [8, 16, 32, 64, 128, 256, 300]
    return [(r.uniform_int(min_log, max_log), widths[r.next() % len(widths)]) for _ in range(n)]


def shard_commit_fri(ctx, shapes=None, log_fri=22, log_stop=6, queries=100, pow_bits=16,
                     quotient_log_h=19, quotient_chunks=4):
    """One shard's commits and FRI on the device; returns a timing dict."""
    f = core.KOALABEAR
    shapes = shapes or shard_shapes()
    mats = [ctx.random_matrix(f, (3 << 32) | (100 + i), 1 << L, w) for i, (L, w) in enumerate(shapes)]
    q = ctx.random_matrix(f, (3 << 32) | 200, quotient_chunks << quotient_log_h, 4)  # flattened ext4
    v = ctx.random_matrix(f, (3 << 32) | 300, 1 << log_fri, 4)
    core.device_synchronize()
    cells = sum((1 << L) * w for L, w in shapes)
    t = {}
    t0 = time.perf_counter()
    pcs = core.TwoAdicFriPcs(ctx, f, 1)
    root, main = pcs.commit([(pcs.natural_domain_for_degree(m.rows), m) for m in mats])
    t1 = time.perf_counter()
    qpcs = core.TwoAdicFriPcs(ctx, f, 3)
    p = core.PRIMES[f]
    # the quotient domain is the disjoint coset GEN <w>; chunk i lives on GEN w^i <w^d>
    qdom = core.TwoAdicMultiplicativeCoset(quotient_log_h + quotient_chunks.bit_length() - 1,
                                           core.GENERATORS[f], f)
    chunks = qdom.split_evals(quotient_chunks, q)
    qroot, quot = qpcs.commit(list(zip(qdom.split_domains(quotient_chunks), chunks)))
    for c in chunks:
        c.free()
    t2 = time.perf_counter()
    ch = core.DuplexChallenger(f)
    ch.observe(root)
    ch.observe(qroot)
    # Pcs::open's device part: every matrix at zeta and g zeta, the quotient
    # at zeta, then the reduced openings of both batches
    zeta = np.asarray(ch.sample_ext(), np.uint64)
    gz = zeta * core.two_adic_generator(f, max(L for L, _ in shapes)) % p
    pts2, pts1 = np.stack([zeta, gz]), zeta[None, :]
    t_o0 = time.perf_counter()
    main_vals = [pcs.open_values(main, k, pts2) for k in range(len(shapes))]
    quot_vals = [qpcs.open_values(quot, k, pts1) for k in range(quotient_chunks)]
    t_o1 = time.perf_counter()
    ro = pcs.reduce_openings([(main, [pts2] * len(shapes)), (quot, [pts1] * quotient_chunks)], [main_vals, quot_vals],
                             ch.sample_ext())
    t_o2 = time.perf_counter()
    for buf in ro.values():
        buf.free()
    fri = core.fri_commit_phase(ctx, f, ch, {log_fri: v}, log_stop, final_mode=1)
    t3 = time.perf_counter()
    witness = core.fri_grind(ctx, f, ch, pow_bits)
    t4 = time.perf_counter()
    idx = [ch.sample_bits(log_fri) for _ in range(queries)]
    sib, paths = core.fri_answer_queries(fri, idx, log_fri)
    mm = core.MerkleTreeMmcs(ctx, f)
    main_idx = [i >> (log_fri - main.log_height) if main.log_height <= log_fri else i for i in idx]
    mm.open_many(main_idx, main)
    mm.open_many([i >> max(0, log_fri - quot.log_height) for i in idx], quot)
    t5 = time.perf_counter()
    t.update({
        "cells": cells, "matrices": len(shapes), "shapes": shapes,
        "main_commit_ms": (t1 - t0) * 1e3, "quotient_commit_ms": (t2 - t1) * 1e3,
        "open_values_ms": (t_o1 - t_o0) * 1e3, "reduce_openings_ms": (t_o2 - t_o1) * 1e3,
        "fri_commit_phase_ms": (t3 - t2) * 1e3 - (t_o2 - t_o0) * 1e3, "pow_ms": (t4 - t3) * 1e3,
        "queries_ms": (t5 - t4) * 1e3, "total_ms": (t5 - t0) * 1e3,
        "reduced_heights": sorted(ro),
        "fri_rounds": fri.num_rounds, "pow_witness": witness,
    })
    # The whole TwoAdicFriPcs::open of the main round as one pg_pcs_open call
    # (what a Rust adapter makes): values of the 14 matrices at zeta and
    # g zeta, observed; alpha; reduced openings; FRI down to the blowup;
    # PoW; the queries and every opening they need.
    fri.free()
    fri = None
    ch2 = core.DuplexChallenger(f)
    ch2.observe(root)
    t6 = time.perf_counter()
    _, proof = pcs.open([(main, [pts2] * len(shapes))], ch2, queries, pow_bits)
    t["pcs_open_main_round_ms"] = pcs.last_open_seconds * 1e3
    t["pcs_open_main_round_with_unpacking_ms"] = (time.perf_counter() - t6) * 1e3
    t["pcs_open_fri_rounds"] = len(proof["commits"])
    main.free()
    quot.free()
    q.free()
    return t


def shard_ingest(ctx, shapes=None, reps=3):
    """The shard's main commit from traces in host memory: upload every
    matrix then commit, against pg_pcs_commit_host (uploads overlapped with
    the commit of the taller height groups), from page-locked and from
    pageable memory.  Host clock, best of `reps` (single calls of
    pg_pcs_commit_host vary by up to a third: the HIP runtime now and then
    holds an asynchronous copy call on the host for milliseconds); roots
    must agree."""
    f = core.KOALABEAR
    shapes = shapes or shard_shapes()
    pcs = core.TwoAdicFriPcs(ctx, f, 1)
    pinned = []
    for i, (L, w) in enumerate(shapes):
        m = ctx.random_matrix(f, (3 << 32) | (100 + i), 1 << L, w)
        pm = ctx.host_matrix(1 << L, w)
        core.check(core.lib().pg_download(ctx.handle, m.handle, 0, 1 << L, pm.ptr, -1, 0))
        m.free()
        pinned.append(pm)
    pageable = [np.array(pm.array) for pm in pinned]
    doms = [pcs.natural_domain_for_degree(1 << L) for L, _ in shapes]
    cells = sum((1 << L) * w for L, w in shapes)
    nbytes = 4 * cells
    out = {"cells": cells, "bytes": nbytes}
    roots = []
    for mem, arrays in (("pinned", [pm.array for pm in pinned]), ("pageable", pageable)):
        best = {}
        for _ in range(reps):
            core.device_synchronize()
            t0 = time.perf_counter()
            devs = [ctx.upload(a) for a in arrays]
            t1 = time.perf_counter()
            r1, pd = pcs.commit(list(zip(doms, devs)))
            t2 = time.perf_counter()
            pd.free()
            for d in devs:
                d.free()
            t3 = time.perf_counter()
            r2, pd = pcs.commit_host(list(zip(doms, arrays)), canonical=False)
            t4 = time.perf_counter()
            pd.free()
            roots += [r1, r2]
            cur = {"upload_ms": (t1 - t0) * 1e3, "upload_then_commit_ms": (t2 - t0) * 1e3,
                   "commit_host_ms": (t4 - t3) * 1e3}
            best = cur if not best or cur["commit_host_ms"] < best["commit_host_ms"] else best
        best["upload_GB_per_s"] = nbytes / (best["upload_ms"] / 1e3) / 1e9
        best["commit_host_cells_per_s"] = cells / (best["commit_host_ms"] / 1e3)
        out[mem] = best
    out["roots_identical"] = bool(all(np.array_equal(r, roots[0]) for r in roots))
    for pm in pinned:
        pm.free()
    return out


# ------------------------------------------------------------ chip tables --
def chip_table_shapes(seed_base=6 << 32, n=40, min_log=10, max_log=22):
    """One shard's chip tables: mostly small, a few tall (log-height
    min_log + floor((max_log - min_log + 1) u^3)), widths 8..300."""
    r = SplitMix64(seed_base)
    span = max_log - min_log + 1
    out = []
    for _ in range(n):
        u = r.uniform01()
        out.append((min(max_log, min_log + int(span * u ** 3)), r.uniform_int(8, 300)))
    return out


def lde_launches(heights):
    """LDE pass launches for matrices of these log-heights: one per matrix
    (1 pass up to 2^11 rows, 3 above) when committed one by one; one per
    distinct height in a single Pcs::commit."""
    per = [1 if L <= 11 else 3 for L in heights]
    return sum(per), sum(1 if L <= 11 else 3 for L in set(heights))


def chip_tables(ctx, shapes=None, reps=3, log_blowup=1):
    """One Pcs::commit of a shard's chip tables (their LDEs queued without
    host synchronisation, equal heights sharing launches) against the same
    commit composed one matrix at a time (an LDE call per matrix, each
    waiting for its kernels, then one MMCS commit)."""
    f = core.KOALABEAR
    shapes = shapes or chip_table_shapes()
    mats = [ctx.random_matrix(f, (6 << 32) | (100 + i), 1 << L, w) for i, (L, w) in enumerate(shapes)]
    pcs = core.TwoAdicFriPcs(ctx, f, log_blowup)
    dft = core.Radix2DitParallel(ctx, f)
    mmcs = core.MerkleTreeMmcs(ctx, f)
    gen = core.GENERATORS[f]
    cells = sum((1 << L) * w for L, w in shapes)

    def together():
        r, pd = pcs.commit([(pcs.natural_domain_for_degree(m.rows), m) for m in mats])
        pd.free()
        return r

    def one_by_one():
        ldes = [dft.coset_lde_batch(m, log_blowup, gen) for m in mats]
        r, pd = mmcs.commit(ldes, take_ownership=True)
        pd.free()
        return r

    out = {"tables": len(shapes), "cells": cells, "log_heights": sorted({L for L, _ in shapes})}
    for name, fn in (("one_by_one", one_by_one), ("together", together)):
        root = fn()  # warm
        ts = []
        for _ in range(reps):
            core.device_synchronize()
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        out[name + "_ms"] = min(ts) * 1e3
        out[name + "_root"] = [int(x) for x in root]
    out["roots_identical"] = out["one_by_one_root"] == out["together_root"]
    out["lde_launches_one_by_one"], out["lde_launches_together"] = lde_launches([L for L, _ in shapes])
    out["speedup"] = out["one_by_one_ms"] / out["together_ms"]
    for m in mats:
        m.free()
    return out


# --------------------------------------------------------------- quotient --
def quotient_eval(ctx, log_n=20, width=256, log_blowup=1, reps=3):
    """compute_quotient_values on the device for one wide chip: a 2^log_n x
    width trace committed at blowup 2^log_blowup, one degree-3 transition
    constraint per column (main[c] main[c+1] main[c+2] - next.main[c] under
    is_transition: log quotient degree 1, |Q| = 2^(log_n+1)), against the
    readback of the same rows that the host-side evaluation needs
    (get_evaluations_on_domain, prover.rs:467-490)."""
    from . import air
    f = core.KOALABEAR
    b = air.AirBuilder(f)
    for c in range(width):
        x, y, z = b.main(c), b.main((c + 1) % width), b.main((c + 2) % width)
        b.when_transition().assert_zero(x * y * z - b.main(c, next=True))
    prog = b.compile()
    trace = ctx.random_matrix(f, (2 << 32) | 0, 1 << log_n, width)
    pcs = core.TwoAdicFriPcs(ctx, f, log_blowup)
    _, pd = pcs.commit([(pcs.natural_domain_for_degree(1 << log_n), trace)])
    trace.free()
    alpha = [31, 41, 59, 26]
    q = air.quotient_values(ctx, prog, pd, 0, log_blowup, 1, alpha)  # warm
    q.free()
    ts = []
    for _ in range(reps):
        core.device_synchronize()
        t0 = time.perf_counter()
        q = air.quotient_values(ctx, prog, pd, 0, log_blowup, 1, alpha)
        ts.append(time.perf_counter() - t0)
        q.free()
    ctx.profile(True)
    ctx.profile_reset()
    q = air.quotient_values(ctx, prog, pd, 0, log_blowup, 1, alpha)
    ctx.profile(False)
    kms = ctx.profile_get("quotient")[0]
    q.free()
    rows_q = 1 << (log_n + 1)
    host = np.zeros((rows_q, width), np.uint32)
    rb = []
    for _ in range(reps):
        t0 = time.perf_counter()
        core.check(core.lib().pg_pcs_evaluations(ctx.handle, pd.handle, 0, rows_q, core._ptr(host)))
        rb.append(time.perf_counter() - t0)
    del host
    pd.free()
    return {"trace": [1 << log_n, width], "quotient_rows": rows_q, "constraints": width,
            "program_instructions": prog.ninstr, "program_registers": prog.nregs,
            "ms": min(ts) * 1e3, "kernel_ms": kms, "rows_per_s": rows_q / min(ts),
            "readback_it_replaces_ms": min(rb) * 1e3}


def lookup_air(width, interactions, batch):
    """The LogUp constraints Pico builds for a chip's lookups
    (vm/src/machine/permutation.rs:183-346) for `interactions` interactions
    over a `width`-column trace in batches of `batch` (the reference's batch
    size at blowup 2 keeps the degree at 3): interaction i sends (kind i % 8,
    main[2i], main[2i+1]) with multiplicity main[2i+2] (every other one is a
    receive); per batch rlc_1 ... rlc_b entry = sum_i m_i prod_{j != i}
    rlc_j; the running sum phi with its first-row, transition and last-row
    constraints."""
    from . import air
    f = core.KOALABEAR
    b = air.AirBuilder(f)
    alpha, beta = b.perm_challenge(0), b.perm_challenge(1)
    beta2 = beta * beta
    nb = (interactions + batch - 1) // batch
    col = lambda i: b.main(i % width)
    for k in range(nb):
        rlcs, mults = [], []
        for i in range(k * batch, min(interactions, (k + 1) * batch)):
            rlcs.append(alpha + b.constant(i % 8) + beta * col(2 * i) + beta2 * col(2 * i + 1))
            m = col(2 * i + 2)
            mults.append(m if i % 2 == 0 else -m)
        prod, num = None, None
        for i, (m, r) in enumerate(zip(mults, rlcs)):
            prod = r if prod is None else prod * r
            others = None
            for j, o in enumerate(rlcs):
                if j != i:
                    others = o if others is None else others * o
            term = m * others if others is not None else m
            num = term if num is None else num + term
        b.assert_zero(prod * b.permutation(k) - num)
    entries = [b.permutation(k) for k in range(nb)]
    entries_n = [b.permutation(k, next=True) for k in range(nb)]
    phi, phi_n = b.permutation(nb), b.permutation(nb, next=True)
    s_loc, s_nxt = entries[0], entries_n[0]
    for k in range(1, nb):
        s_loc, s_nxt = s_loc + entries[k], s_nxt + entries_n[k]
    b.when_first_row().assert_zero(phi - s_loc)
    b.when_transition().assert_zero(phi_n - phi - s_nxt)
    b.when_last_row().assert_zero(phi - b.regional_sum())
    return b, nb + 1


def quotient_eval_lookups(ctx, log_n=20, width=256, interactions=32, batch=2, log_blowup=1, reps=3):
    """The permutation (LogUp) side of compute_quotient_values on the device
    for a lookup-heavy chip: `interactions` interactions in batches of
    `batch` over the 2^log_n x width main trace and a permutation trace of
    interactions / batch + 1 ext4 columns, both committed at blowup
    2^log_blowup (random values: the timing does not depend on validity);
    log quotient degree 1."""
    from . import air
    f = core.KOALABEAR
    b, perm_cols = lookup_air(width, interactions, batch)
    prog = b.compile()
    n = 1 << log_n
    pcs = core.TwoAdicFriPcs(ctx, f, log_blowup)
    dom = pcs.natural_domain_for_degree(n)
    trace = ctx.random_matrix(f, (2 << 32) | 0, n, width)
    perm = ctx.random_matrix(f, (2 << 32) | 1, n, 4 * perm_cols)
    _, md = pcs.commit([(dom, trace)])
    _, qd = pcs.commit([(dom, perm)])
    trace.free()
    perm.free()
    alpha = [31, 41, 59, 26]
    chal = [[2, 7, 1, 8], [3, 1, 4, 1]]
    kw = dict(perm_data=qd, challenges=chal, regional_sum=[5, 9, 2, 6])
    air.quotient_values(ctx, prog, md, 0, log_blowup, 1, alpha, **kw).free()  # warm
    ts = []
    for _ in range(reps):
        core.device_synchronize()
        t0 = time.perf_counter()
        q = air.quotient_values(ctx, prog, md, 0, log_blowup, 1, alpha, **kw)
        ts.append(time.perf_counter() - t0)
        q.free()
    ctx.profile(True)
    ctx.profile_reset()
    air.quotient_values(ctx, prog, md, 0, log_blowup, 1, alpha, **kw).free()
    ctx.profile(False)
    kms = ctx.profile_get("quotient")[0]
    md.free()
    qd.free()
    rows_q = 1 << (log_n + 1)
    return {"trace": [n, width], "interactions": interactions, "batch": batch, "perm_ext_columns": perm_cols,
            "quotient_rows": rows_q, "constraints": len(b.asserts), "program_instructions": prog.ninstr,
            "program_registers": prog.nregs, "ms": min(ts) * 1e3, "kernel_ms": kms, "rows_per_s": rows_q / min(ts)}


# ---------------------------------------------------------------- batched --
def batched_shapes(seed_base=4 << 32, n=64):
    r = SplitMix64(seed_base)
    return [(r.uniform_int(10, 18), r.uniform_int(8, 512), r.uniform_int(1, 3)) for _ in range(n)]


def batched_jobs(ctx, shapes=None, reps=3):
    """Batched (graph replay) vs one-by-one commits of the same jobs."""
    f = core.KOALABEAR
    shapes = shapes or batched_shapes()
    mats = [ctx.random_matrix(f, (4 << 32) | (1000 + i), 1 << L, w) for i, (L, w, _) in enumerate(shapes)]
    cells = sum((1 << L) * w for L, w, _ in shapes)

    def one_by_one():
        roots = []
        for m, (L, w, lb) in zip(mats, shapes):
            pcs = core.TwoAdicFriPcs(ctx, f, lb)
            r, pd = pcs.commit([(pcs.natural_domain_for_degree(m.rows), m)])
            pd.free()
            roots.append(r)
        return roots

    seq_roots = one_by_one()  # warm
    t0 = time.perf_counter()
    for _ in range(reps):
        one_by_one()
    t_seq = (time.perf_counter() - t0) / reps
    batch = core.CommitBatch(ctx, f, [(m, lb) for m, (_, _, lb) in zip(mats, shapes)], use_graph=True)
    roots = batch.run()  # warm
    t0 = time.perf_counter()
    for _ in range(reps):
        roots = batch.run()
    t_bat = (time.perf_counter() - t0) / reps
    same = all(np.array_equal(a, b) for a, b in zip(roots, seq_roots))
    out = {"jobs": len(shapes), "cells": cells, "launches": batch.num_launches,
           "one_by_one_ms": t_seq * 1e3, "batched_ms": t_bat * 1e3,
           "one_by_one_cells_per_s": cells / t_seq, "batched_cells_per_s": cells / t_bat,
           "speedup": t_seq / t_bat, "roots_identical": bool(same)}
    batch.free()
    return out


# ----------------------------------------------------------------- served --
def served_jobs(seed_base=5 << 32, n=40):
    """(field, log_h, width, log_blowup) per job.  A KoalaBear LDE has at most
    2^24 rows (two-adicity 24), so KoalaBear jobs with log_h + b > 24 use b = 1."""
    r = SplitMix64(seed_base)
    jobs = []
    for _ in range(n):
        L = r.uniform_int(16, 22)
        w = r.uniform_int(16, 512)
        f = core.KOALABEAR if r.next() & 1 else core.BABYBEAR
        b = 1 if r.next() & 1 else 3
        if f == core.KOALABEAR and L + b > core.TWO_ADICITY[f]:
            b = 1
        jobs.append((f, L, w, b))
    return jobs


def served_host_traces(ctx, jobs):
    """The jobs' traces in ordinary (pageable) host memory, Montgomery words:
    the same seeded matrices the synthetic jobs fill on the device, as a
    shard's records reach the prover from the emulator."""
    out = []
    for k, (f, L, w, _) in enumerate(jobs):
        m = ctx.random_matrix(f, (5 << 32) | (10000 + k), 1 << L, w)
        out.append(m.to_host())
        m.free()
    return out


def served_stream(ctx, rate, jobs, budget=200 * 10 ** 9, workers=3, openings=50, seed=5, host_traces=None):
    """Poisson arrivals at `rate` jobs/s into the commit service; each job is
    finished when its 50 openings have been answered.  Returns throughput and
    latency percentiles (arrival -> last opening).  With host_traces the jobs
    arrive as host arrays and their upload is part of each job."""
    svc = core.CommitService(ctx, budget=budget, workers=workers)
    rng = SplitMix64((5 << 32) | seed)
    arrivals, t = [], 0.0
    for _ in jobs:
        t += -math.log(1.0 - rng.uniform01()) / rate
        arrivals.append(t)
    lat = [None] * len(jobs)
    ids = [None] * len(jobs)
    ready = threading.Semaphore(0)
    start = time.perf_counter()

    def submitter():
        for k, (f, L, w, b) in enumerate(jobs):
            delay = start + arrivals[k] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            if host_traces is not None:
                ids[k] = svc.submit_host(f, host_traces[k], b, canonical=False)
            else:
                ids[k] = svc.submit_synthetic(f, (5 << 32) | (10000 + k), 1 << L, w, b)
            ready.release()

    def finisher(k):
        f, L, w, b = jobs[k]
        svc.wait(ids[k], f)
        q = np.random.default_rng(k).integers(0, 1 << (L + b), openings)
        svc.open(ids[k], f, q, w, L + b)
        lat[k] = time.perf_counter() - (start + arrivals[k])
        svc.release(ids[k])

    th = threading.Thread(target=submitter)
    th.start()
    fins = []
    for k in range(len(jobs)):
        ready.acquire()
        ft = threading.Thread(target=finisher, args=(k,))
        ft.start()
        fins.append(ft)
    th.join()
    for ft in fins:
        ft.join()
    wall = time.perf_counter() - start
    stats = svc.stats()
    svc.close()
    cells = sum((1 << L) * w for _, L, w, _ in jobs)
    return {"rate": rate, "jobs": len(jobs), "wall_s": wall, "jobs_per_s": len(jobs) / wall,
            "cells_per_s": cells / wall, "p50_ms": percentile(lat, 50) * 1e3,
            "p99_ms": percentile(lat, 99) * 1e3, "peak_running": stats["peak_running"]}


def served_sequential(ctx, jobs, openings=50, host_traces=None):
    """The plain path: each job committed, opened 50 times and released before
    the next starts (all jobs present at time 0); from host_traces through
    the host-memory commit when given."""
    start = time.perf_counter()
    lat = []
    for k, (f, L, w, b) in enumerate(jobs):
        pcs = core.TwoAdicFriPcs(ctx, f, b)
        dom = pcs.natural_domain_for_degree(1 << L)
        if host_traces is not None:
            _, pd = pcs.commit_host([(dom, host_traces[k])], canonical=False)
        else:
            m = ctx.random_matrix(f, (5 << 32) | (10000 + k), 1 << L, w)
            _, pd = pcs.commit([(dom, m)])
            m.free()
        q = np.random.default_rng(k).integers(0, 1 << (L + b), openings)
        core.MerkleTreeMmcs(ctx, f).open_many(q, pd)
        pd.free()
        lat.append(time.perf_counter() - start)
    wall = time.perf_counter() - start
    cells = sum((1 << L) * w for _, L, w, _ in jobs)
    return {"jobs": len(jobs), "wall_s": wall, "jobs_per_s": len(jobs) / wall, "cells_per_s": cells / wall,
            "p50_ms": percentile(lat, 50) * 1e3, "p99_ms": percentile(lat, 99) * 1e3}
