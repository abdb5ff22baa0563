"""Writes the golden fixtures of the large commit workloads, computed by the
CPU oracle (oracle/oracle.c) on the seeded inputs the GPU tests and the
benchmark draw on the device (pg_fill_random: the same SplitMix64 stream as
oracle.random_matrix).

    python tests/golden/make_golden.py headline   # 2^20 x 256 KoalaBear commit
    python tests/golden/make_golden.py batched    # 64 single-matrix jobs
    python tests/golden/make_golden.py served     # the KoalaBear/BabyBear service mix
    python tests/golden/make_golden.py shard      # 14-matrix shard + quotient chunks
    python tests/golden/make_golden.py all

Roots come from oracle.pcs_commit_seeded (LDEs in column blocks, pinned
against oracle.MerkleTree of whole oracle.coset_lde outputs in
tests/test_oracle.py).  All values are canonical.  Takes about 30 minutes on
8 cores for `all`.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402

KB, BB = oracle.KOALABEAR, oracle.BABYBEAR


def _write(name, obj):
    path = os.path.join(HERE, name)
    with open(path, "w") as f:
        json.dump(obj, f, indent=1)
        f.write("\n")
    print(f"wrote {path}", flush=True)


def _ints(a):
    return [int(x) for x in np.asarray(a).ravel()]


def headline():
    """The benchmark's step: Pcs::commit of one KoalaBear 2^20 x 256 trace
    (seed (2 << 32) | 0) at log_blowup 1 (vm/src/machine/prover.rs:238-244,
    parameters vm/src/configs/stark_config/kb_poseidon2.rs:60-93).  Besides
    the root: LDE samples (committed row order) of 8 columns x 40 rows."""
    t0 = time.time()
    f, log_h, w, lb, seed = KB, 20, 256, 1, (2 << 32) | 0
    h = 1 << log_h
    root = oracle.pcs_commit_seeded(f, [seed], [h], [w], lb)
    trace = oracle.random_matrix(f, seed, h, w)
    cols = [0, 1, 7, 8, 100, 128, 201, 255]
    lde = oracle.coset_lde(f, np.ascontiguousarray(trace[:, cols]), lb, oracle.generator(f))
    del trace
    rng = np.random.default_rng(2024)
    rows = sorted(set([0, 1, (h << lb) - 1, h] + list(rng.integers(0, h << lb, 36))))
    samples = [[int(r), int(c), int(lde[r, j])] for r in rows for j, c in enumerate(cols)]
    _write("headline_kb_2p20x256.json", {
        "_source": "oracle/oracle.c (or_pcs_commit_seeded, or_coset_lde) on the seeded trace; "
                   "tests/golden/make_golden.py headline",
        "field": "koalabear", "seed": seed, "rows": h, "cols": w, "log_blowup": lb,
        "root": _ints(root), "lde_samples_row_col_value": samples})
    print(f"headline: {time.time() - t0:.1f} s", flush=True)


def batched():
    """64 independent single-matrix commits (log_h 10..18, widths 8..512,
    log_blowup 1..3), matrix j seeded (4 << 32) | (1000 + j)."""
    from pico_amd import workloads
    t0 = time.time()
    shapes = workloads.batched_shapes()
    roots = []
    for j, (L, w, lb) in enumerate(shapes):
        roots.append(_ints(oracle.pcs_commit_seeded(KB, [(4 << 32) | (1000 + j)], [1 << L], [w], lb)))
        print(f"  batched job {j}: 2^{L} x {w} b={lb} ({time.time() - t0:.0f} s)", flush=True)
    _write("batched_kb_64jobs.json", {
        "_source": "oracle/oracle.c (or_pcs_commit_seeded); tests/golden/make_golden.py batched",
        "field": "koalabear", "seed_rule": "(4 << 32) | (1000 + j)",
        "jobs_log_h_width_log_blowup": [list(s) for s in shapes], "roots": roots})


def served_mix():
    """Shard commit jobs as the service receives them: KoalaBear and BabyBear,
    log_blowup 1 or 3, widths 16..512, heights 2^16..2^18 plus one 2^20 job;
    job k seeded (5 << 32) | (20000 + k)."""
    from pico_amd.workloads import SplitMix64
    r = SplitMix64((5 << 32) | 77)
    jobs = []
    for _ in range(9):
        L = r.uniform_int(16, 18)
        w = r.uniform_int(16, 512)
        f = KB if r.next() & 1 else BB
        b = 1 if r.next() & 1 else 3
        jobs.append((f, L, w, b))
    jobs.insert(4, (BB, 20, 200, 3))
    return jobs


def served():
    t0 = time.time()
    jobs = served_mix()
    roots = []
    for k, (f, L, w, b) in enumerate(jobs):
        roots.append(_ints(oracle.pcs_commit_seeded(f, [(5 << 32) | (20000 + k)], [1 << L], [w], b)))
        print(f"  served job {k}: field {f} 2^{L} x {w} b={b} ({time.time() - t0:.0f} s)", flush=True)
    _write("served_mix.json", {
        "_source": "oracle/oracle.c (or_pcs_commit_seeded); tests/golden/make_golden.py served",
        "seed_rule": "(5 << 32) | (20000 + k)", "fields": {"0": "koalabear", "1": "babybear"},
        "jobs_field_log_h_width_log_blowup": [list(j) for j in jobs], "roots": roots})


def shard():
    """One shard: the 14 main traces (log_h 14..22, widths 8..300, matrix i
    seeded (3 << 32) | (100 + i)) in one mixed-height commit at log_blowup 1,
    and a 2^21-row ext4 quotient (seed (3 << 32) | 200, flattened to 4 base
    columns) split into 4 chunks (rows i, i + 4, ...) on the domains
    GEN w_{2^21}^i <w_{2^19}> (vm/src/machine/prover.rs:513-523), committed
    at log_blowup 1."""
    from pico_amd import workloads
    t0 = time.time()
    f, lb = KB, 1
    shapes = workloads.shard_shapes()
    seeds = [(3 << 32) | (100 + i) for i in range(len(shapes))]
    main = oracle.pcs_commit_seeded(f, seeds, [1 << L for L, _ in shapes], [w for _, w in shapes], lb)
    print(f"  shard main commit ({time.time() - t0:.0f} s)", flush=True)
    p = oracle.PRIMES[f]
    g = oracle.generator(f)
    q = oracle.random_matrix(f, (3 << 32) | 200, 4 << 19, 4)
    w21 = oracle.two_adic_generator(f, 21)
    ldes = []
    for i in range(4):
        s = g * pow(w21, i, p) % p
        ldes.append(oracle.coset_lde(f, np.ascontiguousarray(q[i::4]), lb, g * pow(s, p - 2, p) % p))
    quot = oracle.MerkleTree(f, ldes).root
    _write("shard_kb_14.json", {
        "_source": "oracle/oracle.c (or_pcs_commit_seeded; or_coset_lde + or_mmcs_commit for the quotient "
                   "chunks); tests/golden/make_golden.py shard",
        "field": "koalabear", "log_blowup": lb, "seed_rule": "(3 << 32) | (100 + i)",
        "main_log_h_width": [list(s) for s in shapes], "main_root": _ints(main),
        "quotient": {"seed": (3 << 32) | 200, "rows": 4 << 19, "cols": 4, "chunks": 4},
        "quotient_root": _ints(quot)})
    print(f"shard: {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    todo = ["headline", "served", "batched", "shard"] if what == ["all"] else what
    for name in todo:
        globals()[name]()
