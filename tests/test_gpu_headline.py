"""The benchmark's workload at full size against the oracle's golden values:
Pcs::commit (vm/src/machine/prover.rs:238-244; parameters
vm/src/configs/stark_config/kb_poseidon2.rs:60-93) of a KoalaBear 2^20 x 256
trace drawn on the device from seed (2 << 32) | 0, at log_blowup 1.  The root
and 320 LDE samples (committed row order) were computed by the CPU oracle
(tests/golden/make_golden.py headline)."""
import json
import os

import numpy as np
import pytest

import pico_amd
from pico_amd import KOALABEAR

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "headline_kb_2p20x256.json")


def test_headline_commit_matches_golden(ctx):
    with open(GOLDEN) as fh:
        g = json.load(fh)
    f = KOALABEAR
    print(f"headline commit {g['rows']}x{g['cols']} seed {g['seed']:#x}", flush=True)
    trace = ctx.random_matrix(f, g["seed"], g["rows"], g["cols"])
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, g["log_blowup"])
    root, pd = pcs.commit([(pcs.natural_domain_for_degree(g["rows"]), trace)])
    try:
        assert [int(x) for x in root] == g["root"]
        lde = pico_amd.MerkleTreeMmcs(ctx, f).get_matrices(pd)[0]
        assert (lde.rows, lde.cols) == (g["rows"] << g["log_blowup"], g["cols"])
        by_row = {}
        for r, c, v in g["lde_samples_row_col_value"]:
            by_row.setdefault(r, []).append((c, v))
        for r, cv in by_row.items():
            row = lde.to_host(f, row0=r, nrows=1)[0]
            for c, v in cv:
                assert int(row[c]) == v, (r, c)
        # the seeded trace itself: the device fill draws the oracle's stream
        head = trace.to_host(f, row0=0, nrows=2)
        from oracle import oracle
        assert np.array_equal(head, oracle.random_matrix(f, g["seed"], 2, g["cols"]))
    finally:
        pd.free()
        trace.free()
