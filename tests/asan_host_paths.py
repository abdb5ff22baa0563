"""The C ABI's host-only paths (challenger, verify_batch, the tile-shape
table parser, argument errors) in a process whose library was built with
AddressSanitizer and UBSan on its host code (make -C pico_amd asan).  Run by
tests/test_host_sanitized.py with the sanitizer runtime preloaded; no GPU is
touched.  usage: python3 asan_host_paths.py LIBRARY"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import pico_amd._lib as L
L._lib = None
L.LIB_PATH = sys.argv[1]
L.lib()
import numpy as np
import pico_amd
from pico_amd import KOALABEAR, BABYBEAR
from oracle import oracle
# challenger
for f in (KOALABEAR, BABYBEAR):
    ch = pico_amd.DuplexChallenger(f)
    for i in range(50):
        ch.observe(list(range(i % 11)))
        ch.sample_ext(); ch.sample_bits(17)
    ch.check_witness(5, 1234)
# verify_batch on oracle trees, good and corrupted
f = KOALABEAR
mats = [oracle.random_matrix(f, 7 + i, 1 << L, w) for i, (L, w) in enumerate([(5, 9), (4, 3), (5, 8), (3, 17)])]
tree = oracle.MerkleTree(f, mats)
mm = pico_amd.MerkleTreeMmcs.__new__(pico_amd.MerkleTreeMmcs); mm.field = f
dims = [(m.shape[1], m.shape[0]) for m in mats]
for idx in range(32):
    rows, path = tree.open(idx)
    assert mm.verify_batch(tree.root, dims, idx, rows, path)
    bad = [r.copy() for r in rows]; bad[1][0] ^= 1
    assert not mm.verify_batch(tree.root, dims, idx, bad, path)
# tuning table parsing
t = pico_amd.tune_export()
pico_amd.tune_import(t)
for bad in ["0 1 8 0 2\n", "junk\n", "0 1 8 0 1 9\n", "1 2 10 3 1\n" * 1000]:
    try:
        pico_amd.tune_import(bad)
    except pico_amd.PgError:
        pass
buf_needed = pico_amd.tune_export()
# argument errors without a device
lib = L.lib()
assert lib.pg_buf_view(None, None, 0, 1, 0, 1, 1, None) != 0
assert lib.pg_pcs_commit_host(None, 0, None, None, None, None, 0, 1, 0, None, None) != 0
assert lib.pg_pcs_open(None, None, 0, None, None, 1, -1, 0, 0, None, None) != 0
assert lib.pg_autotune(None, 0, 1, 2, None, 0, 1, 1, None, 0) != 0
print(lib.pg_last_error().decode() if lib.pg_last_error() else "")
print("host paths ok")
