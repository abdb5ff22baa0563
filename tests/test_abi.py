"""C-ABI checks that need no GPU: the library loads, exports every function
include/pico_gpu.h declares, and its host-only entry points (the challenger
and verify_batch) agree with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

import pico_amd
from pico_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "pico_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), n
    # and the Python binding covers exactly the declared set
    assert sorted(_lib.SIGNATURES) == names


def test_no_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.lib()


@pytest.mark.parametrize("f", [0, 1])
def test_host_challenger_matches_oracle(oracle, f):
    ch = pico_amd.DuplexChallenger(f)
    oc = oracle.Challenger(f)
    rng = np.random.default_rng(7)
    p = pico_amd.PRIMES[f]
    for _ in range(80):
        op = rng.integers(0, 4)
        if op == 0:
            v = int(rng.integers(0, p))
            ch.observe(v)
            oc.observe(v)
        elif op == 1:
            assert ch.sample() == oc.sample()
        elif op == 2:
            assert list(ch.sample_ext()) == list(oc.sample_ext())
        else:
            assert ch.sample_bits(16) == oc.sample_bits(16)
    c1, c2 = ch.clone(), oc.clone()
    for w in range(40):
        a, b = c1.clone(), c2.clone()
        assert a.check_witness(3, w) == b.check_witness(3, w)


@pytest.mark.parametrize("f", [0, 1])
def test_host_challenger_long_stream_matches_oracle(oracle, f):
    """About 1300 host permutations over states the observed values make
    random: every sample equals the oracle's (the host permutation runs in
    AVX2 where the CPU has it)."""
    ch = pico_amd.DuplexChallenger(f)
    oc = oracle.Challenger(f)
    rng = np.random.default_rng(17 + f)
    p = pico_amd.PRIMES[f]
    for step in range(200):
        for v in rng.integers(0, p, int(rng.integers(1, 60))):
            ch.observe(int(v))
            oc.observe(int(v))
        assert list(ch.sample_ext()) == list(oc.sample_ext()), step


@pytest.mark.parametrize("f", [0, 1])
def test_host_verify_batch_matches_oracle(oracle, f):
    shape = [(8, 3), (8, 9), (4, 8), (1, 2)]
    mats = [oracle.random_matrix(f, 300 + i, h, w) for i, (h, w) in enumerate(shape)]
    tree = oracle.MerkleTree(f, mats)
    mm = pico_amd.MerkleTreeMmcs(None, f)
    dims = [(w, h) for h, w in shape]
    for idx in range(8):
        rows, path = tree.open(idx)
        assert mm.verify_batch(tree.root, dims, idx, rows, path)
        bad = [r.copy() for r in rows]
        bad[1][4] = (int(bad[1][4]) + 1) % pico_amd.PRIMES[f]
        assert not mm.verify_batch(tree.root, dims, idx, bad, path)
        assert not mm.verify_batch(tree.root, dims, idx ^ 1, rows, path)


def test_invalid_arguments_fail_loudly():
    # null context: the ABI rejects it with PG_E_INVALID and a message
    rc = _lib.lib().pg_buf_alloc(None, 4, 4, ctypes.byref(ctypes.c_void_p()))
    assert rc == _lib.PG_E_INVALID
    assert b"null" in _lib.lib().pg_last_error()
    with pytest.raises(pico_amd.PgError):
        _lib.check(rc)
