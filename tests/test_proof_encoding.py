"""The FriProof encoding of the C ABI (pg_open_proof_encode / _decode) on the
host: bytes written by an independent Python encoder from the reference's
struct layout (tests/pyref.py encode_fri_proof) decode to the same parts and
encode back to the same bytes, in Montgomery and canonical words; malformed
encodings are rejected.  The GPU tests check the encoding of real
Pcs::open proofs (tests/test_gpu_open.py)."""
import numpy as np
import pytest

import pico_amd
from pico_amd import BABYBEAR, KOALABEAR, PgError
from tests import pyref


def synthetic_proof(f, seed, widths=((3, 5), (4,)), log_h=(6, 5), fri_levels=(5, 4, 3, 2), nq=3):
    rng = np.random.default_rng(seed)
    p = pyref.P[f]

    def vals(*shape):
        return rng.integers(0, p, shape, dtype=np.uint64)

    commits = vals(len(fri_levels), 8)
    queries = []
    for _ in range(nq):
        inp = [([vals(w) for w in ws], vals(lh, 8)) for ws, lh in zip(widths, log_h)]
        cp = [(vals(4), vals(lv, 8)) for lv in fri_levels]
        queries.append((inp, cp))
    return commits, queries, vals(4), int(rng.integers(0, p))


@pytest.mark.parametrize("f", [KOALABEAR, BABYBEAR])
@pytest.mark.parametrize("canonical", [False, True])
def test_decode_then_encode_matches_python_encoder(f, canonical):
    commits, queries, final, wit = synthetic_proof(f, 10 + f)
    data = pyref.encode_fri_proof(f, commits, queries, final, wit, canonical)
    d = pico_amd.decode_fri_proof(f, data, canonical=canonical)
    assert d["reencoded"] == pyref.encode_fri_proof(f, commits, queries, final, wit, False)
    assert np.array_equal(d["commits"], commits)
    assert np.array_equal(d["final_poly"], final)
    assert d["pow_witness"] == wit
    assert d["num_queries"] == len(queries)
    for q, (inp, cp) in enumerate(queries):
        flat = np.concatenate([np.concatenate([*rows, path.ravel()]) for rows, path in inp])
        assert np.array_equal(d["inputs_flat"][q], flat)
        assert np.array_equal(d["fri_siblings"][q], np.stack([s for s, _ in cp]))
        assert np.array_equal(d["fri_paths_flat"][q], np.concatenate([pth.ravel() for _, pth in cp]))


def test_empty_shapes_roundtrip():
    # no FRI rounds, one round of one zero-width matrix of height 1, one query
    f = KOALABEAR
    commits, queries, final, wit = synthetic_proof(f, 3, widths=((0,),), log_h=(0,), fri_levels=(), nq=1)
    data = pyref.encode_fri_proof(f, commits, queries, final, wit, False)
    assert pico_amd.decode_fri_proof(f, data)["reencoded"] == data
    data = pyref.encode_fri_proof(f, commits, [], final, wit, False)  # no queries
    assert pico_amd.decode_fri_proof(f, data)["reencoded"] == data


def test_malformed_encodings_are_rejected():
    f = KOALABEAR
    commits, queries, final, wit = synthetic_proof(f, 4)
    data = pyref.encode_fri_proof(f, commits, queries, final, wit, False)
    for bad in (data[:-1], data + b"\0", data[:8]):
        with pytest.raises(PgError):
            pico_amd.decode_fri_proof(f, bad)
    # a word that is not a field element
    with pytest.raises(PgError):
        pico_amd.decode_fri_proof(f, data[:-4] + (0xFFFFFFFF).to_bytes(4, "little"))
    # a declared length far beyond the data
    with pytest.raises(PgError):
        pico_amd.decode_fri_proof(f, (1 << 40).to_bytes(8, "little") + data[8:])
    # the queries of one proof must share one shape
    q2 = queries[:1] + [synthetic_proof(f, 5, widths=((3, 6), (4,)))[1][0]]
    with pytest.raises(PgError):
        pico_amd.decode_fri_proof(f, pyref.encode_fri_proof(f, commits, q2, final, wit, False))
