"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  All arrays are
canonical field elements (uint32 in [0, p)).
"""
import ctypes
import os

import numpy as np

KOALABEAR = 0
BABYBEAR = 1
PRIMES = {KOALABEAR: 2130706433, BABYBEAR: 2013265921}

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compile liboracle.so in place (make)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        u32, u64, sz, i = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int
        P = ctypes.c_void_p
        sig = {
            "or_prime": (u32, [i]), "or_add": (u32, [i, u32, u32]),
            "or_sub": (u32, [i, u32, u32]), "or_mul": (u32, [i, u32, u32]),
            "or_pow": (u32, [i, u32, u64]), "or_inv": (u32, [i, u32]),
            "or_generator": (u32, [i]), "or_two_adic_generator": (u32, [i, i]),
            "or_two_adicity": (i, [i]), "or_to_monty": (u32, [i, u32]),
            "or_from_monty": (u32, [i, u32]),
            "or_ext_mul": (None, [i, P, P, P]), "or_ext_inv": (None, [i, P, P]),
            "or_poseidon2": (None, [i, P]), "or_sponge": (None, [i, P, sz, P]),
            "or_compress": (None, [i, P, P, P]),
            "or_fill_random": (None, [i, u64, P, sz]),
            "or_coset_lde": (None, [i, P, sz, sz, i, u32, i, P]),
            "or_coset_lde_naive": (None, [i, P, sz, sz, i, u32, i, P]),
            "or_mmcs_digest_words": (sz, [P, sz]),
            "or_mmcs_commit": (i, [i, P, sz, P, P]),
            "or_mmcs_open": (i, [P, sz, P, sz, P, P]),
            "or_mmcs_verify": (i, [i, P, P, sz, P, sz, P, P]),
            "or_fri_fold": (None, [i, P, sz, P, P]),
            "or_ch_init": (None, [P, i]), "or_ch_observe": (None, [P, u32]),
            "or_ch_sample": (u32, [P]), "or_ch_sample_ext": (None, [P, P]),
            "or_ch_sample_bits": (u32, [P, i]), "or_ch_check_witness": (i, [P, i, u32]),
            "or_fri_commit_phase": (i, [P, P, i, i, P, P, P, P, P]),
            "or_num_threads": (i, []),
            "or_pcs_open_values": (None, [i, P, sz, sz, i, P, P]),
            "or_pcs_reduce": (None, [i, sz, P, P, P, P, P, P, P, P]),
            "or_pcs_commit_seeded": (i, [i, sz, P, P, P, P, i, sz, P]),
            "cb_pcs_commit": (i, [i, P, sz, sz, i, P, P]),
            "cb_pcs_commit_isa": (i, [i, i, P, sz, sz, i, P, P]),
            "cb_isa": (i, []),
            "cb_to_monty": (None, [i, P, P, sz]),
            "cb_num_threads": (i, []),
            "cb_last_phase_seconds": (None, [P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ field --
def prime(field):
    return PRIMES[field]


def generator(field):
    return lib().or_generator(field)


def two_adic_generator(field, log_n):
    return lib().or_two_adic_generator(field, log_n)


def to_monty(field, a):
    p = PRIMES[field]
    return ((np.asarray(a, dtype=np.uint64) << np.uint64(32)) % np.uint64(p)).astype(np.uint32)


def from_monty(field, a):
    p = PRIMES[field]
    rinv = pow(2 ** 32, p - 2, p)
    return ((np.asarray(a, dtype=np.uint64) * np.uint64(rinv)) % np.uint64(p)).astype(np.uint32)


def ext_mul(field, a, b):
    a, b, out = _u32(a), _u32(b), np.zeros(4, np.uint32)
    lib().or_ext_mul(field, _ptr(a), _ptr(b), _ptr(out))
    return out


def ext_inv(field, a):
    a, out = _u32(a), np.zeros(4, np.uint32)
    lib().or_ext_inv(field, _ptr(a), _ptr(out))
    return out


# -------------------------------------------------------------- poseidon2 --
def poseidon2(field, state):
    s = _u32(state).copy()
    assert s.shape == (16,)
    lib().or_poseidon2(field, _ptr(s))
    return s


def sponge(field, values):
    v = _u32(values).ravel()
    out = np.zeros(8, np.uint32)
    lib().or_sponge(field, _ptr(v), v.size, _ptr(out))
    return out


def compress(field, left, right):
    l, r, out = _u32(left), _u32(right), np.zeros(8, np.uint32)
    lib().or_compress(field, _ptr(l), _ptr(r), _ptr(out))
    return out


def random_matrix(field, seed, h, w):
    out = np.empty((h, w), np.uint32)
    lib().or_fill_random(field, seed, _ptr(out), out.size)
    return out


# -------------------------------------------------------------------- LDE --
def coset_lde(field, mat, added_bits, shift, bit_reverse=True, naive=False):
    m = _u32(mat)
    h, w = m.shape
    out = np.empty((h << added_bits, w), np.uint32)
    fn = lib().or_coset_lde_naive if naive else lib().or_coset_lde
    fn(field, _ptr(m), h, w, added_bits, shift, int(bit_reverse), _ptr(out))
    return out


# ------------------------------------------------------------------- MMCS --
class _Mat(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("height", ctypes.c_size_t),
                ("width", ctypes.c_size_t)]


class MerkleTree:
    """Oracle commitment: keeps the matrices and every digest layer."""

    def __init__(self, field, mats):
        self.field = field
        self.mats = [_u32(m) for m in mats]
        self._c = (_Mat * len(self.mats))(*[
            _Mat(_ptr(m).value, m.shape[0], m.shape[1]) for m in self.mats])
        words = lib().or_mmcs_digest_words(self._c, len(self.mats))
        if words == 0:
            raise ValueError("heights must be non-zero powers of two")
        self.digests = np.zeros(words, np.uint32)
        self.root = np.zeros(8, np.uint32)
        rc = lib().or_mmcs_commit(field, self._c, len(self.mats), _ptr(self.digests),
                                  _ptr(self.root))
        assert rc == 0
        self.max_height = max(m.shape[0] for m in self.mats)
        self.log_max = self.max_height.bit_length() - 1

    def layer(self, j):
        off = sum(self.max_height >> t for t in range(j))
        n = self.max_height >> j
        return self.digests[8 * off: 8 * (off + n)].reshape(n, 8)

    def open(self, index):
        wtot = sum(m.shape[1] for m in self.mats)
        rows = np.zeros(max(wtot, 1), np.uint32)
        path = np.zeros(8 * max(self.log_max, 1), np.uint32)
        rc = lib().or_mmcs_open(self._c, len(self.mats), _ptr(self.digests), index,
                                _ptr(rows), _ptr(path))
        assert rc == 0
        out, pos = [], 0
        for m in self.mats:
            out.append(rows[pos:pos + m.shape[1]].copy())
            pos += m.shape[1]
        return out, path[:8 * self.log_max].reshape(self.log_max, 8)


def mmcs_verify(field, heights, widths, root, index, rows, path):
    hs = np.asarray(heights, np.uint64).astype(np.uintp)
    ws = np.asarray(widths, np.uint64).astype(np.uintp)
    r = _u32(np.concatenate([np.asarray(x, np.uint32).ravel() for x in rows])
             if len(rows) else np.zeros(1, np.uint32))
    pth = _u32(path).ravel()
    if pth.size == 0:
        pth = np.zeros(8, np.uint32)
    rt = _u32(root)
    return bool(lib().or_mmcs_verify(field, _ptr(hs), _ptr(ws), len(heights), _ptr(rt), index,
                                     _ptr(r), _ptr(pth)))


def pcs_commit_seeded(field, seeds, heights, widths, log_blowup, domain_shifts=None, block_cols=64):
    """Root of TwoAdicFriPcs::commit over seeded matrices (random_matrix(field,
    seeds[k], heights[k], widths[k]) on domains domain_shifts[k] <w_h>,
    canonical shifts, default 1), without holding any LDE whole."""
    n = len(seeds)
    sd = np.ascontiguousarray(seeds, dtype=np.uint64)
    hs = np.ascontiguousarray(heights, dtype=np.uint64).astype(np.uintp)
    ws = np.ascontiguousarray(widths, dtype=np.uint64).astype(np.uintp)
    sh = _u32(np.ones(n, np.uint32) if domain_shifts is None else domain_shifts)
    root = np.zeros(8, np.uint32)
    rc = lib().or_pcs_commit_seeded(field, n, _ptr(sd), _ptr(hs), _ptr(ws), _ptr(sh), log_blowup,
                                    block_cols, _ptr(root))
    if rc != 0:
        raise ValueError("heights must be powers of two")
    return root


# -------------------------------------------------------------------- FRI --
def pcs_open_values(field, lde, log_blowup, z):
    """Values (w x 4) of a committed LDE (committed row order) at ext4 z."""
    a = _u32(lde)
    n, w = a.shape
    out = np.zeros((w, 4), np.uint32)
    lib().or_pcs_open_values(field, _ptr(a), n, w, log_blowup, _ptr(_u32(z)), _ptr(out))
    return out


def pcs_reduce(field, mats, npts, points, values, alpha):
    """Reduced openings.  mats: committed LDEs in opening order; npts[m]
    points for matrix m; points: (sum npts) x 4; values: per (matrix, point)
    a w_m x 4 array, in order.  Returns {log_height: (2^lh x 4) array}."""
    mats = [_u32(m) for m in mats]
    hs = np.array([m.shape[0] for m in mats], np.uintp)
    ws = np.array([m.shape[1] for m in mats], np.uintp)
    ptrs = (ctypes.c_void_p * len(mats))(*[_ptr(m).value for m in mats])
    np_ = np.ascontiguousarray(npts, dtype=np.uint32)
    pts = _u32(np.asarray(points).reshape(-1, 4)) if len(points) else np.zeros((1, 4), np.uint32)
    vals = _u32(np.concatenate([np.asarray(v).reshape(-1, 4) for v in values])) if values else np.zeros((1, 4), np.uint32)
    outs = {}
    arr = (ctypes.c_void_p * 64)()
    for m, k in zip(mats, np_):
        lh = m.shape[0].bit_length() - 1
        if k and lh not in outs:
            outs[lh] = np.zeros((m.shape[0], 4), np.uint32)
            arr[lh] = _ptr(outs[lh]).value
    lib().or_pcs_reduce(field, len(mats), ptrs, _ptr(hs), _ptr(ws), _ptr(np_), _ptr(pts), _ptr(vals),
                        _ptr(_u32(alpha)), arr)
    return outs


def fri_fold(field, v, beta):
    v = _u32(v).reshape(-1, 4)
    n = v.shape[0]
    out = np.zeros((n // 2, 4), np.uint32)
    b = _u32(beta)
    lib().or_fri_fold(field, _ptr(v), n, _ptr(b), _ptr(out))
    return out


class _Ch(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int), ("state", ctypes.c_uint32 * 16),
                ("inp", ctypes.c_uint32 * 8), ("n_in", ctypes.c_int),
                ("out", ctypes.c_uint32 * 8), ("n_out", ctypes.c_int)]


class Challenger:
    def __init__(self, field):
        self._c = _Ch()
        self.field = field
        lib().or_ch_init(ctypes.byref(self._c), field)

    def observe(self, v):
        for x in np.atleast_1d(np.asarray(v, np.uint32)).ravel():
            lib().or_ch_observe(ctypes.byref(self._c), int(x))

    def sample(self):
        return lib().or_ch_sample(ctypes.byref(self._c))

    def sample_ext(self):
        out = np.zeros(4, np.uint32)
        lib().or_ch_sample_ext(ctypes.byref(self._c), _ptr(out))
        return out

    def sample_bits(self, bits):
        return lib().or_ch_sample_bits(ctypes.byref(self._c), bits)

    def check_witness(self, bits, witness):
        return bool(lib().or_ch_check_witness(ctypes.byref(self._c), bits, witness))

    def clone(self):
        c = Challenger(self.field)
        ctypes.memmove(ctypes.byref(c._c), ctypes.byref(self._c), ctypes.sizeof(_Ch))
        return c

    def state(self):
        return np.array(self._c.state, np.uint32)


def fri_commit_phase(ch, v, log_stop, inputs=None):
    """Returns (roots[r,8], layers (list of committed vectors), final[.,4], betas[r,4])."""
    v = _u32(v).reshape(-1, 4)
    log_n = v.shape[0].bit_length() - 1
    rounds = log_n - log_stop
    roots = np.zeros((max(rounds, 1), 8), np.uint32)
    betas = np.zeros((max(rounds, 1), 4), np.uint32)
    total = sum((1 << (log_n - r)) for r in range(rounds))
    layers = np.zeros((max(total, 1), 4), np.uint32)
    final = np.zeros((1 << log_stop, 4), np.uint32)
    keep = []
    arr = (ctypes.c_void_p * (log_n + 1))()
    if inputs:
        for lg, vec in inputs.items():
            a = _u32(vec)
            keep.append(a)
            arr[lg] = _ptr(a).value
    r = lib().or_fri_commit_phase(ctypes.byref(ch._c), _ptr(v), log_n, log_stop,
                                  arr if inputs else None, _ptr(roots), _ptr(layers),
                                  _ptr(final), _ptr(betas))
    assert r == rounds
    out_layers, off = [], 0
    for rr in range(rounds):
        n = 1 << (log_n - rr)
        out_layers.append(layers[off:off + n].copy())
        off += n
    return roots[:rounds], out_layers, final, betas[:rounds]


def num_threads():
    return lib().or_num_threads()


# --------------------------------------------------- CPU baseline (timing) --
# cpu_baseline.c: the reference's CPU commit path (Radix2DitParallel LDE +
# MerkleTreeMmcs, Montgomery form, AVX-512 or AVX2, OpenMP), timed by
# bench.py and pinned to the functions above by tests/test_oracle.py.
CPU_ISAS = {1: "avx2", 2: "avx512"}


def cpu_isa():
    """The instruction set the timed commit uses on this host: 2 (AVX-512) or 1 (AVX2)."""
    return lib().cb_isa()
def cpu_to_monty(field, mat):
    """Canonical matrix -> Montgomery words (the input form the timed commit takes)."""
    a = _u32(mat)
    out = np.empty_like(a)
    lib().cb_to_monty(field, _ptr(a), _ptr(out), a.size)
    return out


def cpu_pcs_commit_mont(field, mont, log_blowup, want_lde=False, isa=0):
    """TwoAdicFriPcs::commit of one trace held in Montgomery words:
    (root, LDE or None), both Montgomery, the LDE in committed row order.
    isa 0: the best this host has; 1 AVX2; 2 AVX-512."""
    mont = _u32(mont)
    h, w = mont.shape
    root = np.zeros(8, np.uint32)
    lde = np.empty((h << log_blowup, w), np.uint32) if want_lde else None
    rc = lib().cb_pcs_commit_isa(isa, field, _ptr(mont), h, w, log_blowup, _ptr(root),
                                 _ptr(lde) if want_lde else None)
    if rc != 0:
        raise MemoryError("cb_pcs_commit failed (shape or allocation)")
    return root, lde


def cpu_pcs_commit(field, mat, log_blowup, want_lde=False, isa=0):
    """The same from canonical values; canonical outputs."""
    root, lde = cpu_pcs_commit_mont(field, cpu_to_monty(field, mat), log_blowup, want_lde, isa)
    return from_monty(field, root), (from_monty(field, lde) if want_lde else None)


def cpu_num_threads():
    return lib().cb_num_threads()


def cpu_last_phase_seconds():
    """(LDE seconds, Merkle seconds) of the last cpu_pcs_commit* call."""
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().cb_last_phase_seconds(ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


# ---------------------------------------------------------- quotient values --
# compute_quotient_values (vm/src/machine/utils.rs:201-356) restated over the
# expression DAG directly, vectorised over the rows of the quotient coset
# Q = GEN <w_Q> in natural order; the fold acc = acc * alpha + c follows
# vm/src/machine/folder.rs:222-226 (assert_zero and assert_zero_ext alike)
# and the selectors vm/src/compiler/recursion/circuit/domain.rs:46-60.  Node
# kinds as in the C ABI (pg_air_node): 0 main, 1 preprocessed (a = row
# offset, b = column), 2 public (b), 3 is_first_row, 4 is_last_row,
# 5 is_transition, 6 constant (a), 7 add, 8 sub, 9 neg (a), 10 mul (a, b =
# earlier nodes), 11 permutation (a = row offset, b = ext4 column: base
# columns 4b..4b+3), 12 permutation challenge (b), 13 regional cumulative
# sum, 14 global cumulative sum coordinate (b).  Every value is held as an
# element of F[X]/(X^4 - W) (base values with zero higher coordinates), so
# sums, products and the fold are extension arithmetic throughout.
_W = {KOALABEAR: 3, BABYBEAR: 11}


def _ext_mul_rows(f, a, b):
    """a (n x 4) times b (n x 4 or 4), mod p, in F[x]/(x^4 - W)."""
    p = np.uint64(PRIMES[f])
    w = np.uint64(_W[f])
    a = a.astype(np.uint64)
    b = np.broadcast_to(np.asarray(b, np.uint64), a.shape)
    c = [np.zeros(a.shape[0], np.uint64) for _ in range(7)]
    for i in range(4):
        for j in range(4):
            c[i + j] = (c[i + j] + a[:, i] * b[:, j] % p) % p
    out = np.empty_like(a)
    for k in range(4):
        out[:, k] = (c[k] + (w * c[k + 4] % p if k + 4 < 7 else 0)) % p
    return out


def quotient_values(field, nodes, asserts, main_q, log_n, log_qd, alpha, publics=(), pre_q=None, perm_q=None,
                    challenges=(), regional_sum=None, global_sum=()):
    """nodes (k x 3), asserts: the DAG; main_q / pre_q / perm_q: the traces'
    evaluations on Q, natural order (|Q| x width, canonical; perm_q has the
    permutation trace's base columns, 4 per ext4 column); challenges: ext4
    each; regional_sum: ext4; global_sum: base values.  Returns the |Q| x 4
    quotient values (canonical)."""
    p = PRIMES[field]
    P = np.uint64(p)
    q = main_q.shape[0]
    assert q == 1 << (log_n + log_qd)
    step = 1 << log_qd
    g = generator(field)
    wq = two_adic_generator(field, log_n + log_qd)
    xs = np.empty(q, np.uint64)
    cur = g
    for i in range(q):
        xs[i] = cur
        cur = cur * wq % p
    n = 1 << log_n
    zh = np.array([(pow(int(x), n, p) - 1) % p for x in xs], np.uint64)
    inv = np.vectorize(lambda v: pow(int(v), p - 2, p), otypes=[np.uint64])
    w_last = pow(two_adic_generator(field, log_n), p - 2, p)
    trans = (xs + P - np.uint64(w_last)) % P
    first = zh * inv((xs + P - np.uint64(1)) % P) % P
    last = zh * inv(trans) % P
    nxt = (np.arange(q) + step) % q
    mains = [main_q.astype(np.uint64), main_q[nxt].astype(np.uint64)]
    pres = [pre_q.astype(np.uint64), pre_q[nxt].astype(np.uint64)] if pre_q is not None else None
    perms = [perm_q.astype(np.uint64), perm_q[nxt].astype(np.uint64)] if perm_q is not None else None

    def base(v):  # (q,) base values -> (q, 4)
        out = np.zeros((q, 4), np.uint64)
        out[:, 0] = v
        return out

    def const(c):  # one ext4 (or base) value for every row
        out = np.zeros((q, 4), np.uint64)
        out[:, :] = np.asarray([int(x) % p for x in (c if np.ndim(c) else [c, 0, 0, 0])], np.uint64)
        return out

    vals = []
    for kind, a, b in np.asarray(nodes, np.uint64).reshape(-1, 3).tolist():
        if kind == 0:
            v = base(mains[a][:, b])
        elif kind == 1:
            v = base(pres[a][:, b])
        elif kind == 2:
            v = const(publics[b])
        elif kind == 3:
            v = base(first)
        elif kind == 4:
            v = base(last)
        elif kind == 5:
            v = base(trans)
        elif kind == 6:
            v = const(a)
        elif kind == 7:
            v = (vals[a] + vals[b]) % P
        elif kind == 8:
            v = (vals[a] + P - vals[b]) % P
        elif kind == 9:
            v = (P - vals[a]) % P
        elif kind == 10:
            v = _ext_mul_rows(field, vals[a], vals[b])
        elif kind == 11:
            v = perms[a][:, 4 * b:4 * b + 4] % P
        elif kind == 12:
            v = const(challenges[b])
        elif kind == 13:
            v = const(regional_sum)
        elif kind == 14:
            v = const(global_sum[b])
        else:
            raise ValueError(kind)
        vals.append(v)
    acc = np.zeros((q, 4), np.uint64)
    for k in asserts:
        acc = (_ext_mul_rows(field, acc, alpha) + vals[k]) % P
    return (acc * inv(zh)[:, None] % P).astype(np.uint32)


def constraints_at_point(field, nodes, asserts, main_local, main_next, zeta, log_n, alpha, publics=(),
                         pre_local=None, pre_next=None, perm_local=None, perm_next=None, challenges=(),
                         regional_sum=None, global_sum=()):
    """The verifier's side at one ext4 point (vm/src/machine/verifier.rs:
    307-356): the folded constraints and 1 / Z_H(zeta), from the trace
    polynomials' values at zeta and zeta w_n (ext4 per column; for the
    permutation trace, per base column of its flattened form, an ext4 column
    being sum_e X^e v_{4c+e})."""
    p = PRIMES[field]

    def mul(a, b):
        return [int(x) for x in ext_mul(field, a, b)]

    def add(a, b):
        return [(x + y) % p for x, y in zip(a, b)]

    def sub(a, b):
        return [(x - y) % p for x, y in zip(a, b)]

    def base(v):
        return [int(v) % p, 0, 0, 0]

    def ext_col(vals4):  # sum_e X^e v_e
        acc = [0, 0, 0, 0]
        for e, v in enumerate(vals4):
            mono = [0, 0, 0, 0]
            mono[e] = 1
            acc = add(acc, mul(mono, [int(x) for x in v]))
        return acc

    z = [int(x) for x in zeta]
    zn = z
    for _ in range(log_n):
        zn = mul(zn, zn)
    zh = sub(zn, base(1))
    w_last = pow(two_adic_generator(field, log_n), p - 2, p)
    trans = sub(z, base(w_last))
    first = mul(zh, [int(x) for x in ext_inv(field, sub(z, base(1)))])
    last = mul(zh, [int(x) for x in ext_inv(field, trans)])
    mains = [main_local, main_next]
    pres = [pre_local, pre_next]
    perms = [perm_local, perm_next]
    vals = []
    for kind, a, b in np.asarray(nodes, np.uint64).reshape(-1, 3).tolist():
        if kind == 0:
            v = [int(x) for x in mains[a][b]]
        elif kind == 1:
            v = [int(x) for x in pres[a][b]]
        elif kind == 2:
            v = base(publics[b])
        elif kind == 3:
            v = first
        elif kind == 4:
            v = last
        elif kind == 5:
            v = trans
        elif kind == 6:
            v = base(a)
        elif kind == 7:
            v = add(vals[a], vals[b])
        elif kind == 8:
            v = sub(vals[a], vals[b])
        elif kind == 9:
            v = sub(base(0), vals[a])
        elif kind == 10:
            v = mul(vals[a], vals[b])
        elif kind == 11:
            v = ext_col(perms[a][4 * b:4 * b + 4])
        elif kind == 12:
            v = [int(x) % p for x in challenges[b]]
        elif kind == 13:
            v = [int(x) % p for x in regional_sum]
        elif kind == 14:
            v = base(global_sum[b])
        else:
            raise ValueError(kind)
        vals.append(v)
    acc = base(0)
    for k in asserts:
        acc = add(mul(acc, [int(x) for x in alpha]), vals[k])
    return acc, [int(x) for x in ext_inv(field, zh)]
