"""Python mirror of the Plonky3 prover interfaces that Pico's StarkGenericConfig
plugs in (vm/src/configs/stark_config/kb_poseidon2.rs:15-27, 82-93), backed by
the gfx950 HIP library through its C ABI.

Names follow the reference's types: ``Radix2DitParallel.coset_lde_batch``,
``MerkleTreeMmcs.commit / open_batch / get_matrices / verify_batch``,
``TwoAdicFriPcs.commit / get_evaluations_on_domain``, the FRI commit phase and
``DuplexChallenger``.  Field elements cross this Python API as canonical
integers (numpy uint32 in [0, p)); the conversion to the Montgomery form the
device and the ABI use happens at the boundary.  Errors raise ``PgError``
where the reference panics.
"""
import ctypes
import os
import time

import numpy as np

from ._lib import BABYBEAR, KOALABEAR, PG_E_INVALID, Challenger as _ChStruct, PgError, check, lib

PRIMES = {KOALABEAR: 2130706433, BABYBEAR: 2013265921}
GENERATORS = {KOALABEAR: 3, BABYBEAR: 31}
TWO_ADICITY = {KOALABEAR: 24, BABYBEAR: 27}


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def to_monty(field, a):
    p = np.uint64(PRIMES[field])
    return ((np.asarray(a, dtype=np.uint64) << np.uint64(32)) % p).astype(np.uint32)


def from_monty(field, a):
    p = PRIMES[field]
    rinv = np.uint64(pow(2 ** 32, p - 2, p))
    return ((np.asarray(a, dtype=np.uint64) * rinv) % np.uint64(p)).astype(np.uint32)


def device_synchronize(device=0):
    check(lib().pg_device_synchronize(device))


def two_adic_generator(field, log_n):
    p = PRIMES[field]
    return pow(GENERATORS[field], (p - 1) >> log_n, p)


class Context:
    """One device, an HBM budget (bytes, 0 = all free memory) and table caches."""

    def __init__(self, device=0, hbm_budget=0):
        h = ctypes.c_void_p()
        check(lib().pg_create(device, hbm_budget, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().pg_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def bytes_in_use(self):
        return lib().pg_bytes_in_use(self.handle)

    def profile(self, enable=True):
        check(lib().pg_profile_enable(self.handle, int(enable)))

    def debug_fail_alloc(self, after):
        """Make the device allocation `after` allocations from now fail once
        with PG_E_OOM (negative: disarm).  For tests of error paths."""
        check(lib().pg_debug_fail_alloc(self.handle, int(after)))

    def debug_host_blocks(self, block_bytes=0, min_cols=0):
        """Column-block sizing of commit_host (0: the defaults).  For tests
        of the blocked path on small matrices."""
        check(lib().pg_debug_host_blocks(self.handle, int(block_bytes), int(min_cols)))

    def profile_reset(self):
        check(lib().pg_profile_reset(self.handle))

    def profile_get(self, family):
        """(total milliseconds, launches) recorded for a kernel family."""
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        check(lib().pg_profile_get(self.handle, family.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def matrix(self, rows, cols):
        h = ctypes.c_void_p()
        check(lib().pg_buf_alloc(self.handle, rows, cols, ctypes.byref(h)))
        return DeviceMatrix(self, h, owned=True)

    def upload(self, host, field=None):
        """Copy a 2-D uint32 array to a new device matrix.  With `field`, the
        host holds canonical values, converted to Montgomery form on device;
        without it the values are copied as they are (Montgomery)."""
        a = np.ascontiguousarray(host, dtype=np.uint32)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        m = self.matrix(a.shape[0], a.shape[1])
        try:
            check(lib().pg_upload(self.handle, m.handle, _ptr(a), -1 if field is None else field,
                                  int(field is not None)))
        except Exception:
            m.free()
            raise
        return m

    def random_matrix(self, field, seed, rows, cols):
        m = self.matrix(rows, cols)
        check(lib().pg_fill_random(self.handle, m.handle, field, seed))
        return m

    def host_matrix(self, rows, cols):
        """A rows x cols uint32 array in page-locked host memory (pg_host_alloc);
        uploads from it run at the link's rate.  Free with .free()."""
        return PinnedMatrix(self, rows, cols)

    def wrap(self, device_ptr, rows, cols):
        h = ctypes.c_void_p()
        check(lib().pg_buf_wrap(self.handle, ctypes.c_void_p(device_ptr), rows, cols, ctypes.byref(h)))
        return DeviceMatrix(self, h, owned=True)  # the view object itself is freed; memory is not


# ------------------------------------------------------------- autotune --
LDE_PASSES = {"single": 0, "a": 1, "bc": 2, "d": 3}


def autotune(ctx, field, log_h_min, log_h_max, widths, log_blowup=1, reps=3):
    """Time both tile shapes of every LDE pass for these heights and widths
    and keep the faster (process-wide); returns the decision report."""
    w = np.ascontiguousarray(widths, dtype=np.uint32)
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().pg_autotune(ctx.handle, field, log_h_min, log_h_max, _ptr(w), w.size, log_blowup, reps, buf,
                            len(buf)))
    return buf.value.decode()


TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune", "lde_shapes.txt")


def load_tuning(path=None):
    """Import a tile-shape table (default: the shipped MI355X table, or the
    file named by PICO_AMD_LDE_TUNING; "none" keeps the built-in shapes).
    Host-side only; returns the path used or None."""
    path = path or os.environ.get("PICO_AMD_LDE_TUNING", TUNING_FILE)
    if path == "none" or not os.path.exists(path):
        return None
    with open(path) as fh:
        tune_import(fh.read())
    return path


def tune_export():
    """The tile-shape table as text ("field kind log_s width_bucket variant" lines)."""
    n = ctypes.c_size_t()
    check(lib().pg_tune_export(None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    check(lib().pg_tune_export(buf, n.value, ctypes.byref(n)))
    return buf.value.decode()


def tune_import(text):
    check(lib().pg_tune_import(text.encode()))


def tune_set(field, kind, log_s, width_bucket, variant):
    check(lib().pg_tune_set(field, kind, log_s, width_bucket, variant))


def tune_get(field, kind, log_s, width):
    v = ctypes.c_int()
    check(lib().pg_tune_get(field, kind, log_s, width, ctypes.byref(v)))
    return v.value


class PinnedMatrix:
    """Page-locked host buffer viewed as a numpy uint32 matrix (.array)."""

    def __init__(self, ctx, rows, cols):
        p = ctypes.c_void_p()
        check(lib().pg_host_alloc(ctx.handle, rows * cols * 4, ctypes.byref(p)))
        self.ptr = p
        buf = (ctypes.c_uint32 * (rows * cols)).from_address(p.value)
        self.array = np.ctypeslib.as_array(buf).reshape(rows, cols)

    def free(self):
        if self.ptr:
            self.array = None
            lib().pg_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceMatrix:
    """Row-major matrix of Montgomery-form u32 in HBM (a pg_buf)."""

    def __init__(self, ctx, handle, owned):
        self.ctx = ctx
        self.handle = handle
        self._owned = owned
        self.rows = lib().pg_buf_rows(handle)
        self.cols = lib().pg_buf_cols(handle)

    @property
    def device_ptr(self):
        return lib().pg_buf_device_ptr(self.handle)

    def height(self):
        return self.rows

    def width(self):
        return self.cols

    @property
    def ld(self):
        """Row pitch in words (cols unless this is a strided view)."""
        return lib().pg_buf_ld(self.handle)

    def view(self, row0=0, row_step=1, col0=0, rows=None, cols=None):
        """Rows row0 + row_step * r and columns col0 .. col0 + cols - 1, as a
        matrix that shares this one's memory (it keeps this one alive).  Views
        are LDE/commit inputs and can be downloaded; nothing writes to them."""
        if rows is None:
            rows = 0 if row0 >= self.rows else (self.rows - 1 - row0) // row_step + 1
        cols = self.cols - col0 if cols is None else cols
        h = ctypes.c_void_p()
        check(lib().pg_buf_view(self.ctx.handle, self.handle, row0, row_step, col0, rows, cols,
                                ctypes.byref(h)))
        v = DeviceMatrix(self.ctx, h, owned=True)
        v._parent = self
        return v

    def to_host(self, field=None, row0=0, nrows=None):
        """Rows [row0, row0 + nrows): canonical values of `field`, or the raw
        Montgomery words when no field is given."""
        nrows = self.rows - row0 if nrows is None else nrows
        out = np.empty((nrows, self.cols), np.uint32)
        check(lib().pg_download(self.ctx.handle, self.handle, row0, nrows, _ptr(out),
                                -1 if field is None else field, int(field is not None)))
        return out

    def free(self):
        if self._owned and self.handle:
            lib().pg_buf_free(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ DFT --
class Radix2DitParallel:
    """TwoAdicSubgroupDft on the device (p3 Radix2DitParallel's role)."""

    def __init__(self, ctx, field):
        self.ctx, self.field = ctx, field

    def coset_lde_batch(self, mat, added_bits, shift, bit_reversed=True):
        """LDE of `mat` (values on <w_h>) onto shift * <w_N>; rows in committed
        (bit-reversed) order unless bit_reversed=False.  `shift` is canonical."""
        h = ctypes.c_void_p()
        s = int(to_monty(self.field, shift))
        check(lib().pg_coset_lde(self.ctx.handle, self.field, mat.handle, added_bits, s,
                                 int(bool(bit_reversed)), ctypes.byref(h)))
        return DeviceMatrix(self.ctx, h, owned=True)


# ----------------------------------------------------------------- MMCS --
class ProverData:
    """MMCS prover data: the committed matrices and every digest layer."""

    def __init__(self, ctx, field, handle, keep=()):
        self.ctx, self.field, self.handle = ctx, field, handle
        self._keep = list(keep)  # caller matrices the tree references
        self.log_height = lib().pg_tree_log_height(handle)
        self.total_width = lib().pg_tree_total_width(handle)
        self.num_matrices = lib().pg_tree_num_matrices(handle)

    def root(self):
        r = np.zeros(8, np.uint32)
        check(lib().pg_tree_root(self.handle, _ptr(r)))
        return from_monty(self.field, r)

    def digest_layer(self, layer):
        n = 1 << (self.log_height - layer)
        out = np.zeros((n, 8), np.uint32)
        check(lib().pg_tree_digest_layer(self.ctx.handle, self.handle, layer, _ptr(out)))
        return from_monty(self.field, out)

    def free(self):
        if self.handle:
            lib().pg_tree_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class MerkleTreeMmcs:
    """Mmcs with PaddingFreeSponge<Poseidon2,16,8,8> leaves and
    TruncatedPermutation<Poseidon2,2,8,16> compression (kb_poseidon2.rs:17-20)."""

    def __init__(self, ctx, field):
        self.ctx, self.field = ctx, field

    def commit(self, mats, take_ownership=False):
        arr = (ctypes.c_void_p * len(mats))(*[m.handle for m in mats])
        root = np.zeros(8, np.uint32)
        h = ctypes.c_void_p()
        check(lib().pg_mmcs_commit(self.ctx.handle, self.field, arr, len(mats), int(take_ownership),
                                   _ptr(root), ctypes.byref(h)))
        if take_ownership:
            for m in mats:
                m._owned = False  # the tree frees them
        pd = ProverData(self.ctx, self.field, h, keep=mats)
        return from_monty(self.field, root), pd

    def commit_matrix(self, mat):
        return self.commit([mat])

    def get_matrices(self, data):
        return [DeviceMatrix(self.ctx, lib().pg_tree_matrix(data.handle, k), owned=False)
                for k in range(data.num_matrices)]

    def open_many(self, indices, data):
        """Batched open_batch: (rows[nq, total_width], proofs[nq, log_h, 8]), canonical."""
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        nq = idx.size
        rows = np.zeros((nq, max(data.total_width, 1)), np.uint32)
        paths = np.zeros((nq, max(data.log_height, 1), 8), np.uint32)
        check(lib().pg_mmcs_open_batch(self.ctx.handle, data.handle, _ptr(idx), nq, _ptr(rows),
                                       _ptr(paths)))
        return (from_monty(self.field, rows[:, :data.total_width]),
                from_monty(self.field, paths[:, :data.log_height]))

    def open_batch(self, index, data):
        """(opened rows per matrix in original order, sibling path)."""
        rows, paths = self.open_many([index], data)
        widths = [m.cols for m in self.get_matrices(data)]
        out, pos = [], 0
        for w in widths:
            out.append(rows[0, pos:pos + w])
            pos += w
        return out, paths[0]

    def verify_batch(self, commit, dimensions, index, opened_values, proof):
        """dimensions: [(width, height)] in original order."""
        hs = np.array([d[1] for d in dimensions], np.uintp)
        ws = np.array([d[0] for d in dimensions], np.uintp)
        flat = [np.asarray(v, np.uint32).ravel() for v in opened_values]
        rows = to_monty(self.field, np.concatenate(flat) if flat else np.zeros(1, np.uint32))
        rows = np.ascontiguousarray(rows if rows.size else np.zeros(1, np.uint32))
        path = np.ascontiguousarray(to_monty(self.field, np.asarray(proof, np.uint32).reshape(-1)))
        if path.size == 0:
            path = np.zeros(8, np.uint32)
        root = np.ascontiguousarray(to_monty(self.field, commit))
        rc = lib().pg_mmcs_verify_batch(self.field, _ptr(hs), _ptr(ws), len(dimensions), _ptr(root),
                                        index, _ptr(rows), _ptr(path))
        if rc < 0:
            check(rc)
        return rc == 1


# ------------------------------------------------------------------ PCS --
class TwoAdicMultiplicativeCoset:
    """shift * <w_{2^log_n}>, shift canonical.  `field` is needed only by
    split_domains (the chunk shifts are field elements)."""

    def __init__(self, log_n, shift=1, field=None):
        self.log_n, self.shift, self.field = log_n, shift, field

    def size(self):
        return 1 << self.log_n

    def split_domains(self, num_chunks, field=None):
        """Chunk i: shift * w_N^i * <w_{N / num_chunks}> (N = this size), the
        domains of the quotient chunks (vm/src/machine/prover.rs:513-523;
        verifier side vm/src/compiler/recursion/circuit/stark.rs:269-270)."""
        f = self.field if field is None else field
        lc = num_chunks.bit_length() - 1
        if num_chunks != 1 << lc or lc > self.log_n:
            raise ValueError("num_chunks must be a power of two <= the domain size")
        p = PRIMES[f]
        w = two_adic_generator(f, self.log_n)
        return [TwoAdicMultiplicativeCoset(self.log_n - lc, self.shift * pow(w, i, p) % p, f)
                for i in range(num_chunks)]

    def split_evals(self, num_chunks, evals):
        """Row i of `evals` (a DeviceMatrix of this domain's size) goes to
        chunk i mod num_chunks: strided views of the same device memory, no
        copy (the reference builds new host matrices, prover.rs:520)."""
        if evals.rows != self.size() or num_chunks < 1 or self.size() % num_chunks:
            raise ValueError("evaluations do not match the domain or the chunk count")
        return [evals.view(row0=i, row_step=num_chunks, rows=self.size() // num_chunks)
                for i in range(num_chunks)]


class TwoAdicFriPcs:
    """Pcs::commit and get_evaluations_on_domain on device-resident data."""

    def __init__(self, ctx, field, log_blowup):
        self.ctx, self.field, self.log_blowup = ctx, field, log_blowup

    def natural_domain_for_degree(self, n):
        return TwoAdicMultiplicativeCoset(n.bit_length() - 1, 1, self.field)

    def commit(self, evaluations):
        """evaluations: [(domain, DeviceMatrix)] -> (root, ProverData)."""
        n = len(evaluations)
        arr = (ctypes.c_void_p * n)(*[m.handle for _, m in evaluations])
        shifts = np.ascontiguousarray(to_monty(self.field, [d.shift for d, _ in evaluations]))
        root = np.zeros(8, np.uint32)
        h = ctypes.c_void_p()
        check(lib().pg_pcs_commit(self.ctx.handle, self.field, arr, _ptr(shifts), n, self.log_blowup,
                                  _ptr(root), ctypes.byref(h)))
        return from_monty(self.field, root), ProverData(self.ctx, self.field, h)

    def commit_host(self, evaluations, canonical=True):
        """evaluations: [(domain, 2-D uint32 host array)] -> (root, ProverData).
        The arrays stay on the host until the library uploads them, a height
        group at a time, overlapped with the commit of the taller groups.
        canonical=False means the arrays already hold Montgomery words."""
        n = len(evaluations)
        arrs = [np.ascontiguousarray(a, dtype=np.uint32) for _, a in evaluations]
        arrs = [a.reshape(-1, 1) if a.ndim == 1 else a for a in arrs]
        ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in arrs])
        rows = np.array([a.shape[0] for a in arrs], np.uint64)
        cols = np.array([a.shape[1] for a in arrs], np.uint64)
        shifts = np.ascontiguousarray(to_monty(self.field, [d.shift for d, _ in evaluations]))
        root = np.zeros(8, np.uint32)
        h = ctypes.c_void_p()
        check(lib().pg_pcs_commit_host(self.ctx.handle, self.field, ptrs, _ptr(rows), _ptr(cols), _ptr(shifts), n,
                                       self.log_blowup, int(bool(canonical)), _ptr(root), ctypes.byref(h)))
        return from_monty(self.field, root), ProverData(self.ctx, self.field, h)

    def get_evaluations_on_domain(self, data, idx, domain):
        """Rows of committed matrix `idx` on `domain` (natural order).  As in
        p3 TwoAdicFriPcs, only domains with shift GENERATOR (a sub-coset of
        the committed LDE coset of a trace domain) are available."""
        if domain.shift % PRIMES[self.field] != GENERATORS[self.field]:
            raise PgError(PG_E_INVALID, "get_evaluations_on_domain: domain shift must be the field generator")
        m = lib().pg_tree_matrix(data.handle, idx)
        cols = lib().pg_buf_cols(m)
        out = np.zeros((domain.size(), cols), np.uint32)
        check(lib().pg_pcs_evaluations(self.ctx.handle, data.handle, idx, domain.size(), _ptr(out)))
        return from_monty(self.field, out)

    # ---- open -------------------------------------------------------------
    def open_values(self, data, idx, points):
        """Matrix `idx` of committed `data` at canonical ext4 points
        (npts x 4) -> canonical values, npts x width x 4."""
        pts = np.ascontiguousarray(to_monty(self.field, np.asarray(points, np.uint64).reshape(-1, 4)))
        m = lib().pg_tree_matrix(data.handle, idx)
        w = lib().pg_buf_cols(m)
        out = np.zeros((len(pts), w, 4), np.uint32)
        check(lib().pg_pcs_open_values(self.ctx.handle, data.handle, idx, self.log_blowup, _ptr(pts),
                                       len(pts), _ptr(out)))
        return from_monty(self.field, out)

    def reduce_openings(self, rounds, values, alpha):
        """rounds: [(ProverData, [points of matrix 0, points of matrix 1, ...])]
        with each matrix's points an (n x 4) canonical array (n <= 2);
        values: the opened values in the same nesting (per matrix, n x w x 4).
        Returns {log_height: DeviceMatrix (2^lh x 4, ext4 rows)}."""
        trees = (ctypes.c_void_p * len(rounds))(*[d.handle for d, _ in rounds])
        npts, pts, vals = [], [], []
        for (d, mat_points), mat_values in zip(rounds, values):
            for p, v in zip(mat_points, mat_values):
                p = np.asarray(p, np.uint64).reshape(-1, 4)
                npts.append(len(p))
                pts.append(p)
                vals.append(np.asarray(v, np.uint64).reshape(-1, 4))
        npts = np.ascontiguousarray(npts, dtype=np.uint32)
        pts = np.ascontiguousarray(to_monty(self.field, np.concatenate(pts) if pts else np.zeros((1, 4))))
        vals = np.ascontiguousarray(to_monty(self.field, np.concatenate(vals) if vals else np.zeros((1, 4))))
        a = np.ascontiguousarray(to_monty(self.field, np.asarray(alpha, np.uint64)))
        outs = (ctypes.c_void_p * 33)()
        check(lib().pg_pcs_reduce_openings(self.ctx.handle, trees, len(rounds), _ptr(npts), _ptr(pts), _ptr(vals),
                                           _ptr(a), outs))
        return {lh: DeviceMatrix(self.ctx, outs[lh], owned=True) for lh in range(33) if outs[lh]}

    def open(self, rounds, challenger, num_queries, pow_bits, log_final=None):
        """TwoAdicFriPcs::open through pg_pcs_open: rounds = [(ProverData,
        [points per matrix])].  Evaluates, has the challenger observe every
        value (round -> matrix -> point -> column, 4 coefficients each),
        samples alpha, reduces, runs the FRI commit phase down to 2^log_final
        (default log_blowup, the reference's stop, final vector constant),
        grinds, samples the query indices and answers them.  Returns
        (opened_values, proof dict), all canonical."""
        f = self.field
        npts, pts, widths = [], [], []
        for d, mat_points in rounds:
            ws = [lib().pg_buf_cols(lib().pg_tree_matrix(d.handle, k)) for k in range(d.num_matrices)]
            if len(mat_points) != len(ws):
                raise ValueError("one list of points per committed matrix")
            widths.append(ws)
            for p in mat_points:
                p = np.asarray(p, np.uint64).reshape(-1, 4)
                npts.append(len(p))
                pts.append(p)
        trees = (ctypes.c_void_p * len(rounds))(*[d.handle for d, _ in rounds])
        npts_a = np.ascontiguousarray(npts, dtype=np.uint32)
        pts_a = np.ascontiguousarray(to_monty(f, np.concatenate(pts) if pts else np.zeros((0, 4), np.uint64)))
        h = ctypes.c_void_p()
        t0 = time.perf_counter()
        check(lib().pg_pcs_open(self.ctx.handle, trees, len(rounds), _ptr(npts_a), _ptr(pts_a), self.log_blowup,
                                -1 if log_final is None else int(log_final), int(num_queries), int(pow_bits),
                                ctypes.byref(challenger.c), ctypes.byref(h)))
        # duration of the native call alone (the unpacking below is Python's)
        self.last_open_seconds = time.perf_counter() - t0
        try:
            sz = [ctypes.c_size_t() for _ in range(5)]
            nr = ctypes.c_int()
            check(lib().pg_open_proof_sizes(h, ctypes.byref(sz[0]), ctypes.byref(nr), ctypes.byref(sz[1]),
                                            ctypes.byref(sz[2]), ctypes.byref(sz[3]), ctypes.byref(sz[4])))
            nvals, nfinal, nq, inw, fpw = (x.value for x in sz)
            R = nr.value
            values = np.zeros(max(nvals, 1), np.uint32)
            alpha = np.zeros(4, np.uint32)
            commits = np.zeros((max(R, 1), 8), np.uint32)
            betas = np.zeros((max(R, 1), 4), np.uint32)
            final = np.zeros((nfinal // 4, 4), np.uint32)
            wit = ctypes.c_uint32()
            qidx = np.zeros(max(nq, 1), np.uint64)
            inputs = np.zeros(max(nq * inw, 1), np.uint32)
            sib = np.zeros((max(nq, 1), max(R, 1), 4), np.uint32)
            fpaths = np.zeros(max(nq * fpw, 1), np.uint32)
            check(lib().pg_open_proof_read(h, _ptr(values), _ptr(alpha), _ptr(commits), _ptr(betas), _ptr(final),
                                           ctypes.byref(wit), _ptr(qidx), _ptr(inputs), _ptr(sib), _ptr(fpaths)))
            # the FriProof as the reference's bincode holds it (constant final
            # vector only: p3's FriProof keeps one final value)
            bincode = _encode_open_proof(h, canonical=False) if log_final is None else None
        finally:
            lib().pg_open_proof_free(h)
        values = from_monty(f, values[:nvals])
        opened, off, mi = [], 0, 0
        for ws in widths:
            per_mat = []
            for w in ws:
                n = npts[mi] * w * 4
                per_mat.append(values[off:off + n].reshape(npts[mi], w, 4))
                off += n
                mi += 1
            opened.append(per_mat)
        log_heights = [d.log_height for d, _ in rounds]
        input_openings = []
        inputs = from_monty(f, inputs[:nq * inw]).reshape(nq, inw) if nq else inputs[:0]
        for q in range(nq):
            per_round, o = [], 0
            for ws, lh in zip(widths, log_heights):
                rows = []
                for w in ws:
                    rows.append(inputs[q, o:o + w])
                    o += w
                path = inputs[q, o:o + 8 * lh].reshape(lh, 8)
                o += 8 * lh
                per_round.append((rows, path))
            input_openings.append(per_round)
        log_max = max(log_heights)  # pg_pcs_open rejects a round taller than every opened matrix
        plens = [log_max - 1 - r for r in range(R)]
        fpaths = from_monty(f, fpaths[:nq * fpw]).reshape(nq, fpw) if nq else fpaths[:0]
        fri_paths = []
        for q in range(nq):
            lst, o = [], 0
            for L in plens:
                lst.append(fpaths[q, o:o + 8 * L].reshape(L, 8))
                o += 8 * L
            fri_paths.append(lst)
        proof = {"alpha": from_monty(f, alpha), "commits": from_monty(f, commits[:R]),
                 "betas": from_monty(f, betas[:R]), "final_poly": from_monty(f, final), "pow_witness": wit.value,
                 "query_indices": [int(x) for x in qidx[:nq]], "input_openings": input_openings,
                 "fri_siblings": from_monty(f, sib[:nq, :R]), "fri_paths": fri_paths, "bincode": bincode}
        return opened, proof


def _encode_open_proof(handle, canonical):
    n = ctypes.c_size_t()
    check(lib().pg_open_proof_encode(handle, int(bool(canonical)), None, 0, ctypes.byref(n)))
    buf = (ctypes.c_uint8 * n.value)()
    check(lib().pg_open_proof_encode(handle, int(bool(canonical)), buf, n.value, ctypes.byref(n)))
    return bytes(buf)


def decode_fri_proof(field, data, canonical=False):
    """pg_open_proof_decode: bincode bytes of a FriProof (see
    TwoAdicFriPcs.open's proof["bincode"]) -> its parts, canonical values:
    commits (R x 8), input_openings[query][round] = (rows per matrix, path
    levels x 8), fri_siblings (queries x R x 4), fri_paths[query][round]
    (levels x 8), final_poly (4), pow_witness; plus "reencoded", the bytes
    pg_open_proof_encode writes for the decoded proof (Montgomery words)."""
    raw = (ctypes.c_uint8 * max(len(data), 1)).from_buffer_copy(bytes(data) or b"\0")
    h = ctypes.c_void_p()
    check(lib().pg_open_proof_decode(field, raw, len(data), int(bool(canonical)), ctypes.byref(h)))
    try:
        sz = [ctypes.c_size_t() for _ in range(5)]
        nr = ctypes.c_int()
        check(lib().pg_open_proof_sizes(h, ctypes.byref(sz[0]), ctypes.byref(nr), ctypes.byref(sz[1]),
                                        ctypes.byref(sz[2]), ctypes.byref(sz[3]), ctypes.byref(sz[4])))
        _, nfinal, nq, inw, fpw = (x.value for x in sz)
        R = nr.value
        commits = np.zeros((max(R, 1), 8), np.uint32)
        final = np.zeros(max(nfinal, 4), np.uint32)
        wit = ctypes.c_uint32()
        inputs = np.zeros(max(nq * inw, 1), np.uint32)
        sib = np.zeros((max(nq, 1), max(R, 1), 4), np.uint32)
        fpaths = np.zeros(max(nq * fpw, 1), np.uint32)
        check(lib().pg_open_proof_read(h, None, None, _ptr(commits), None, _ptr(final), ctypes.byref(wit), None,
                                       _ptr(inputs), _ptr(sib), _ptr(fpaths)))
        reencoded = _encode_open_proof(h, canonical=False)
    finally:
        lib().pg_open_proof_free(h)
    return {"commits": from_monty(field, commits[:R]), "final_poly": from_monty(field, final[:4]),
            "pow_witness": wit.value,
            "inputs_flat": from_monty(field, inputs[:nq * inw]).reshape(nq, inw),
            "fri_siblings": from_monty(field, sib[:nq, :R]),
            "fri_paths_flat": from_monty(field, fpaths[:nq * fpw]).reshape(nq, fpw),
            "num_queries": nq, "reencoded": reencoded}


class CommitBatch:
    """Many single-matrix Pcs::commit jobs as one launch sequence (optionally a
    captured hipGraph).  jobs: [(DeviceMatrix, log_blowup)] or
    [(DeviceMatrix, log_blowup, domain_shift)].  The inputs must stay alive
    while the batch exists."""

    def __init__(self, ctx, field, jobs, use_graph=True):
        self.ctx, self.field, self.n = ctx, field, len(jobs)
        self._inputs = [j[0] for j in jobs]
        arr = (ctypes.c_void_p * self.n)(*[j[0].handle for j in jobs])
        lbs = np.ascontiguousarray([j[1] for j in jobs], dtype=np.uint32)
        shifts = None
        if any(len(j) > 2 for j in jobs):
            shifts = np.ascontiguousarray(to_monty(field, [j[2] if len(j) > 2 else 1 for j in jobs]))
        h = ctypes.c_void_p()
        check(lib().pg_batch_create(ctx.handle, field, arr, None if shifts is None else _ptr(shifts),
                                    _ptr(lbs), self.n, int(use_graph), ctypes.byref(h)))
        self.handle = h
        self.num_launches = lib().pg_batch_num_launches(h)

    def run(self):
        roots = np.zeros((self.n, 8), np.uint32)
        check(lib().pg_batch_run(self.ctx.handle, self.handle, _ptr(roots)))
        return from_monty(self.field, roots)

    def take(self, j):
        h = ctypes.c_void_p()
        check(lib().pg_batch_take_tree(self.handle, j, ctypes.byref(h)))
        return ProverData(self.ctx, self.field, h)

    def free(self):
        if self.handle:
            lib().pg_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CommitService:
    """Device-resident commit service (pg_service): submit shard commit jobs as
    they arrive, wait for roots, open from the resident prover data, release."""

    def __init__(self, ctx, budget=0, workers=3):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().pg_service_create(ctx.handle, budget, workers, ctypes.byref(h)))
        self.handle = h
        self._held = {}  # host traces of submitted jobs, until waited for

    def submit(self, field, trace, log_blowup):
        """Hands `trace` (a DeviceMatrix) to the service; returns a job id."""
        jid = ctypes.c_uint64()
        check(lib().pg_service_submit(self.handle, field, trace.handle, log_blowup, ctypes.byref(jid)))
        trace._owned = False
        trace.handle = None
        return jid.value

    def submit_host(self, field, trace, log_blowup, canonical=True):
        """A 2-D uint32 host array (canonical values, or Montgomery words with
        canonical=False); the upload is part of the job.  The service keeps a
        reference to the array until the job is waited for."""
        arr = np.ascontiguousarray(trace, dtype=np.uint32)
        arr = arr.reshape(-1, 1) if arr.ndim == 1 else arr
        jid = ctypes.c_uint64()
        check(lib().pg_service_submit_host(self.handle, field, _ptr(arr), arr.shape[0], arr.shape[1],
                                           int(bool(canonical)), log_blowup, ctypes.byref(jid)))
        self._held[jid.value] = arr
        return jid.value

    def submit_synthetic(self, field, seed, rows, cols, log_blowup):
        jid = ctypes.c_uint64()
        check(lib().pg_service_submit_synthetic(self.handle, field, seed, rows, cols, log_blowup,
                                                ctypes.byref(jid)))
        return jid.value

    def wait(self, job_id, field):
        """(root canonical[8], (t_submit, t_start, t_done))"""
        root = np.zeros(8, np.uint32)
        times = np.zeros(3, np.float64)
        try:
            check(lib().pg_service_wait(self.handle, job_id, _ptr(root), _ptr(times)))
        finally:
            self._held.pop(job_id, None)  # the job no longer reads its host trace
        return from_monty(field, root), tuple(times)

    def open(self, job_id, field, indices, total_width, log_height):
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        rows = np.zeros((idx.size, total_width), np.uint32)
        paths = np.zeros((idx.size, max(log_height, 1), 8), np.uint32)
        check(lib().pg_service_open(self.handle, job_id, _ptr(idx), idx.size, _ptr(rows), _ptr(paths)))
        return from_monty(field, rows), from_monty(field, paths[:, :log_height])

    def release(self, job_id):
        check(lib().pg_service_release(self.handle, job_id))

    def cancel(self, job_id):
        """Take a job that has not started off the queue (wait() then raises
        PgError with PG_E_CANCELLED); release it afterwards."""
        check(lib().pg_service_cancel(self.handle, job_id))

    def stats(self):
        r, p, k = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().pg_service_stats(self.handle, ctypes.byref(r), ctypes.byref(p), ctypes.byref(k)))
        return {"reserved_bytes": r.value, "pending": p.value, "peak_running": k.value}

    def close(self):
        if self.handle:
            lib().pg_service_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------- challenger --
class DuplexChallenger:
    """DuplexChallenger<F, Poseidon2, 16, 8>; values are canonical here."""

    def __init__(self, field):
        self.field = field
        self.c = _ChStruct()
        lib().pg_challenger_init(ctypes.byref(self.c), field)

    def observe(self, values):
        for v in np.atleast_1d(np.asarray(values, np.uint64)).ravel():
            lib().pg_challenger_observe(ctypes.byref(self.c), int(to_monty(self.field, int(v))))

    def sample(self):
        return int(from_monty(self.field, lib().pg_challenger_sample(ctypes.byref(self.c))))

    def sample_ext(self):
        out = np.zeros(4, np.uint32)
        lib().pg_challenger_sample_ext(ctypes.byref(self.c), _ptr(out))
        return from_monty(self.field, out)

    def sample_bits(self, bits):
        return int(lib().pg_challenger_sample_bits(ctypes.byref(self.c), bits))

    def check_witness(self, bits, witness):
        return bool(lib().pg_challenger_check_witness(ctypes.byref(self.c), bits, witness))

    def clone(self):
        c = DuplexChallenger(self.field)
        ctypes.memmove(ctypes.byref(c.c), ctypes.byref(self.c), ctypes.sizeof(_ChStruct))
        return c


# ------------------------------------------------------------------ FRI --
def fri_fold(ctx, field, v, beta):
    """fold_even_odd of an n x 4 ext4 device vector by canonical beta[4]."""
    b = np.ascontiguousarray(to_monty(field, beta))
    h = ctypes.c_void_p()
    check(lib().pg_fri_fold(ctx.handle, field, v.handle, _ptr(b), ctypes.byref(h)))
    return DeviceMatrix(ctx, h, owned=True)


class FriProverData:
    def __init__(self, ctx, field, handle, roots, betas, final):
        self.ctx, self.field, self.handle = ctx, field, handle
        self.commits, self.betas, self.final_poly = roots, betas, final
        self.num_rounds = lib().pg_fri_num_rounds(handle)

    def round_data(self, r):
        t = lib().pg_fri_round_tree(self.handle, r)
        pd = ProverData(self.ctx, self.field, t)
        pd.free = lambda: None  # owned by the FRI data
        return pd

    def free(self):
        if self.handle:
            lib().pg_fri_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fri_commit_phase(ctx, field, challenger, inputs, log_stop, final_mode=0):
    """inputs: {log_len: DeviceMatrix (2^log_len x 4)}.  Returns FriProverData
    (commits, betas, final vector canonical)."""
    max_log = max(inputs)
    arr = (ctypes.c_void_p * (max_log + 1))()
    for k, m in inputs.items():
        arr[k] = m.handle
    rounds = max_log - log_stop
    roots = np.zeros((max(rounds, 1), 8), np.uint32)
    betas = np.zeros((max(rounds, 1), 4), np.uint32)
    final = np.zeros((1 << log_stop, 4), np.uint32)
    nr = ctypes.c_int()
    h = ctypes.c_void_p()
    check(lib().pg_fri_commit_phase(ctx.handle, field, ctypes.byref(challenger.c), arr, max_log,
                                    log_stop, final_mode, _ptr(roots), _ptr(betas), _ptr(final),
                                    ctypes.byref(nr), ctypes.byref(h)))
    r = nr.value
    return FriProverData(ctx, field, h, from_monty(field, roots[:r]), from_monty(field, betas[:r]),
                         from_monty(field, final))


def fri_grind(ctx, field, challenger, bits):
    w = ctypes.c_uint32()
    check(lib().pg_fri_grind(ctx.handle, field, ctypes.byref(challenger.c), bits, ctypes.byref(w)))
    return w.value


def fri_answer_queries(data, indices, log_max):
    """Returns (siblings[nq, R, 4], paths: list over queries of list over rounds
    of [log_max-1-r, 8] arrays), canonical."""
    idx = np.ascontiguousarray(indices, dtype=np.uint64)
    nq, R = idx.size, data.num_rounds
    plens = [log_max - 1 - r for r in range(R)]
    per_q = 8 * sum(plens)
    sib = np.zeros((nq, max(R, 1), 4), np.uint32)
    paths = np.zeros(max(nq * per_q, 1), np.uint32)
    check(lib().pg_fri_answer_queries(data.ctx.handle, data.handle, _ptr(idx), nq, _ptr(sib),
                                      _ptr(paths)))
    sib = from_monty(data.field, sib[:, :R])
    if not nq:
        return sib, []
    paths = from_monty(data.field, paths[:nq * per_q]).reshape(nq, per_q)
    per_round, off = [], 0  # one (nq, L, 8) view per round
    for L in plens:
        per_round.append(paths[:, off:off + 8 * L].reshape(nq, L, 8))
        off += 8 * L
    return sib, [[pr[q] for pr in per_round] for q in range(nq)]


__all__ = [
    "KOALABEAR", "BABYBEAR", "PRIMES", "GENERATORS", "Context", "DeviceMatrix",
    "Radix2DitParallel", "MerkleTreeMmcs", "ProverData", "TwoAdicFriPcs",
    "TwoAdicMultiplicativeCoset", "DuplexChallenger", "fri_fold", "fri_commit_phase",
    "fri_grind", "fri_answer_queries", "to_monty", "from_monty", "two_adic_generator", "PgError",
    "device_synchronize", "CommitBatch", "CommitService", "PinnedMatrix",
    "autotune", "tune_export", "tune_import", "tune_set", "tune_get", "LDE_PASSES", "load_tuning",
    "TUNING_FILE", "decode_fri_proof",
]
