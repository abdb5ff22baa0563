"""Small pure-Python restatements used to pin the oracle on tiny inputs.

Written independently of oracle/oracle.c from the definitions (polynomial
evaluation, the verify_batch walk of vm/src/compiler/recursion/circuit/fri.rs:389-436,
the fold line of circuit/fri.rs:345-383, the duplex rules of
circuit/challenger.rs:89-227).  Python ints, so only for small sizes.
"""
P = {0: 2130706433, 1: 2013265921}
GEN = {0: 3, 1: 31}
W = {0: 3, 1: 11}


def rev(x, bits):
    r = 0
    for i in range(bits):
        r |= ((x >> i) & 1) << (bits - 1 - i)
    return r


def root_of_unity(f, log_n):
    p = P[f]
    return pow(GEN[f], (p - 1) >> log_n, p)


def interpolate_coeffs(f, vals):
    """Coefficients of the polynomial taking vals[j] at w_h^j (Lagrange via the
    inverse DFT definition, O(h^2))."""
    p = P[f]
    h = len(vals)
    lh = h.bit_length() - 1
    winv = pow(root_of_unity(f, lh), p - 2, p)
    hinv = pow(h, p - 2, p)
    return [sum(vals[j] * pow(winv, j * k, p) for j in range(h)) * hinv % p for k in range(h)]


def evaluate(f, coeffs, x):
    p = P[f]
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % p
    return acc


def coset_lde_column(f, vals, added_bits, shift, bit_reverse=True):
    p = P[f]
    coeffs = interpolate_coeffs(f, vals)
    n = len(vals) << added_bits
    ln = n.bit_length() - 1
    wn = root_of_unity(f, ln)
    out = []
    for i in range(n):
        e = rev(i, ln) if bit_reverse else i
        out.append(evaluate(f, coeffs, shift * pow(wn, e, p) % p))
    return out


def ext_mul(f, a, b):
    p, w = P[f], W[f]
    c = [0] * 7
    for i in range(4):
        for j in range(4):
            c[i + j] += a[i] * b[j]
    return [(c[k] + (w * c[k + 4] if k + 4 < 7 else 0)) % p for k in range(4)]


def ext_inv_by_search_check(f, a, inv):
    return ext_mul(f, a, inv) == [1, 0, 0, 0]


def fold_pair(f, a, b, beta, x):
    """Line through (x, a) and (-x, b) evaluated at beta (ext4 a, b, beta)."""
    p = P[f]
    # a + (beta - x) * (b - a) / (-x - x)
    inv = pow((-2 * x) % p, p - 2, p)
    slope = [((b[k] - a[k]) * inv) % p for k in range(4)]
    bm = [(beta[0] - x) % p, beta[1], beta[2], beta[3]]
    t = ext_mul(f, bm, slope)
    return [(a[k] + t[k]) % p for k in range(4)]


def merkle_root(f, mats, sponge, compress):
    """Mixed-height commitment by the verify_batch definition: layer by layer,
    compress pairs, inject the hash of same-height rows."""
    heights = [len(m) for m in mats]
    H = max(heights)

    def rows_hash(hgt, i):
        vals = []
        for m in mats:
            if len(m) == hgt:
                vals.extend(m[i])
        return list(sponge(f, vals))

    layer = [rows_hash(H, i) for i in range(H)]
    layers = [layer]
    n = H
    while n > 1:
        n //= 2
        nxt = []
        for i in range(n):
            d = list(compress(f, layer[2 * i], layer[2 * i + 1]))
            if n in heights:
                d = list(compress(f, d, rows_hash(n, i)))
            nxt.append(d)
        layer = nxt
        layers.append(layer)
    return layer[0], layers


def verify_batch(f, heights, widths, root, index, rows, path, sponge, compress):
    """verify_batch of circuit/fri.rs:389-436 for power-of-two heights."""
    H = max(heights)
    order = sorted(range(len(heights)), key=lambda k: -heights[k])
    cur = list(sponge(f, [v for k in order if heights[k] == H for v in rows[k]]))
    h = H
    for j, sib in enumerate(path):
        cur = list(compress(f, sib, cur) if (index >> j) & 1 else compress(f, cur, sib))
        h //= 2
        if h in heights:
            cur = list(compress(f, cur, sponge(f, [v for k in order if heights[k] == h for v in rows[k]])))
    return [int(x) for x in cur] == [int(x) for x in root]


class PyChallenger:
    """Duplex rules of circuit/challenger.rs:89-227 around a given permutation."""

    def __init__(self, f, permute):
        self.f, self.permute = f, permute
        self.state, self.inp, self.out = [0] * 16, [], []

    def _duplex(self):
        for i, v in enumerate(self.inp):
            self.state[i] = v
        self.inp = []
        self.state = [int(x) for x in self.permute(self.f, self.state)]
        self.out = self.state[:8]

    def observe(self, v):
        self.out = []
        self.inp.append(int(v))
        if len(self.inp) == 8:
            self._duplex()

    def sample(self):
        if self.inp or not self.out:
            self._duplex()
        return self.out.pop()

    def sample_ext(self):
        return [self.sample() for _ in range(4)]

    def sample_bits(self, bits):
        return self.sample() & ((1 << bits) - 1)


def splitmix_element(f, seed, i):
    mask = (1 << 64) - 1
    p = P[f]
    x = 0
    for t in range(4):
        z = (seed + (4 * i + t + 1) * 0x9E3779B97F4A7C15) & mask
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        x = z >> 33
        if x < p:
            return x
    return x % p


def encode_fri_proof(f, commits, queries, final_poly, pow_witness, canonical):
    """Bincode 1 bytes of p3-fri's FriProof, written from the struct layout
    the reference's recursion verifier reads (FriProofVariable and friends,
    vm/src/compiler/recursion/circuit/types.rs:29-54): u64 little-endian
    sequence lengths, u32 words, fixed-size arrays and extension elements
    without a length.  All inputs canonical; `queries` is a list of
    (input_proof, commit_phase_openings) with input_proof a list of
    (rows per matrix, path digests) per round and commit_phase_openings a
    list of (sibling ext element, path digests).  Words are written
    canonical, or in Montgomery form (x 2^32 mod p) when canonical is False."""
    import struct
    p = P[f]

    def word(x):
        x = int(x) % p
        return struct.pack("<I", x if canonical else (x << 32) % p)

    def length(n):
        return struct.pack("<Q", n)

    def digests(ds):
        return length(len(ds)) + b"".join(word(x) for d in ds for x in d)

    out = [length(len(commits))] + [word(x) for c in commits for x in c]
    out.append(length(len(queries)))
    for input_proof, cp in queries:
        out.append(length(len(input_proof)))
        for rows, path in input_proof:
            out.append(length(len(rows)))
            for row in rows:
                out.append(length(len(row)) + b"".join(word(x) for x in row))
            out.append(digests(path))
        out.append(length(len(cp)))
        for sibling, path in cp:
            out.append(b"".join(word(x) for x in sibling) + digests(path))
    out.append(b"".join(word(x) for x in final_poly) + word(pow_witness))
    return b"".join(out)
