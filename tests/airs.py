"""AIRs for the quotient tests, each with a generator of valid traces
(canonical values, n rows).  Written with pico_amd.air.AirBuilder, whose
vocabulary is p3's AirBuilder as Pico's chips use it.

  toy        the reference's ToyChip (vm/src/chips/chips/toys/toy.rs:150-161):
             assert_bool(is_add); a + is_add b - (1 - is_add) b - result = 0.
             Degree 2, log quotient degree 0.
  fibonacci  first row (left, right) = (pv0, pv1); transition next.left =
             right, next.right = left + right; last row right = pv2.  Every
             selector and the next row; degree 2.
  cube       first row (x, y) = (pv0, pv1); z = x^2 y on every row;
             transition next.x = y, next.y = z; last row z = pv2.  Degree 3:
             log quotient degree 1, so the next row is 2 steps ahead on Q.
  gated      a preprocessed 0/1 column s: s (a - b) = 0 on every row and
             (1 - s) (next.a - a) = 0 on transitions.
"""
import numpy as np

from pico_amd.air import AirBuilder

P = {0: 2130706433, 1: 2013265921}


def toy_air(f):
    b = AirBuilder(f)
    a, bb, res, is_add = (b.main(i) for i in range(4))
    b.assert_bool(is_add)
    b.assert_zero(a + is_add * bb - (1 - is_add) * bb - res)
    return b, 0  # builder, log quotient degree


def toy_trace(f, n, rng):
    p = P[f]
    a = rng.integers(0, 256, n)
    bb = rng.integers(0, 256, n)
    is_add = rng.integers(0, 2, n)
    res = np.where(is_add == 1, a + bb, (a - bb) % p)
    t = np.stack([a, bb, res, is_add], 1).astype(np.uint64) % p
    t[n - n // 4:] = 0  # padding rows, as pad_to_power_of_two leaves them
    return t.astype(np.uint32), []


def fibonacci_air(f):
    b = AirBuilder(f)
    left, right = b.main(0), b.main(1)
    nleft, nright = b.main(0, next=True), b.main(1, next=True)
    b.when_first_row().assert_eq(left, b.public(0))
    b.when_first_row().assert_eq(right, b.public(1))
    b.when_transition().assert_eq(right, nleft)
    b.when_transition().assert_eq(left + right, nright)
    b.when_last_row().assert_eq(right, b.public(2))
    return b, 0


def fibonacci_trace(f, n, rng):
    p = P[f]
    x, y = int(rng.integers(0, p)), int(rng.integers(0, p))
    pv = [x, y]
    rows = []
    for _ in range(n):
        rows.append((x, y))
        x, y = y, (x + y) % p
    pv.append(rows[-1][1])
    return np.array(rows, np.uint32), pv


def cube_air(f):
    b = AirBuilder(f)
    x, y, z = b.main(0), b.main(1), b.main(2)
    b.when_first_row().assert_eq(x, b.public(0))
    b.when_first_row().assert_eq(y, b.public(1))
    b.assert_eq(z, x * x * y)
    b.when_transition().assert_eq(b.main(0, next=True), y)
    b.when_transition().assert_eq(b.main(1, next=True), z)
    b.when_last_row().assert_eq(z, b.public(2))
    return b, 1


def cube_trace(f, n, rng):
    p = P[f]
    x, y = int(rng.integers(0, p)), int(rng.integers(0, p))
    pv = [x, y]
    rows = []
    for _ in range(n):
        z = x * x % p * y % p
        rows.append((x, y, z))
        x, y = y, z
    pv.append(rows[-1][2])
    return np.array(rows, np.uint32), pv


def gated_air(f):
    b = AirBuilder(f)
    s, a, bb = b.preprocessed(0), b.main(0), b.main(1)
    b.assert_zero(s * (a - bb))
    b.when_transition().assert_zero((1 - s) * (b.main(0, next=True) - a))
    return b, 0


def gated_trace(f, n, rng):
    """(main, preprocessed, publics)"""
    p = P[f]
    s = rng.integers(0, 2, n)
    a = np.zeros(n, np.uint64)
    cur = int(rng.integers(0, p))
    for i in range(n):
        a[i] = cur
        if s[i] == 1:
            cur = int(rng.integers(0, p))  # a row with s = 1 may change a on the next row
    bb = np.where(s == 1, a, rng.integers(0, p, n).astype(np.uint64))
    return np.stack([a, bb], 1).astype(np.uint32), s.astype(np.uint32).reshape(n, 1)


def random_air(f, rng, width, pre_width, nconstraints, npublics=3):
    """Random constraints over `width` main and `pre_width` preprocessed
    columns: sums, differences and products of up to four leaves (columns at
    the local or next row spread over every column window, constants,
    public values, selectors), some under a when_* filter.  Degree <= 5 with
    a filter, so log quotient degree 2 covers it."""
    b = AirBuilder(f)
    p = P[f]

    def leaf():
        r = rng.integers(0, 10)
        if r < 6 or (r < 8 and not pre_width):
            return b.main(int(rng.integers(0, width)), next=bool(rng.integers(0, 2)))
        if r < 8:
            return b.preprocessed(int(rng.integers(0, pre_width)), next=bool(rng.integers(0, 2)))
        if r == 8:
            return b.public(int(rng.integers(0, npublics)))
        return b.constant(int(rng.integers(0, p)))

    for _ in range(nconstraints):
        e = leaf()
        for _ in range(int(rng.integers(0, 4))):
            op = rng.integers(0, 3)
            e = e * leaf() if op == 0 else e + leaf() if op == 1 else e - leaf()
        if rng.integers(0, 5) == 0:
            e = -e
        sel = rng.integers(0, 5)
        if sel == 0:
            b.when_first_row().assert_zero(e)
        elif sel == 1:
            b.when_last_row().assert_zero(e)
        elif sel == 2:
            b.when_transition().assert_zero(e)
        else:
            b.assert_zero(e)
    return b, 2


def lookup_air(f):
    """The permutation (LogUp) constraints Pico builds for a chip's lookups
    (vm/src/machine/permutation.rs:183-346), here for one batch of two
    interactions over main columns (a, b, m1, m2, g): a send of (kind 1, a)
    with multiplicity m1 and a receive of (kind 2, a, b) with multiplicity
    m2.  rlc_i = alpha + kind_i + beta a (+ beta^2 b) with (alpha, beta) the
    permutation challenges; the permutation trace has the batch entry and
    the running sum phi (two ext4 columns):
      rlc_1 rlc_2 entry = m1 rlc_2 - m2 rlc_1
      first row: phi = entry;  transition: phi' - phi = entry'
      last row: phi = regional cumulative sum
    plus the global-sum check of the last row, g = global sum coordinate 0.
    Degree 3 (log quotient degree 1)."""
    b = AirBuilder(f)
    a, bb, m1, m2, g = (b.main(i) for i in range(5))
    alpha, beta = b.perm_challenge(0), b.perm_challenge(1)
    entry, phi = b.permutation(0), b.permutation(1)
    entry_n, phi_n = b.permutation(0, next=True), b.permutation(1, next=True)
    rlc1 = alpha + b.constant(1) + beta * a
    rlc2 = alpha + b.constant(2) + beta * a + beta * beta * bb
    b.assert_zero(rlc1 * rlc2 * entry - (m1 * rlc2 - m2 * rlc1))
    b.when_first_row().assert_zero(phi - entry)
    b.when_transition().assert_zero(phi_n - phi - entry_n)
    b.when_last_row().assert_zero(phi - b.regional_sum())
    b.when_last_row().assert_eq(g, b.global_sum(0))
    return b, 1


def lookup_trace(f, n, rng, oracle):
    """(main n x 5, permutation n x 8 (ext4 columns entry, phi flattened),
    challenges [alpha, beta], regional sum, global sum (14 coordinates)),
    all canonical; a valid trace of lookup_air."""
    p = P[f]
    main = rng.integers(0, p, (n, 5)).astype(np.uint64)
    main[:, 2:4] = rng.integers(0, 5, (n, 2))  # small multiplicities
    alpha = [int(x) for x in rng.integers(0, p, 4)]
    beta = [int(x) for x in rng.integers(0, p, 4)]

    def ext(v):
        return [int(v) % p, 0, 0, 0]

    def add(x, y):
        return [(u + v) % p for u, v in zip(x, y)]

    def mul(x, y):
        return [int(t) for t in oracle.ext_mul(f, x, y)]

    beta2 = mul(beta, beta)
    perm = np.zeros((n, 8), np.uint64)
    phi = [0, 0, 0, 0]
    for i in range(n):
        a, bb, m1, m2 = (int(x) for x in main[i, :4])
        rlc1 = add(add(alpha, ext(1)), mul(beta, ext(a)))
        rlc2 = add(add(add(alpha, ext(2)), mul(beta, ext(a))), mul(beta2, ext(bb)))
        entry = add(mul(ext(m1), [int(t) for t in oracle.ext_inv(f, rlc1)]),
                    mul(ext(p - m2 if m2 else 0), [int(t) for t in oracle.ext_inv(f, rlc2)]))
        phi = add(phi, entry)
        perm[i, 0:4] = entry
        perm[i, 4:8] = phi
    gsum = [int(x) for x in rng.integers(0, p, 14)]
    gsum[0] = int(main[n - 1, 4])
    return main.astype(np.uint32), perm.astype(np.uint32), [alpha, beta], list(phi), gsum


def random_ext_air(f, rng, width, perm_cols, nconstraints):
    """Random extension-field constraints: sums, differences and products of
    up to four leaves drawn from main columns, permutation ext4 columns (local
    or next), two challenges, the regional sum, global-sum coordinates and
    constants, some under a when_* filter; every coordinate of the lowered
    extension arithmetic shows."""
    b = AirBuilder(f)
    p = P[f]

    def leaf():
        r = rng.integers(0, 10)
        if r < 3:
            return b.main(int(rng.integers(0, width)), next=bool(rng.integers(0, 2)))
        if r < 6:
            return b.permutation(int(rng.integers(0, perm_cols)), next=bool(rng.integers(0, 2)))
        if r == 6:
            return b.perm_challenge(int(rng.integers(0, 2)))
        if r == 7:
            return b.regional_sum()
        if r == 8:
            return b.global_sum(int(rng.integers(0, 14)))
        return b.constant(int(rng.integers(0, p)))

    for _ in range(nconstraints):
        e = leaf()
        for _ in range(int(rng.integers(0, 4))):
            op = rng.integers(0, 3)
            e = e * leaf() if op == 0 else e + leaf() if op == 1 else e - leaf()
        if rng.integers(0, 5) == 0:
            e = -e
        sel = rng.integers(0, 4)
        if sel == 0:
            b.when_first_row().assert_zero(e)
        elif sel == 1:
            b.when_transition().assert_zero(e)
        else:
            b.assert_zero(e)
    return b, 2


AIRS = {"toy": (toy_air, toy_trace), "fibonacci": (fibonacci_air, fibonacci_trace),
        "cube": (cube_air, cube_trace), "gated": (gated_air, None)}
