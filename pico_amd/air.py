"""AIR constraints for the device quotient evaluator (pg_air_compile /
pg_quotient_values).

An AirBuilder collects constraints the way Pico's SymbolicConstraintFolder
does (vm/src/machine/folder.rs:30-120): expressions over main and
preprocessed columns at the local and next row, public values, constants and
the three selectors, each assert_zero appending one constraint (the fold
order).  The permutation side (PermutationBuilder, folder.rs:246-276) adds
extension-valued leaves: the permutation trace's ext4 columns, the
permutation challenges and the regional cumulative sum, plus the global
cumulative sum's coordinates; an expression with an extension-valued
operand is extension-valued and its assert is assert_zero_ext.
Expressions are hash-consed into the node DAG the C ABI takes (one
pg_air_node per distinct expression, operands first).  assert_bool and the
when_* filters follow p3's AirBuilder: assert_bool(x) asserts x (x - 1), and
a filtered assert multiplies the constraint by its condition.
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .core import PRIMES, DeviceMatrix, _ptr, from_monty, to_monty

(MAIN, PREPROCESSED, PUBLIC, IS_FIRST_ROW, IS_LAST_ROW, IS_TRANSITION, CONSTANT, ADD, SUB, NEG, MUL, PERMUTATION,
 PERM_CHALLENGE, REGIONAL_SUM, GLOBAL_SUM) = range(15)


class Expr:
    __slots__ = ("b", "i")

    def __init__(self, builder, index):
        self.b, self.i = builder, index

    def _lift(self, other):
        return other if isinstance(other, Expr) else self.b.constant(other)

    def __add__(self, o):
        return self.b._node(ADD, self.i, self._lift(o).i)

    def __radd__(self, o):
        return self._lift(o) + self

    def __sub__(self, o):
        return self.b._node(SUB, self.i, self._lift(o).i)

    def __rsub__(self, o):
        return self._lift(o) - self

    def __mul__(self, o):
        return self.b._node(MUL, self.i, self._lift(o).i)

    def __rmul__(self, o):
        return self._lift(o) * self

    def __neg__(self):
        return self.b._node(NEG, self.i, 0)


class _Filtered:
    """assert_* under a condition: the constraint is multiplied by it."""

    def __init__(self, builder, cond):
        self.b, self.cond = builder, cond

    def assert_zero(self, x):
        self.b.assert_zero(self.cond * x)

    def assert_eq(self, x, y):
        self.assert_zero(self.b._lift(x) - y)

    def assert_bool(self, x):
        x = self.b._lift(x)
        self.assert_zero(x * (x - 1))


class AirBuilder:
    def __init__(self, field):
        self.field = field
        self.nodes = []  # (kind, a, b)
        self._memo = {}
        self.asserts = []

    def _node(self, kind, a, b):
        key = (kind, a, b)
        if key not in self._memo:
            self._memo[key] = len(self.nodes)
            self.nodes.append(key)
        return Expr(self, self._memo[key])

    def _lift(self, x):
        return x if isinstance(x, Expr) else self.constant(x)

    # ---- leaves -----------------------------------------------------------
    def main(self, col, next=False):
        return self._node(MAIN, int(bool(next)), col)

    def preprocessed(self, col, next=False):
        return self._node(PREPROCESSED, int(bool(next)), col)

    def public(self, index):
        return self._node(PUBLIC, 0, index)

    def constant(self, value):
        return self._node(CONSTANT, int(value) % PRIMES[self.field], 0)

    def permutation(self, col, next=False):
        """ext4 column `col` of the permutation trace (base columns 4 col..4 col + 3)"""
        return self._node(PERMUTATION, int(bool(next)), col)

    def perm_challenge(self, index):
        return self._node(PERM_CHALLENGE, 0, index)

    def regional_sum(self):
        return self._node(REGIONAL_SUM, 0, 0)

    def global_sum(self, coordinate):
        """coordinate 0..6 of the global cumulative sum's x, 7..13 of its y"""
        return self._node(GLOBAL_SUM, 0, coordinate)

    @property
    def is_first_row(self):
        return self._node(IS_FIRST_ROW, 0, 0)

    @property
    def is_last_row(self):
        return self._node(IS_LAST_ROW, 0, 0)

    @property
    def is_transition(self):
        return self._node(IS_TRANSITION, 0, 0)

    # ---- constraints --------------------------------------------------------
    def assert_zero(self, x):
        self.asserts.append(self._lift(x).i)

    def assert_eq(self, x, y):
        self.assert_zero(self._lift(x) - y)

    def assert_bool(self, x):
        x = self._lift(x)
        self.assert_zero(x * (x - 1))

    def when(self, cond):
        return _Filtered(self, self._lift(cond))

    def when_first_row(self):
        return _Filtered(self, self.is_first_row)

    def when_last_row(self):
        return _Filtered(self, self.is_last_row)

    def when_transition(self):
        return _Filtered(self, self.is_transition)

    # ---- the C ABI form -----------------------------------------------------
    def node_array(self):
        return np.ascontiguousarray(np.array(self.nodes, dtype=np.uint32).reshape(-1, 3))

    def compile(self):
        return AirProgram(self.field, self.node_array(), np.ascontiguousarray(self.asserts, dtype=np.uint32))


class AirProgram:
    """pg_air_compile's register program (freed with the object)."""

    def __init__(self, field, nodes, asserts):
        self.field = field
        self.nodes, self.asserts = nodes, asserts
        h = ctypes.c_void_p()
        nregs, ninstr = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().pg_air_compile(field, _ptr(nodes) if nodes.size else None, len(nodes),
                                   _ptr(asserts) if asserts.size else None, len(asserts), ctypes.byref(h),
                                   ctypes.byref(nregs), ctypes.byref(ninstr)))
        self.handle, self.nregs, self.ninstr = h, nregs.value, ninstr.value

    def free(self):
        if self.handle:
            lib().pg_air_program_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def quotient_values(ctx, program, main_data, main_idx, log_blowup, log_quotient_degree, alpha, publics=(),
                    pre_data=None, pre_idx=0, perm_data=None, perm_idx=0, challenges=(), regional_sum=None,
                    global_sum=()):
    """compute_quotient_values on the device: the |Q| x 4 quotient (ext4,
    natural order on the quotient coset) as a new DeviceMatrix, from the
    committed main (and preprocessed, permutation) LDEs.  alpha, each
    challenge and regional_sum: canonical ext4; publics, global_sum:
    canonical values."""
    f = program.field

    def mont(vals, shape=None):
        if vals is None or len(vals) == 0:
            return None
        return np.ascontiguousarray(to_monty(f, np.asarray(vals, np.uint64)).reshape(-1))

    pub, al = mont(publics), mont(alpha)
    chal = mont(np.asarray(challenges, np.uint64).reshape(-1)) if len(challenges) else None
    rsum, gsum = mont(regional_sum), mont(global_sum)
    h = ctypes.c_void_p()
    check(lib().pg_quotient_values_perm(
        ctx.handle, program.handle, main_data.handle, main_idx, pre_data.handle if pre_data is not None else None,
        pre_idx, perm_data.handle if perm_data is not None else None, perm_idx, log_blowup, log_quotient_degree,
        _ptr(pub) if pub is not None else None, len(publics), _ptr(chal) if chal is not None else None,
        len(challenges), _ptr(rsum) if rsum is not None else None, _ptr(gsum) if gsum is not None else None,
        len(global_sum), _ptr(al), ctypes.byref(h)))
    return DeviceMatrix(ctx, h, owned=True)


__all__ = ["AirBuilder", "AirProgram", "Expr", "quotient_values", "from_monty"]
