"""Fragment-state file formats of the reference and their conversion to the
layout the density build takes (occupation lists of every determinant in
lexical order, core included; coefficients n_states x n_configs).

* general-XRCC (general-XRCC/Be631g.py:36-52): Z_{N}e.npy holds the
  coefficients as (n_configs x n_states), configs_{N}e-stable.npy the
  occupation lists in lexical order: load_npy_states.
* hermitian-XRCC fixtures (tests/ref_data/check_{configs,states}_{m}_{chg}.pkl,
  written and read by hermitian-XRCC/mains/workflow.py:80-85): the configs
  as a list of bit-packed integers (bit p = spin orbital p, alpha orbitals
  first) in ascending integer order -- the colexicographic order of the
  occupation lists -- and the states as a list of 1-D coefficient arrays
  over those configs: load_packed_states, colex_to_lex.

read_pickled_data decodes such pickle files without executing anything
from them: it walks the opcode stream (pickletools.genops, a disassembler)
and accepts only lists, tuples, integers, strings, bytes, booleans, None
and the numpy array reconstruction pattern of numpy's own pickles, whose
global names it compares against a fixed list and emulates (it never
imports or calls them).  Any other opcode or global raises.
"""
import itertools
import math
import pickletools

import numpy as np


# ---------------------------------------------------------------------------
# configuration orderings
# ---------------------------------------------------------------------------

def packed_to_occupations(packed):
    """Ascending spin orbitals of a bit-packed determinant (bit p = spin orbital p)."""
    x, p, out = int(packed), 0, []
    while x:
        if x & 1:
            out.append(p)
        x >>= 1
        p += 1
    return out


def occupations_to_packed(occ):
    v = 0
    for p in occ:
        v |= 1 << int(p)
    return v


def lex_configs(n_orbs, n_elec, n_core=0):
    """Every determinant with the core spin orbitals (the first n_core
    spatial orbitals of each spin) filled, in lexical order over the valence
    spin orbitals (the order general-XRCC/density_tensors.c:29-62 ranks)."""
    core = list(range(n_core)) + list(range(n_orbs, n_orbs + n_core))
    valence = [p for p in range(2 * n_orbs) if p not in core]
    n_val = n_elec - 2 * n_core
    if n_val < 0 or n_val > len(valence):
        return np.zeros((0, n_elec), dtype=np.int64)
    out = np.empty((math.comb(len(valence), n_val), n_elec), dtype=np.int64)
    for k, combo in enumerate(itertools.combinations(valence, n_val)):
        out[k] = sorted(core + list(combo))
    return out


def colex_to_lex(packed_configs, coeffs, n_orbs, n_core=0):
    """(configs, coeffs) in the lexical layout from bit-packed configs in
    any order (the reference fixtures use ascending integers, i.e. colex)
    and coefficients (n_states x len(packed_configs)).  Determinants the
    list lacks get zero coefficients, so the result spans the complete
    space; a determinant that leaves a core orbital empty, or a repeated
    one, raises."""
    coeffs = np.asarray(coeffs, dtype=np.float64)
    occs = [packed_to_occupations(x) for x in packed_configs]
    n_elec = {len(o) for o in occs}
    if len(n_elec) != 1:
        raise ValueError("configs with different electron counts")
    n_elec = n_elec.pop()
    lex = lex_configs(n_orbs, n_elec, n_core)
    where = {occupations_to_packed(row): k for k, row in enumerate(lex)}
    out = np.zeros((coeffs.shape[0], lex.shape[0]))
    seen = set()
    for j, packed in enumerate(packed_configs):
        k = where.get(int(packed))
        if k is None:
            raise ValueError(f"config {packed:#x} is not a determinant of this space (core {n_core})")
        if k in seen:
            raise ValueError(f"config {packed:#x} repeats")
        seen.add(k)
        out[:, k] = coeffs[:, j]
    return lex, out


def lex_to_colex(configs, coeffs):
    """Inverse direction: bit-packed configs in ascending order (colex) and
    the coefficient columns in that order."""
    packed = np.array([occupations_to_packed(row) for row in np.asarray(configs)], dtype=np.int64)
    order = np.argsort(packed, kind="stable")
    return packed[order].tolist(), np.asarray(coeffs)[:, order]


# ---------------------------------------------------------------------------
# loaders
# ---------------------------------------------------------------------------

class _States:
    """.configs (n_configs x n_elec occupation lists, lexical) and .coeffs
    (n_states x n_configs), the objects build_density_tensors reads."""

    def __init__(self, configs, coeffs):
        self.configs = configs
        self.coeffs = coeffs


def load_npy_states(states_file, configs_file, n_states=None):
    """general-XRCC/Be631g.py:_load_states: Z_{N}e.npy is (n_configs x
    n_states); configs_{N}e-stable.npy the occupation lists.  n_states
    keeps the first n of them."""
    coeffs = np.load(states_file, allow_pickle=False).T
    configs = np.load(configs_file, allow_pickle=False)
    if n_states is not None:
        coeffs = coeffs[:int(n_states)]
    return _States(np.ascontiguousarray(configs, dtype=np.int64), np.ascontiguousarray(coeffs))


def load_packed_states(states_path, configs_path, n_orbs, n_core):
    """The check_states / check_configs pair of one fragment and charge,
    converted to the lexical layout."""
    packed = read_pickled_data(configs_path)
    vectors = read_pickled_data(states_path)
    coeffs = np.stack([np.asarray(v, dtype=np.float64) for v in vectors])
    configs, coeffs = colex_to_lex(packed, coeffs, n_orbs, n_core)
    return _States(configs, coeffs)


# ---------------------------------------------------------------------------
# pickle data without execution
# ---------------------------------------------------------------------------

_ARRAY_GLOBALS = {("numpy._core.multiarray", "_reconstruct"), ("numpy.core.multiarray", "_reconstruct"),
                  ("numpy", "ndarray"), ("numpy", "dtype"), ("numpy._core.numeric", "_frombuffer"),
                  ("numpy.core.numeric", "_frombuffer"), ("_codecs", "encode")}


class _Global:
    def __init__(self, module, name):
        if (module, name) not in _ARRAY_GLOBALS:
            raise ValueError(f"pickle names {module}.{name}, which is not array data")
        self.name = name


class _ArraySkeleton:
    """The empty array numpy's _reconstruct stands for, filled by BUILD."""

    def __init__(self):
        self.array = None


class _DType:
    def __init__(self, code):
        self.code = code
        self.dtype = None


def _reduce(func, args):
    if not isinstance(func, _Global):
        raise ValueError("REDUCE of something other than a whitelisted array constructor")
    if func.name == "_reconstruct":
        return _ArraySkeleton()
    if func.name == "dtype":
        return _DType(args[0])
    if func.name == "encode":            # protocol 2 bytes: (latin-1 text, 'latin1')
        text, codec = args
        if codec != "latin1" or not isinstance(text, str):
            raise ValueError("unexpected encoded bytes")
        return text.encode("latin1")
    if func.name == "_frombuffer":       # protocol 5 arrays: (buffer, dtype, shape, order)
        buf, dt, shape, order = args
        if not isinstance(dt, _DType) or dt.dtype is None or not isinstance(buf, (bytes, bytearray)):
            raise ValueError("array without a plain dtype or buffer")
        sk = _ArraySkeleton()
        sk.array = np.frombuffer(bytes(buf), dtype=dt.dtype).copy().reshape(shape, order=order)
        return sk
    raise ValueError(f"REDUCE of {func.name}")


def _build(obj, state):
    if isinstance(obj, _DType):
        # (version, byte order, subarray, names, fields, itemsize, alignment, flags)
        order = state[1] if state[1] in ("<", ">", "|", "=") else "<"
        obj.dtype = np.dtype(obj.code).newbyteorder(order) if order in "<>" else np.dtype(obj.code)
        return obj
    if isinstance(obj, _ArraySkeleton):
        version, shape, dt, fortran, raw = state
        if not isinstance(dt, _DType) or dt.dtype is None:
            raise ValueError("array without a plain dtype")
        if not isinstance(raw, (bytes, bytearray)):
            raise ValueError("array data is not raw bytes")
        a = np.frombuffer(bytes(raw), dtype=dt.dtype).copy()
        a = a.reshape(shape, order="F" if fortran else "C")
        obj.array = a
        return obj
    raise ValueError("BUILD on an unexpected object")


def _finish(x):
    if isinstance(x, _ArraySkeleton):
        return x.array
    if isinstance(x, list):
        return [_finish(v) for v in x]
    if isinstance(x, tuple):
        return tuple(_finish(v) for v in x)
    return x


def read_pickled_data(path):
    """The data of a pickle holding lists / tuples of ints, floats, strings,
    bytes and numpy arrays, decoded opcode by opcode; nothing in the file is
    imported or called."""
    data = open(path, "rb").read()
    stack, memo, marks = [], {}, []
    for op, arg, _ in pickletools.genops(data):
        n = op.name
        if n in ("PROTO", "FRAME"):
            continue
        if n == "STOP":
            break
        if n == "MARK":
            marks.append(len(stack))
        elif n in ("EMPTY_LIST",):
            stack.append([])
        elif n in ("EMPTY_TUPLE",):
            stack.append(())
        elif n in ("BININT", "BININT1", "BININT2", "INT", "LONG1", "LONG", "BINFLOAT", "FLOAT"):
            stack.append(arg)
        elif n in ("SHORT_BINUNICODE", "BINUNICODE", "UNICODE", "BINUNICODE8"):
            stack.append(arg)
        elif n in ("SHORT_BINBYTES", "BINBYTES", "BINBYTES8", "BYTEARRAY8"):
            stack.append(bytes(arg))
        elif n == "NONE":
            stack.append(None)
        elif n == "NEWTRUE":
            stack.append(True)
        elif n == "NEWFALSE":
            stack.append(False)
        elif n == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif n in ("BINPUT", "LONG_BINPUT", "PUT"):
            memo[arg] = stack[-1]
        elif n in ("BINGET", "LONG_BINGET", "GET"):
            stack.append(memo[arg])
        elif n == "APPEND":
            v = stack.pop()
            stack[-1].append(v)
        elif n == "APPENDS":
            m = marks.pop()
            items = stack[m:]
            del stack[m:]
            stack[-1].extend(items)
        elif n == "TUPLE":
            m = marks.pop()
            t = tuple(stack[m:])
            del stack[m:]
            stack.append(t)
        elif n in ("TUPLE1", "TUPLE2", "TUPLE3"):
            k = int(n[-1])
            t = tuple(stack[-k:])
            del stack[-k:]
            stack.append(t)
        elif n == "STACK_GLOBAL":
            name = stack.pop()
            module = stack.pop()
            stack.append(_Global(module, name))
        elif n == "GLOBAL":
            module, name = arg.split(" ")
            stack.append(_Global(module, name))
        elif n == "REDUCE":
            args = stack.pop()
            func = stack.pop()
            stack.append(_reduce(func, args))
        elif n == "BUILD":
            state = stack.pop()
            stack.append(_build(stack.pop(), state))
        else:
            raise ValueError(f"pickle opcode {n} is not plain data")
    if len(stack) != 1:
        raise ValueError("malformed pickle data")
    return _finish(stack[0])
