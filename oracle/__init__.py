"""ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, as the checker or as the reported CPU
baseline.  The product never imports it.

The functions restate the reference's algorithms; see lcs_oracle.c for the
reference lines each follows and how each is pinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libpermref.so")

_i32p = ctypes.POINTER(ctypes.c_int32)


def build(with_ref: bool = True) -> None:
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    if with_ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build(with_ref=False)
    lib = ctypes.CDLL(_LIB_PATH)
    i32 = ctypes.c_int32
    lib.orc_comb_antidiag.argtypes = [_i32p, i32, _i32p, i32, _i32p, _i32p]
    lib.orc_comb_rows.argtypes = [_i32p, i32, _i32p, i32, _i32p, _i32p]
    lib.orc_comb_mt.argtypes = [_i32p, i32, _i32p, i32, _i32p, _i32p, i32]
    lib.orc_comb_simd.argtypes = [_i32p, i32, _i32p, i32, _i32p, _i32p, i32, i32]
    lib.orc_comb_simd.restype = i32
    lib.orc_fixup.argtypes = [_i32p, _i32p, _i32p, _i32p, i32, i32, i32, i32]
    lib.orc_from_strands.argtypes = [_i32p, i32, _i32p, i32, _i32p, _i32p]
    lib.orc_hash.argtypes = [_i32p, i32]
    lib.orc_hash.restype = ctypes.c_int64
    lib.orc_llcs_from_strands.argtypes = [_i32p, i32, i32]
    lib.orc_llcs_from_strands.restype = i32
    lib.orc_string_substring.argtypes = [_i32p, i32, i32, _i32p]
    lib.orc_llcs_dp.argtypes = [_i32p, i32, _i32p, i32]
    lib.orc_llcs_dp.restype = i32
    lib.orc_llcs_bitpar.argtypes = [_i32p, i32, _i32p, i32]
    lib.orc_llcs_bitpar.restype = i32
    lib.orc_semilocal.argtypes = [_i32p, i32, i32, i32, _i32p, _i32p, i32, _i32p]
    lib.orc_staggered_multiply.argtypes = [_i32p, i32, _i32p, i32, i32, i32, _i32p, _i32p]
    lib.orc_staggered_multiply.restype = i32
    lib.orc_sticky_bruteforce.argtypes = [_i32p, _i32p, i32, _i32p]
    lib.orc_steady_ant.argtypes = [_i32p, _i32p, i32, _i32p]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _c(x: np.ndarray):
    return x.ctypes.data_as(_i32p)


def _arr(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


ISA_AUTO, ISA_AVX2, ISA_AVX512, ISA_SCALAR = 0, 1, 2, 3


def simd_isa_available(isa: int) -> bool:
    """Whether this host can run orc_comb_simd with `isa`."""
    h, v = np.empty(1, np.int32), np.empty(1, np.int32)
    a = np.zeros(1, np.int32)
    return lib().orc_comb_simd(_c(a), 1, _c(a), 1, _c(h), _c(v), 1, isa) >= 0


def comb(a, b, order: str = "rows", threads: int = 0, isa: int = ISA_AUTO):
    """Strands (h_out[m], v_out[n]) of the seaweed comb of a against b.
    order: "rows" (sequential), "antidiag" (the reference's order), "mt"
    (scalar block wavefront) or "simd" (vectorised stripes, `isa`)."""
    a, b = _arr(a), _arr(b)
    m, n = a.size, b.size
    h = np.empty(max(m, 1), np.int32)
    v = np.empty(max(n, 1), np.int32)
    if order == "antidiag":
        lib().orc_comb_antidiag(_c(a), m, _c(b), n, _c(h), _c(v))
    elif order == "mt":
        lib().orc_comb_mt(_c(a), m, _c(b), n, _c(h), _c(v), threads or os.cpu_count() or 1)
    elif order == "simd":
        used = lib().orc_comb_simd(_c(a), m, _c(b), n, _c(h), _c(v), threads or os.cpu_count() or 1, isa)
        if used < 0:
            raise RuntimeError(f"orc_comb_simd: this host cannot run isa {isa}")
    else:
        lib().orc_comb_rows(_c(a), m, _c(b), n, _c(h), _c(v))
    return h[:m].copy(), v[:n].copy()


def fixup(a_prepared, b_prepared, h, v, i0, isize, j0, jsize):
    """In-place comb of a sub-rectangle (Lcs_Semi_Fixup semantics)."""
    ap, bp = _arr(a_prepared), _arr(b_prepared)
    assert h.dtype == np.int32 and v.dtype == np.int32
    lib().orc_fixup(_c(ap), _c(bp), _c(h), _c(v), i0, isize, j0, jsize)


def from_strands(h, v):
    h, v = _arr(h), _arr(v)
    size = h.size + v.size
    r2c = np.empty(max(size, 1), np.int32)
    c2r = np.empty(max(size, 1), np.int32)
    lib().orc_from_strands(_c(h), h.size, _c(v), v.size, _c(r2c), _c(c2r))
    return r2c[:size].copy(), c2r[:size].copy()


def perm_hash(col_to_row) -> int:
    c2r = _arr(col_to_row)
    return int(lib().orc_hash(_c(c2r), c2r.size))


def llcs_from_strands(v, m: int) -> int:
    v = _arr(v)
    return int(lib().orc_llcs_from_strands(_c(v), m, v.size))


def string_substring(row_to_col, m: int, n: int) -> np.ndarray:
    r2c = _arr(row_to_col)
    out = np.empty((n + 1) * (n + 1), np.int32)
    lib().orc_string_substring(_c(r2c), m, n, _c(out))
    return out.reshape(n + 1, n + 1)


def llcs_dp(a, b) -> int:
    a, b = _arr(a), _arr(b)
    return int(lib().orc_llcs_dp(_c(a), a.size, _c(b), b.size))


def llcs_bitpar(a, b) -> int:
    """LCS length by the word-parallel recurrence (64-bit words, one thread)."""
    a, b = _arr(a), _arr(b)
    return int(lib().orc_llcs_bitpar(_c(a), a.size, _c(b), b.size))


def semilocal(row_to_col, m: int, n: int, kind: int, q0, q1) -> np.ndarray:
    """Quadrant queries (kind 0..4, see lcs_oracle.h) by direct count."""
    p, q0, q1 = _arr(row_to_col), _arr(q0), _arr(q1)
    out = np.empty(max(q0.size, 1), np.int32)
    lib().orc_semilocal(_c(p), m, n, kind, _c(q0), _c(q1), q0.size, _c(out))
    return out[: q0.size].copy()


def staggered_multiply(p_r2c, q_r2c, k: int, row_glue: bool):
    p, q = _arr(p_r2c), _arr(q_r2c)
    size = p.size + q.size - k
    out = np.empty(max(size, 1), np.int32)
    c2r = np.empty(max(size, 1), np.int32)
    lib().orc_staggered_multiply(_c(p), p.size, _c(q), q.size, k, int(bool(row_glue)), _c(out), _c(c2r))
    return out[:size].copy()


def steady_ant(p_r2c, q_r2c):
    p, q = _arr(p_r2c), _arr(q_r2c)
    out = np.empty(max(p.size, 1), np.int32)
    lib().orc_steady_ant(_c(p), _c(q), p.size, _c(out))
    return out[: p.size].copy()


def sticky_bruteforce(p_r2c, q_r2c):
    p, q = _arr(p_r2c), _arr(q_r2c)
    out = np.empty(max(p.size, 1), np.int32)
    lib().orc_sticky_bruteforce(_c(p), _c(q), p.size, _c(out))
    return out[: p.size].copy()


def permutation(a, b, order: str = "rows"):
    """row_to_col of P(a, b)."""
    h, v = comb(a, b, order)
    return from_strands(h, v)[0]


# ---------------------------------------------------------------- _ref
_ref = None


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref_from_strands_hash(h, v):
    """FromStrands + hash computed by the reference's own PermutationMatrix."""
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(_REF_PATH)
        _ref.ref_perm_from_strands_hash.argtypes = [_i32p, ctypes.c_int32, _i32p, ctypes.c_int32, _i32p, _i32p]
        _ref.ref_perm_from_strands_hash.restype = ctypes.c_int64
    h, v = _arr(h).copy(), _arr(v).copy()
    size = h.size + v.size
    r2c = np.empty(max(size, 1), np.int32)
    c2r = np.empty(max(size, 1), np.int32)
    hv = int(_ref.ref_perm_from_strands_hash(_c(h), h.size, _c(v), v.size, _c(r2c), _c(c2r)))
    return r2c[:size].copy(), c2r[:size].copy(), hv
