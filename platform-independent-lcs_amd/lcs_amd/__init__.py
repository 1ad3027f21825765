"""Python bindings (ctypes) of the MI355X semi-local LCS engine.

The engine itself is the C++/HIP library ``lib/liblcs_hip.so`` built from
``csrc/`` for gfx950; this module only marshals numpy arrays through its C ABI
(``include/lcs_hip.h``).  There is no fallback: when the library is missing
or no GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "liblcs_hip.so")
# LCS_HIP_LIB names another build of the same library (A/B timing of kernel builds)
LIB_PATH = os.environ.get("LCS_HIP_LIB", LIB_PATH)

LCS_OK = 0
LCS_ETIMEDOUT = -110
DEFAULT_TIMEOUT_TICKS = 500_000_000   # 5 s of the 100 MHz device clock
ERRORS = {
    -2: "ENOENT", -5: "EIO", -11: "EAGAIN", -12: "ENOMEM", -16: "EBUSY", -22: "EINVAL",
    -34: "ERANGE", -110: "ETIMEDOUT", -125: "ECANCELED",
}

VARIANT_AUTO, VARIANT_CHAIN, VARIANT_CHAIN_LDS, VARIANT_PACKED16, VARIANT_CHAIN_BYTES = 0, 1, 2, 3, 4
TUNE_SINGLE, TUNE_BATCH = 0, 1
SCHED_FROM = {0: "default", 1: "table", 2: "model", 3: "autotune", 4: "caller"}
SEMI_STRING_SUBSTRING, SEMI_SUBSTRING_STRING, SEMI_PREFIX_SUFFIX, SEMI_SUFFIX_PREFIX, SEMI_DOMINANCE = 0, 1, 2, 3, 4

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class LcsError(RuntimeError):
    def __init__(self, fn: str, rc: int):
        super().__init__(f"{fn} failed: {rc} ({ERRORS.get(rc, '?')})")
        self.rc = rc


class Schedule(ctypes.Structure):
    _fields_ = [("tile_m", ctypes.c_int32), ("waves_per_simd", ctypes.c_int32),
                ("variant", ctypes.c_int32), ("strips", ctypes.c_int32)]

    def __repr__(self):
        return (f"Schedule(tile_m={self.tile_m}, waves_per_simd={self.waves_per_simd}, "
                f"variant={self.variant}, strips={self.strips})")


def build() -> None:
    """Compile the HIP library in-tree (hipcc cross-compiles for gfx950)."""
    subprocess.run(["make", "-s", "-j8", "-C", PKG_ROOT], check=True)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with make -C {PKG_ROOT}")
    L = ctypes.CDLL(LIB_PATH)
    i32, i64, vp, sz = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t
    sched_p = ctypes.POINTER(Schedule)
    L.lcs_ctx_create.argtypes = [ctypes.c_int, sz, ctypes.POINTER(vp)]
    L.lcs_ctx_destroy.argtypes = [vp]
    L.lcs_ctx_device_name.argtypes = [vp, ctypes.c_char_p, sz]
    L.lcs_ctx_info.argtypes = [vp, _i32p, ctypes.POINTER(ctypes.c_double)]
    L.lcs_ctx_status.argtypes = [vp]
    L.lcs_ctx_chain_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.lcs_ctx_set_timeout.argtypes = [vp, ctypes.c_uint64]
    L.lcs_ctx_memory.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)]
    L.lcs_comb.argtypes = [vp, _i32p, i32, _i32p, i32, sched_p, _i32p, _i32p, _i32p]
    L.lcs_comb_permutation.argtypes = [vp, _i32p, i32, _i32p, i32, sched_p, _i32p, _i32p, _i32p]
    L.lcs_comb_device.argtypes = [vp, vp, i32, vp, i32, sched_p, vp, vp, vp]
    L.lcs_comb_fixup.argtypes = [vp, _i32p, i32, _i32p, i32, _i32p, _i32p, i32, i32, i32, i32]
    L.lcs_permutation_from_strands.argtypes = [_i32p, i32, _i32p, i32, _i32p, _i32p]
    L.lcs_permutation_hash.argtypes = [_i32p, i32]
    L.lcs_permutation_hash.restype = i64
    L.lcs_braid_multiply.argtypes = [vp, _i32p, i32, _i32p, i32, i32, i32, _i32p]
    L.lcs_batch_comb.argtypes = [vp, i32, _i32p, _i64p, _i32p, _i32p, _i64p, _i32p, sched_p, i32, _i32p, _i32p]
    L.lcs_batch_comb_device.argtypes = [vp, i32, vp, _i64p, _i32p, vp, _i64p, _i32p, sched_p, i32, vp, vp]
    L.lcs_debug_band_trace.argtypes = [vp, i32, vp]
    f32p = ctypes.POINTER(ctypes.c_float)
    L.lcs_mlp_create.argtypes = [vp, i32, i32, i32, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.lcs_mlp_destroy.argtypes = [vp]
    L.lcs_mlp_param_count.argtypes = [vp, _i32p]
    L.lcs_mlp_get_params.argtypes = [vp, f32p]
    L.lcs_mlp_set_params.argtypes = [vp, f32p]
    L.lcs_mlp_predict.argtypes = [vp, f32p, i32, f32p]
    L.lcs_mlp_loss_grad.argtypes = [vp, f32p, f32p, f32p, i32, f32p, f32p]
    L.lcs_mlp_train.argtypes = [vp, f32p, f32p, f32p, i32, i32, ctypes.c_float, f32p]
    L.lcs_mlp_choose.argtypes = [vp, f32p, i32, _i32p]
    L.lcs_autotune_measure.argtypes = [vp, i32, i32, i32, ctypes.c_float, ctypes.c_uint64, sched_p, i32, i32, i32,
                                       f32p]
    L.lcs_permutation_dominance.argtypes = [vp, _i32p, i32, _i32p, _i32p, i32, _i32p]
    L.lcs_semilocal_string_substring.argtypes = [vp, _i32p, i32, i32, _i32p, _i32p, i32, _i32p]
    L.lcs_llcs.argtypes = [vp, _i32p, i32, _i32p, i32, i32, _i32p]
    L.lcs_llcs_device.argtypes = [vp, vp, i32, vp, i32, i32, vp]
    L.lcs_llcs_batch.argtypes = [vp, i32, _i32p, _i64p, _i32p, _i32p, _i64p, _i32p, i32, _i32p]
    L.lcs_semilocal_query.argtypes = [vp, _i32p, i32, i32, i32, _i32p, _i32p, i32, _i32p]
    L.lcs_dominance_index_build.argtypes = [vp, _i32p, i32, ctypes.POINTER(vp)]
    L.lcs_dominance_index_build_device.argtypes = [vp, vp, i32, ctypes.POINTER(vp)]
    L.lcs_dominance_index_query.argtypes = [vp, i32, i32, i32, _i32p, _i32p, i32, _i32p]
    L.lcs_dominance_index_query_device.argtypes = [vp, i32, i32, i32, vp, vp, i32, vp]
    L.lcs_dominance_index_bytes.argtypes = [vp]
    L.lcs_dominance_index_bytes.restype = i64
    L.lcs_dominance_index_destroy.argtypes = [vp]
    L.lcs_dominance_index_destroy.restype = None
    L.lcs_server_create.argtypes = [vp, sz, i32, ctypes.POINTER(vp)]
    L.lcs_server_destroy.argtypes = [vp]
    L.lcs_server_submit.argtypes = [vp, _i32p, i32, _i32p, i32, sched_p, f32p, _i64p]
    L.lcs_server_poll.argtypes = [vp, i64, _i32p, _i32p, _i32p]
    L.lcs_server_wait.argtypes = [vp, i64, _i32p, _i32p, i32]
    L.lcs_server_cancel.argtypes = [vp, i64]
    L.lcs_server_release.argtypes = [vp, i64]
    L.lcs_server_request_times.argtypes = [vp, i64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.lcs_server_stats.argtypes = [vp, _i64p, _i64p, _i64p, _i32p, ctypes.POINTER(sz)]
    L.lcs_server_set_cost_model.argtypes = [vp, vp, sched_p, i32]
    L.lcs_ctx_add_tuned.argtypes = [vp, i32, i32, i32, sched_p]
    L.lcs_ctx_clear_tuned.argtypes = [vp, i32]
    L.lcs_ctx_set_cost_model.argtypes = [vp, i32, vp, sched_p, i32]
    L.lcs_ctx_set_autotune_on_miss.argtypes = [vp, i32, sched_p, i32, i32]
    L.lcs_ctx_choose_schedule.argtypes = [vp, i32, i32, i32, i32, ctypes.c_float, sched_p, _i32p]
    # workloads
    L.lcs_gen_binary_mt19937.argtypes = [i32, i32, _i32p]
    L.lcs_gen_uniform.argtypes = [ctypes.c_uint64, i32, i32, _i32p]
    L.lcs_gen_zipf.argtypes = [ctypes.c_uint64, i32, i32, ctypes.c_double, _i32p]
    L.lcs_gen_mutate.argtypes = [ctypes.c_uint64, _i32p, i32, i32, ctypes.c_double, ctypes.c_double, i32, _i32p]
    L.lcs_gen_planted.argtypes = [ctypes.c_uint64, _i32p, i32, i32, i32, ctypes.c_double, i32,
                                  ctypes.c_double, _i32p]
    L.lcs_gen_unit.argtypes = [ctypes.c_uint64, i64]
    L.lcs_gen_unit.restype = ctypes.c_double
    _lib = L
    return L


def _p(x: np.ndarray):
    return x.ctypes.data_as(_i32p)


def _i32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _check(fn: str, rc: int) -> None:
    if rc != LCS_OK:
        raise LcsError(fn, rc)


def schedule(tile_m=0, waves_per_simd=0, variant=0, strips=0) -> Schedule:
    return Schedule(tile_m, waves_per_simd, variant, strips)


class Context:
    """An engine context on one GPU (lcs_ctx_create / lcs_ctx_destroy)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._h = ctypes.c_void_p()
        _check("lcs_ctx_create", lib().lcs_ctx_create(device, hbm_budget, ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().lcs_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def device_name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _check("lcs_ctx_device_name", lib().lcs_ctx_device_name(self._h, buf, 256))
        return buf.value.decode()

    def info(self):
        cus = ctypes.c_int32()
        ms = ctypes.c_double()
        _check("lcs_ctx_info", lib().lcs_ctx_info(self._h, ctypes.byref(cus), ctypes.byref(ms)))
        return cus.value, ms.value

    def chain_ms(self) -> float:
        """lcs_ctx_chain_ms: the band-chain launch alone in the last single-pair call."""
        ms = ctypes.c_double()
        _check("lcs_ctx_chain_ms", lib().lcs_ctx_chain_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def status(self) -> int:
        """lcs_ctx_status: waits for the context's kernels; LCS_OK or LCS_ETIMEDOUT."""
        return lib().lcs_ctx_status(self._h)

    def set_timeout(self, ticks: int) -> None:
        _check("lcs_ctx_set_timeout", lib().lcs_ctx_set_timeout(self._h, int(ticks)))

    def memory(self):
        """(held, budget) device bytes."""
        held, budget = ctypes.c_size_t(), ctypes.c_size_t()
        _check("lcs_ctx_memory", lib().lcs_ctx_memory(self._h, ctypes.byref(held), ctypes.byref(budget)))
        return held.value, budget.value

    def comb(self, a, b, sched: Schedule | None = None):
        """(h_out, v_out, llcs) of the seaweed comb of a against b."""
        a, b = _i32(a), _i32(b)
        h = np.empty(max(a.size, 1), np.int32)
        v = np.empty(max(b.size, 1), np.int32)
        llcs = ctypes.c_int32(-1)
        rc = lib().lcs_comb(self._h, _p(a), a.size, _p(b), b.size,
                            ctypes.byref(sched) if sched is not None else None,
                            _p(h), _p(v), ctypes.byref(llcs))
        _check("lcs_comb", rc)
        return h[: a.size].copy(), v[: b.size].copy(), llcs.value

    def comb_permutation(self, a, b, sched: Schedule | None = None):
        """(row_to_col, col_to_row, llcs) of a against b, the permutation built
        on the GPU (lcs_comb_permutation)."""
        a, b = _i32(a), _i32(b)
        size = a.size + b.size
        r2c = np.empty(max(size, 1), np.int32)
        c2r = np.empty(max(size, 1), np.int32)
        llcs = ctypes.c_int32(-1)
        rc = lib().lcs_comb_permutation(self._h, _p(a), a.size, _p(b), b.size,
                                        ctypes.byref(sched) if sched is not None else None,
                                        _p(r2c), _p(c2r), ctypes.byref(llcs))
        _check("lcs_comb_permutation", rc)
        return r2c[:size], c2r[:size], llcs.value

    def comb_fixup(self, a_prepared, b, h, v, i0, isize, j0, jsize):
        ap, b = _i32(a_prepared), _i32(b)
        assert h.dtype == np.int32 and v.dtype == np.int32 and h.flags.c_contiguous and v.flags.c_contiguous
        _check("lcs_comb_fixup", lib().lcs_comb_fixup(self._h, _p(ap), ap.size, _p(b), b.size, _p(h), _p(v),
                                                      i0, isize, j0, jsize))

    def braid_multiply(self, p_r2c, q_r2c, k: int, row_glue: bool) -> np.ndarray:
        p, q = _i32(p_r2c), _i32(q_r2c)
        out = np.empty(max(p.size + q.size - k, 1), np.int32)
        _check("lcs_braid_multiply", lib().lcs_braid_multiply(self._h, _p(p), p.size, _p(q), q.size, k,
                                                              int(bool(row_glue)), _p(out)))
        return out[: p.size + q.size - k].copy()

    def autotune(self, m: int, n: int, alphabet: int, rate: float, candidates, mode: int = 1, reps: int = 2,
                 seed: int = 1) -> np.ndarray:
        """Measured cost (ms per pair) of each candidate schedule for the shape."""
        arr = (Schedule * len(candidates))(*candidates)
        out = np.empty(len(candidates), np.float32)
        _check("lcs_autotune_measure", lib().lcs_autotune_measure(
            self._h, m, n, alphabet, rate, seed, arr, len(candidates), mode, reps, _f(out)))
        return out

    # ---- schedule selection (table, model, autotune on a miss, default)
    def add_tuned(self, mode: int, m: int, n: int, best: Schedule) -> None:
        _check("lcs_ctx_add_tuned", lib().lcs_ctx_add_tuned(self._h, mode, m, n, ctypes.byref(best)))

    def clear_tuned(self, mode: int) -> None:
        _check("lcs_ctx_clear_tuned", lib().lcs_ctx_clear_tuned(self._h, mode))

    def set_cost_model(self, mode: int, mlp, candidates) -> None:
        """Attach (mlp, one output per candidate) for `mode`; mlp None detaches."""
        arr = (Schedule * max(len(candidates), 1))(*candidates) if candidates else None
        _check("lcs_ctx_set_cost_model", lib().lcs_ctx_set_cost_model(
            self._h, mode, mlp._h if mlp is not None else None, arr, len(candidates) if mlp is not None else 0))

    def set_autotune_on_miss(self, mode: int, candidates, reps: int = 1) -> None:
        arr = (Schedule * max(len(candidates), 1))(*candidates) if candidates else None
        _check("lcs_ctx_set_autotune_on_miss", lib().lcs_ctx_set_autotune_on_miss(
            self._h, mode, arr, len(candidates), reps))

    def choose_schedule(self, mode: int, m: int, n: int, alphabet: int = 4, rate: float = 0.1):
        """(Schedule, source) the context would use for a pair of this shape."""
        out, src = Schedule(), ctypes.c_int32()
        _check("lcs_ctx_choose_schedule", lib().lcs_ctx_choose_schedule(
            self._h, mode, m, n, alphabet, rate, ctypes.byref(out), ctypes.byref(src)))
        return out, SCHED_FROM[src.value]

    def string_substring(self, row_to_col, m: int, n: int, qi, qj) -> np.ndarray:
        """LCS(a, b[qi:qj)) per query, read from P(a, b) on the GPU."""
        p, qi, qj = _i32(row_to_col), _i32(qi), _i32(qj)
        out = np.empty(max(qi.size, 1), np.int32)
        _check("lcs_semilocal_string_substring", lib().lcs_semilocal_string_substring(
            self._h, _p(p), m, n, _p(qi), _p(qj), qi.size, _p(out)))
        return out[: qi.size].copy()

    def dominance(self, row_to_col, qr, qc) -> np.ndarray:
        p, qr, qc = _i32(row_to_col), _i32(qr), _i32(qc)
        out = np.empty(max(qr.size, 1), np.int32)
        _check("lcs_permutation_dominance", lib().lcs_permutation_dominance(
            self._h, _p(p), p.size, _p(qr), _p(qc), qr.size, _p(out)))
        return out[: qr.size].copy()

    def llcs(self, a, b, sigma: int) -> int:
        """LCS length by the bit-parallel kernel (symbols in [0, sigma))."""
        a, b = _i32(a), _i32(b)
        out = ctypes.c_int32()
        _check("lcs_llcs", lib().lcs_llcs(self._h, _p(a), a.size, _p(b), b.size, sigma, ctypes.byref(out)))
        return out.value

    def llcs_device(self, d_a: int, m: int, d_b: int, n: int, sigma: int, d_out: int) -> None:
        _check("lcs_llcs_device", lib().lcs_llcs_device(self._h, ctypes.c_void_p(d_a), m, ctypes.c_void_p(d_b), n,
                                                        sigma, ctypes.c_void_p(d_out)))

    def llcs_batch(self, a_list, b_list, sigma: int) -> np.ndarray:
        """LCS length of every pair in one launch."""
        npairs = len(a_list)
        m = np.array([len(x) for x in a_list], np.int32)
        n = np.array([len(x) for x in b_list], np.int32)
        a_off = np.zeros(npairs, np.int64)
        b_off = np.zeros(npairs, np.int64)
        if npairs:
            a_off[1:] = np.cumsum(m[:-1], dtype=np.int64)
            b_off[1:] = np.cumsum(n[:-1], dtype=np.int64)
        a_all = _i32(np.concatenate([np.asarray(x, np.int32) for x in a_list]) if npairs else [])
        b_all = _i32(np.concatenate([np.asarray(x, np.int32) for x in b_list]) if npairs else [])
        a_all = a_all if a_all.size else np.zeros(1, np.int32)
        b_all = b_all if b_all.size else np.zeros(1, np.int32)
        out = np.empty(max(npairs, 1), np.int32)
        _check("lcs_llcs_batch", lib().lcs_llcs_batch(
            self._h, npairs, _p(a_all), a_off.ctypes.data_as(_i64p), _p(m), _p(b_all),
            b_off.ctypes.data_as(_i64p), _p(n), sigma, _p(out)))
        return out[:npairs].copy()

    def semilocal(self, row_to_col, m: int, n: int, kind: int, q0, q1) -> np.ndarray:
        """Quadrant queries (SEMI_*) read from P(a, b) on the GPU."""
        p, q0, q1 = _i32(row_to_col), _i32(q0), _i32(q1)
        out = np.empty(max(q0.size, 1), np.int32)
        _check("lcs_semilocal_query", lib().lcs_semilocal_query(
            self._h, _p(p), m, n, kind, _p(q0), _p(q1), q0.size, _p(out)))
        return out[: q0.size].copy()

    def batch_comb(self, a_list, b_list, scheds=None, use_graph: bool = True):
        """Comb every pair (a_list[p], b_list[p]) in one batched launch."""
        npairs = len(a_list)
        m = np.array([len(x) for x in a_list], np.int32)
        n = np.array([len(x) for x in b_list], np.int32)
        a_off = np.zeros(npairs, np.int64)
        b_off = np.zeros(npairs, np.int64)
        if npairs:
            a_off[1:] = np.cumsum(m[:-1], dtype=np.int64)
            b_off[1:] = np.cumsum(n[:-1], dtype=np.int64)
        a_all = _i32(np.concatenate([np.asarray(x, np.int32) for x in a_list]) if npairs else [])
        b_all = _i32(np.concatenate([np.asarray(x, np.int32) for x in b_list]) if npairs else [])
        h = np.empty(max(a_all.size, 1), np.int32)
        v = np.empty(max(b_all.size, 1), np.int32)
        sarr = None
        if scheds is not None:
            sarr = (Schedule * npairs)(*scheds)
        rc = lib().lcs_batch_comb(self._h, npairs, _p(a_all), a_off.ctypes.data_as(_i64p), _p(m),
                                  _p(b_all), b_off.ctypes.data_as(_i64p), _p(n), sarr, int(use_graph),
                                  _p(h), _p(v))
        _check("lcs_batch_comb", rc)
        return ([h[a_off[p]: a_off[p] + m[p]].copy() for p in range(npairs)],
                [v[b_off[p]: b_off[p] + n[p]].copy() for p in range(npairs)])


class DominanceIndex:
    """A device-resident dominance index of one permutation
    (lcs_dominance_index_*): quadrant queries in O(log size) each."""

    def __init__(self, ctx: Context, row_to_col=None, size: int | None = None, device_ptr: int | None = None):
        self._ctx, self._h = ctx, ctypes.c_void_p()
        if device_ptr is not None:
            _check("lcs_dominance_index_build_device", lib().lcs_dominance_index_build_device(
                ctx._h, ctypes.c_void_p(device_ptr), int(size), ctypes.byref(self._h)))
            self.size = int(size)
        else:
            p = _i32(row_to_col)
            _check("lcs_dominance_index_build", lib().lcs_dominance_index_build(
                ctx._h, _p(p), p.size, ctypes.byref(self._h)))
            self.size = p.size

    def query(self, m: int, n: int, kind: int, q0, q1) -> np.ndarray:
        q0, q1 = _i32(q0), _i32(q1)
        out = np.empty(max(q0.size, 1), np.int32)
        _check("lcs_dominance_index_query", lib().lcs_dominance_index_query(
            self._h, m, n, kind, _p(q0), _p(q1), q0.size, _p(out)))
        return out[: q0.size].copy()

    def query_device(self, m: int, n: int, kind: int, d_q0: int, d_q1: int, nq: int, d_out: int) -> None:
        _check("lcs_dominance_index_query_device", lib().lcs_dominance_index_query_device(
            self._h, m, n, kind, ctypes.c_void_p(d_q0), ctypes.c_void_p(d_q1), nq, ctypes.c_void_p(d_out)))

    @property
    def nbytes(self) -> int:
        return int(lib().lcs_dominance_index_bytes(self._h))

    def close(self) -> None:
        if self._h:
            lib().lcs_dominance_index_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchLayout:
    """Concatenated pairs: a_all / b_all with per-pair offsets and lengths
    (the layout lcs_batch_comb and lcs_batch_comb_device take)."""

    def __init__(self, a_list, b_list):
        self.npairs = len(a_list)
        self.m = np.array([len(x) for x in a_list], np.int32)
        self.n = np.array([len(x) for x in b_list], np.int32)
        self.a_off = np.zeros(self.npairs, np.int64)
        self.b_off = np.zeros(self.npairs, np.int64)
        if self.npairs:
            self.a_off[1:] = np.cumsum(self.m[:-1], dtype=np.int64)
            self.b_off[1:] = np.cumsum(self.n[:-1], dtype=np.int64)
        self.a_all = _i32(np.concatenate([np.asarray(x, np.int32) for x in a_list]) if self.npairs else [])
        self.b_all = _i32(np.concatenate([np.asarray(x, np.int32) for x in b_list]) if self.npairs else [])

    def split(self, h_all, v_all):
        return ([h_all[self.a_off[p]: self.a_off[p] + self.m[p]] for p in range(self.npairs)],
                [v_all[self.b_off[p]: self.b_off[p] + self.n[p]] for p in range(self.npairs)])


def batch_comb_device(ctx: "Context", layout: BatchLayout, d_a, d_b, d_h, d_v, scheds=None,
                      use_graph: bool = True) -> None:
    """lcs_batch_comb_device: d_a, d_b, d_h, d_v are device addresses of the
    concatenated symbols and strands (layout's offsets)."""
    sarr = (Schedule * layout.npairs)(*scheds) if scheds is not None else None
    rc = lib().lcs_batch_comb_device(ctx.handle, layout.npairs, d_a, layout.a_off.ctypes.data_as(_i64p),
                                     _p(layout.m), d_b, layout.b_off.ctypes.data_as(_i64p), _p(layout.n),
                                     sarr, int(use_graph), d_h, d_v)
    _check("lcs_batch_comb_device", rc)


def _f(x: np.ndarray):
    return x.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


class Mlp:
    """The schedule cost model (lcs_mlp_*): fp32 MLP trained on the GPU."""

    def __init__(self, ctx: Context, n_in: int = 4, hidden: int = 64, n_out: int = 6, seed: int = 1):
        self.ctx = ctx
        self.n_in, self.hidden, self.n_out = n_in, hidden, n_out
        self._h = ctypes.c_void_p()
        _check("lcs_mlp_create", lib().lcs_mlp_create(ctx.handle, n_in, hidden, n_out, seed, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().lcs_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def param_count(self) -> int:
        c = ctypes.c_int32()
        _check("lcs_mlp_param_count", lib().lcs_mlp_param_count(self._h, ctypes.byref(c)))
        return c.value

    def get_params(self) -> np.ndarray:
        out = np.empty(self.param_count(), np.float32)
        _check("lcs_mlp_get_params", lib().lcs_mlp_get_params(self._h, _f(out)))
        return out

    def set_params(self, p) -> None:
        p = _f32(p)
        assert p.size == self.param_count()
        _check("lcs_mlp_set_params", lib().lcs_mlp_set_params(self._h, _f(p)))

    def predict(self, X) -> np.ndarray:
        X = _f32(X).reshape(-1, self.n_in)
        Y = np.empty((max(X.shape[0], 1), self.n_out), np.float32)
        _check("lcs_mlp_predict", lib().lcs_mlp_predict(self._h, _f(X), X.shape[0], _f(Y)))
        return Y[: X.shape[0]]

    def loss_grad(self, X, T, M):
        X, T, M = _f32(X), _f32(T), _f32(M)
        loss = np.zeros(1, np.float32)
        g = np.empty(self.param_count(), np.float32)
        _check("lcs_mlp_loss_grad", lib().lcs_mlp_loss_grad(self._h, _f(X), _f(T), _f(M), X.shape[0], _f(loss), _f(g)))
        return float(loss[0]), g

    def train(self, X, T, M, steps: int, lr: float = 3e-3) -> float:
        X, T, M = _f32(X), _f32(T), _f32(M)
        loss = np.zeros(1, np.float32)
        _check("lcs_mlp_train", lib().lcs_mlp_train(self._h, _f(X), _f(T), _f(M), X.shape[0], steps, lr, _f(loss)))
        return float(loss[0])

    def choose(self, X) -> np.ndarray:
        X = _f32(X).reshape(-1, self.n_in)
        c = np.empty(max(X.shape[0], 1), np.int32)
        _check("lcs_mlp_choose", lib().lcs_mlp_choose(self._h, _f(X), X.shape[0], _p(c)))
        return c[: X.shape[0]]


class Server:
    """Continuous-batching request server (lcs_server_*)."""

    def __init__(self, ctx: Context, hbm_budget: int = 0, max_inflight: int = 4):
        self._h = ctypes.c_void_p()
        self.sizes = {}
        _check("lcs_server_create", lib().lcs_server_create(ctx.handle, hbm_budget, max_inflight,
                                                            ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().lcs_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_cost_model(self, mlp: "Mlp | None", candidates):
        arr = (Schedule * len(candidates))(*candidates) if candidates else None
        _check("lcs_server_set_cost_model",
               lib().lcs_server_set_cost_model(self._h, mlp._h if mlp else None, arr, len(candidates)))

    def submit(self, a, b, sched: Schedule | None = None, features=None) -> int:
        a, b = _i32(a), _i32(b)
        rid = ctypes.c_int64()
        f = _f32(features) if features is not None else None
        _check("lcs_server_submit", lib().lcs_server_submit(
            self._h, _p(a), a.size, _p(b), b.size, ctypes.byref(sched) if sched is not None else None,
            _f(f) if f is not None else None, ctypes.byref(rid)))
        self.sizes[rid.value] = (a.size, b.size)
        return rid.value

    def poll(self, rid: int):
        m, n = self.sizes[rid]
        h = np.empty(max(m, 1), np.int32)
        v = np.empty(max(n, 1), np.int32)
        done = ctypes.c_int32()
        _check("lcs_server_poll", lib().lcs_server_poll(self._h, rid, _p(h), _p(v), ctypes.byref(done)))
        return (h[:m].copy(), v[:n].copy()) if done.value else None

    def wait(self, rid: int, timeout_ms: int = -1):
        m, n = self.sizes[rid]
        h = np.empty(max(m, 1), np.int32)
        v = np.empty(max(n, 1), np.int32)
        _check("lcs_server_wait", lib().lcs_server_wait(self._h, rid, _p(h), _p(v), timeout_ms))
        return h[:m].copy(), v[:n].copy()

    def times(self, rid: int):
        q, t = ctypes.c_double(), ctypes.c_double()
        _check("lcs_server_request_times", lib().lcs_server_request_times(self._h, rid, ctypes.byref(q),
                                                                          ctypes.byref(t)))
        return q.value, t.value

    def release(self, rid: int):
        _check("lcs_server_release", lib().lcs_server_release(self._h, rid))
        self.sizes.pop(rid, None)

    def cancel(self, rid: int):
        _check("lcs_server_cancel", lib().lcs_server_cancel(self._h, rid))

    def stats(self):
        p, c, bt = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        mi = ctypes.c_int32()
        fb = ctypes.c_size_t()
        rc = lib().lcs_server_stats(self._h, ctypes.byref(p), ctypes.byref(c), ctypes.byref(bt), ctypes.byref(mi),
                                    ctypes.byref(fb))
        return {"pending": p.value, "completed": c.value, "batches": bt.value, "max_inflight": mi.value,
                "free_bytes": fb.value, "status": rc}


def from_strands(h, v):
    h, v = _i32(h), _i32(v)
    size = h.size + v.size
    r2c = np.empty(max(size, 1), np.int32)
    c2r = np.empty(max(size, 1), np.int32)
    _check("lcs_permutation_from_strands",
           lib().lcs_permutation_from_strands(_p(h), h.size, _p(v), v.size, _p(r2c), _p(c2r)))
    return r2c[:size].copy(), c2r[:size].copy()


def perm_hash(col_to_row) -> int:
    c2r = _i32(col_to_row)
    return int(lib().lcs_permutation_hash(_p(c2r), c2r.size))


# ---------------------------------------------------------------- workloads
def gen_binary(seed: int, length: int) -> np.ndarray:
    out = np.empty(max(length, 1), np.int32)
    lib().lcs_gen_binary_mt19937(seed, length, _p(out))
    return out[:length].copy()


def gen_uniform(seed: int, length: int, alphabet: int) -> np.ndarray:
    out = np.empty(max(length, 1), np.int32)
    lib().lcs_gen_uniform(seed, length, alphabet, _p(out))
    return out[:length].copy()


def gen_zipf(seed: int, length: int, alphabet: int, s: float) -> np.ndarray:
    out = np.empty(max(length, 1), np.int32)
    lib().lcs_gen_zipf(seed, length, alphabet, s, _p(out))
    return out[:length].copy()


def gen_mutate(seed: int, src, alphabet: int, sub_rate: float, indel_rate: float, out_len: int) -> np.ndarray:
    src = _i32(src)
    out = np.empty(max(out_len, 1), np.int32)
    lib().lcs_gen_mutate(seed, _p(src), src.size, alphabet, sub_rate, indel_rate, out_len, _p(out))
    return out[:out_len].copy()


def gen_planted(seed: int, a, n: int, alphabet: int, s: float, copies: int, sub_rate: float) -> np.ndarray:
    a = _i32(a)
    out = np.empty(max(n, 1), np.int32)
    lib().lcs_gen_planted(seed, _p(a), a.size, n, alphabet, s, copies, sub_rate, _p(out))
    return out[:n].copy()


def gen_unit(seed: int, k: int) -> float:
    return float(lib().lcs_gen_unit(seed, k))


@dataclass
class Pair:
    a: np.ndarray
    b: np.ndarray
    alphabet: int
    rate: float = 0.0


def workload_single(name: str, seed: int = 1) -> Pair:
    """The single-pair workloads: 'uniform4096', 'mutated2^20', 'zipf_planted'."""
    if name == "uniform4096":
        return Pair(gen_uniform(seed, 4096, 4), gen_uniform(seed + 1, 4096, 4), 4)
    if name.startswith("mutated"):
        spec = name[len("mutated"):]
        size = 1 << int(spec[2:]) if spec.startswith("2^") else int(spec)
        a = gen_uniform(seed, size, 4)
        b = gen_mutate(seed + 1, a, 4, 0.10, 0.02, size)
        return Pair(a, b, 4, 0.12)
    if name == "zipf_planted":
        a = gen_zipf(seed, 1 << 14, 26, 1.1)
        b = gen_planted(seed + 1, a, 1 << 22, 26, 1.1, 64, 0.05)
        return Pair(a, b, 26, 0.05)
    raise ValueError(name)


def workload_batch(npairs: int, seed: int = 7, lo: int = 256, hi: int = 16384, alphabet=4,
                   rate_lo: float = 0.05, rate_hi: float = 0.20, indel_frac: float = 0.2):
    """Pairs with m, n drawn independently log-uniform in [lo, hi]; B derived
    from A (cyclically extended or cropped to n) with a per-pair mutation
    rate uniform in [rate_lo, rate_hi] of which indel_frac are indels.
    `alphabet` is an int or a tuple to draw from uniformly."""
    pairs = []
    lg_lo, lg_hi = np.log(lo), np.log(hi)
    for p in range(npairs):
        u = [gen_unit(seed, 4 * p + q) for q in range(4)]
        m = int(np.floor(np.exp(lg_lo + (lg_hi - lg_lo) * u[0])))
        n = int(np.floor(np.exp(lg_lo + (lg_hi - lg_lo) * u[1])))
        m, n = min(max(m, lo), hi), min(max(n, lo), hi)
        rate = rate_lo + (rate_hi - rate_lo) * u[2]
        alpha = alphabet if isinstance(alphabet, int) else alphabet[int(u[3] * len(alphabet)) % len(alphabet)]
        a = gen_uniform(seed * 1000003 + 2 * p, m, alpha)
        b = gen_mutate(seed * 1000003 + 2 * p + 1, a, alpha, rate * (1 - indel_frac), rate * indel_frac, n)
        pairs.append(Pair(a, b, alpha, rate))
    return pairs
