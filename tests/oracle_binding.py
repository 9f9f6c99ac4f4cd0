"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the test suite.

The oracle is the checker: a C restatement of the reference's delta and
score arithmetic.  Only tests, the smoke check and the benchmark's
CPU-baseline leg load it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_L = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)


def lib():
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(ORACLE_LIB):
        build()
    L = C.CDLL(ORACLE_LIB)
    vp = C.c_void_p
    sig = {
        "orc_build": (vp, [C.c_int, _i32p, _i32p, _i32p, C.c_int64, _i32p, _i32p, _i32p, C.c_int64]),
        "orc_free": (None, [vp]),
        "orc_score_opt": (C.c_int64, [vp, _i32p]),
        "orc_score_ref": (C.c_int64, [vp, _i32p]),
        "orc_delta_b": (C.c_int32, [vp, _i32p, C.c_int, C.c_int]),
        "orc_delta_g2": (C.c_int32, [vp, _i32p, C.c_int, C.c_int]),
        "orc_delta_g3": (C.c_int32, [vp, _i32p, C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_best_g3": (C.c_int32, [vp, _i32p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
        "orc_apply_g3": (None, [_i32p, C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_delta_p": (C.c_int32, [vp, _i32p, _i32p, _i32p, C.c_int]),
        "orc_next_permutation": (C.c_int, [_i32p, C.c_int]),
        "orc_unrank_perm": (None, [C.c_int, C.c_int64, _i32p]),
        "orc_best_perm": (C.c_int32, [vp, _i32p, _i32p, C.c_int, C.POINTER(C.c_int64)]),
        "orc_apply_perm": (None, [_i32p, _i32p, _i32p, C.c_int]),
        "orc_delta_a": (C.c_int32, [vp, _i32p, C.c_int, C.c_int]),
        "orc_delta_swap_exact": (C.c_int64, [vp, _i32p, C.c_int, C.c_int]),
        "orc_first_improvement": (C.c_int, [vp, _i32p, C.c_int, C.c_int, _i32p, C.c_int64, C.POINTER(C.c_int64),
                                            _i64p, C.c_int64, C.POINTER(C.c_int64), C.c_int,
                                            C.POINTER(C.c_int64)]),
        "orc_row_best_pass": (C.c_int64, [vp, _i32p, _i32p, C.c_int64]),
        "orc_all_pairs": (C.c_int32, [vp, _i32p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), vp, vp, vp]),
        "orc_omp_rowbest_rows": (C.c_int64, [vp, _i32p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
        "orc_row_best_rows": (None, [vp, _i32p, _i32p, C.c_int64, _i32p, _i32p, C.c_int]),
        "orc_delta_rows": (None, [vp, _i32p, _i32p, C.c_int64, _i32p, C.c_int]),
        "orc_first_improvement_rows": (C.c_int64, [vp, _i32p, C.c_int, _i32p, C.c_int64, C.c_int]),
        "orc_best_g3_many": (None, [vp, _i32p, _i32p, C.c_int64, C.c_int, _i32p, _i32p, C.c_int]),
        "orc_best_perm_many": (None, [vp, _i32p, _i32p, C.c_int, C.c_int64, _i32p, _i64p, C.c_int]),
        "orc_rng_next": (C.c_uint64, [C.POINTER(C.c_uint64)]),
        "orc_q3_triples": (C.c_int64, [C.c_int, _i32p, C.c_int, vp]),
        "orc_q3": (C.c_int, [vp, _i32p, C.c_uint64, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double, _i64p]),
        "orc_lex_triples": (C.c_int64, [C.c_int, C.c_int64, C.c_int64, vp]),
        "orc_triples_pass": (C.c_int, [vp, _i32p, C.c_uint64, C.c_int64, C.c_int, _i64p]),
        "orc_vertex_scores": (None, [vp, np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")]),
        "orc_perturb": (C.c_int, [vp, _i32p, C.c_int, C.c_int, C.POINTER(C.c_uint64), _i32p]),
        "orc_row_best_restart": (C.c_int, [vp, _i32p, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int64, _i64p]),
        "orc_quick": (C.c_int, [vp, _i32p, C.c_uint64, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double,
                                C.c_int, _i64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _L = L
    return L


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def rng_next(state):
    """orc_rng_next (splitmix64): returns (value, advanced state)."""
    st = C.c_uint64(state)
    v = lib().orc_rng_next(C.byref(st))
    return v, st.value


class Oracle:
    def __init__(self, n, g1, g2):
        self.n = int(n)
        s1, d1, w1 = (_i32(a) for a in g1)
        s2, d2, w2 = (_i32(a) for a in g2)
        self._h = lib().orc_build(self.n, s1, d1, w1, len(s1), s2, d2, w2, len(s2))
        if not self._h:
            raise ValueError("orc_build rejected the graphs (vertex id outside [0, n])")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    def score_opt(self, m):
        return lib().orc_score_opt(self._h, _i32(m))

    def score_ref(self, m):
        return lib().orc_score_ref(self._h, _i32(m))

    def delta_b(self, m, a, b):
        return lib().orc_delta_b(self._h, _i32(m), a, b)

    def delta_g2(self, m, a, b):
        return lib().orc_delta_g2(self._h, _i32(m), a, b)

    def delta_g3(self, m, m1, m2, m3, p):
        return lib().orc_delta_g3(self._h, _i32(m), m1, m2, m3, p)

    def best_g3(self, m, m1, m2, m3, mode):
        bp = C.c_int()
        d = lib().orc_best_g3(self._h, _i32(m), m1, m2, m3, mode, C.byref(bp))
        return d, bp.value

    def delta_p(self, m, sset, perm):
        return lib().orc_delta_p(self._h, _i32(m), _i32(sset), _i32(perm), len(sset))

    def best_perm(self, m, sset):
        r = C.c_int64()
        d = lib().orc_best_perm(self._h, _i32(m), _i32(sset), len(sset), C.byref(r))
        return d, r.value

    def delta_a(self, m, a, b):
        return lib().orc_delta_a(self._h, _i32(m), a, b)

    def vertex_scores(self):
        out = np.zeros(self.n + 1, np.float32)
        lib().orc_vertex_scores(self._h, out)
        return out

    def perturb(self, m, k, mode, rng_state):
        """Returns (perturbed copy of m, the 2k draws, the advanced generator state)."""
        mm = _i32(m).copy()
        st = C.c_uint64(rng_state)
        sel = np.zeros(2 * k, np.int32)
        if lib().orc_perturb(self._h, mm, k, mode, C.byref(st), sel):
            raise ValueError("orc_perturb: bad k")
        return mm, sel, st.value

    def row_best_restart(self, m, seed=1, initial_swaps=100, restart_swaps=50, idle_passes=2, max_passes=4):
        mm = _i32(m).copy()
        out = np.zeros(6, np.int64)
        if lib().orc_row_best_restart(self._h, mm, seed, initial_swaps, restart_swaps, idle_passes, max_passes, out):
            raise ValueError("orc_row_best_restart failed")
        keys = ("passes", "restarts", "accepted", "initial_score", "final_score", "best_score")
        return mm, dict(zip(keys, (int(x) for x in out)))

    def quick(self, m, seed=1, max_evals=1000, zero_limit=1000, best_window=500, gap=0.002, gap_norm=5154247.0,
              kind=1):
        mm = _i32(m).copy()
        out = np.zeros(9, np.int64)
        if lib().orc_quick(self._h, mm, seed, max_evals, zero_limit, best_window, gap, gap_norm, kind, out):
            raise ValueError("orc_quick failed")
        keys = ("evals", "perturbations", "moves", "positive_moves", "zero_moves", "initial_score", "final_score",
                "best_score", "pairs")
        return mm, dict(zip(keys, (int(x) for x in out)))

    def q3(self, m, seed=1, max_evals=100, zero_limit=1000, best_window=1500, gap=0.002, gap_norm=5154247.0):
        mm = _i32(m).copy()
        out = np.zeros(9, np.int64)
        if lib().orc_q3(self._h, mm, seed, max_evals, zero_limit, best_window, gap, gap_norm, out):
            raise ValueError("orc_q3 failed")
        keys = ("evals", "perturbations", "moves", "positive_moves", "zero_moves", "initial_score", "final_score",
                "best_score", "triples")
        return mm, dict(zip(keys, (int(x) for x in out)))

    def triples_pass(self, m, seed=1, max_evals=100, batch=270000):
        mm = _i32(m).copy()
        out = np.zeros(8, np.int64)
        if lib().orc_triples_pass(self._h, mm, seed, max_evals, batch, out):
            raise ValueError("orc_triples_pass failed")
        keys = ("evals", "passes", "perturbations", "moves", "initial_score", "final_score", "best_score", "triples")
        return mm, dict(zip(keys, (int(x) for x in out)))

    def delta_exact(self, m, a, b):
        return lib().orc_delta_swap_exact(self._h, _i32(m), a, b)

    def first_improvement(self, m, rows_per_sweep=0, max_sweeps=0, log_every=20, acc_cap=1 << 20,
                          log_cap=1 << 20):
        m = _i32(m).copy()
        acc = np.zeros(3 * acc_cap, np.int32)
        logv = np.zeros(3 * log_cap, np.int64)
        na, nl, fs = C.c_int64(), C.c_int64(), C.c_int64()
        sweeps = lib().orc_first_improvement(self._h, m, rows_per_sweep, max_sweeps, acc, acc_cap, C.byref(na),
                                             logv, log_cap, C.byref(nl), log_every, C.byref(fs))
        return {"mapping": m, "accepted": acc[: 3 * min(na.value, acc_cap)].reshape(-1, 3),
                "log": logv[: 3 * min(nl.value, log_cap)].reshape(-1, 3), "sweeps": sweeps,
                "final_score": fs.value}

    def row_best_pass(self, m, acc_cap=1 << 20):
        m = _i32(m).copy()
        acc = np.zeros(3 * acc_cap, np.int32)
        k = lib().orc_row_best_pass(self._h, m, acc, acc_cap)
        return m, acc[: 3 * min(k, acc_cap)].reshape(-1, 3)

    def all_pairs(self, m, matrix=False):
        n1 = self.n + 1
        ba, bb = C.c_int32(), C.c_int32()
        rb = np.zeros(n1, np.int32)
        rd = np.zeros(n1, np.int32)
        dm = np.zeros((n1, n1), np.int32) if matrix else None
        d = lib().orc_all_pairs(self._h, _i32(m), C.byref(ba), C.byref(bb), rb.ctypes.data_as(C.c_void_p),
                                rd.ctypes.data_as(C.c_void_p),
                                dm.ctypes.data_as(C.c_void_p) if dm is not None else None)
        return {"best": (ba.value, bb.value, d), "row_best_b": rb, "row_best_delta": rd, "matrix": dm}

    def row_best_rows(self, m, rows, threads=0):
        rows = _i32(rows)
        b = np.zeros(len(rows), np.int32)
        d = np.zeros(len(rows), np.int32)
        lib().orc_row_best_rows(self._h, _i32(m), rows, len(rows), b, d, threads)
        return b, d

    def delta_rows(self, m, rows, threads=0):
        rows = _i32(rows)
        out = np.zeros(len(rows) * (self.n + 1), np.int32)
        lib().orc_delta_rows(self._h, _i32(m), rows, len(rows), out, threads)
        return out.reshape(len(rows), self.n + 1)

    def first_improvement_rows(self, m, max_rows, threads=0, acc_cap=1 << 20):
        m = _i32(m).copy()
        acc = np.zeros(3 * acc_cap, np.int32)
        k = lib().orc_first_improvement_rows(self._h, m, max_rows, acc, acc_cap, threads)
        return m, acc[: 3 * min(k, acc_cap)].reshape(-1, 3)

    def best_g3_many(self, m, triples, mode, threads=0):
        t = _i32(triples).reshape(-1)
        c = len(t) // 3
        d = np.zeros(c, np.int32)
        p = np.zeros(c, np.int32)
        lib().orc_best_g3_many(self._h, _i32(m), t, c, mode, d, p, threads)
        return d, p

    def best_perm_many(self, m, sets, threads=0):
        s = _i32(sets)
        c, k = s.shape
        d = np.zeros(c, np.int32)
        r = np.zeros(c, np.int64)
        lib().orc_best_perm_many(self._h, _i32(m), np.ascontiguousarray(s.reshape(-1)), k, c, d, r, threads)
        return d, r

    def omp_rowbest_rows(self, m, lo, hi, threads=0):
        cs = C.c_int64()
        ev = lib().orc_omp_rowbest_rows(self._h, _i32(m), lo, hi, threads, C.byref(cs))
        return ev, cs.value


def q3_triples(n, changed):
    c = _i32(changed)
    cnt = lib().orc_q3_triples(n, c, len(c), None)
    out = np.zeros(3 * cnt, np.int32)
    lib().orc_q3_triples(n, c, len(c), out.ctypes.data_as(C.c_void_p))
    return out.reshape(cnt, 3)


def lex_triples(n, start, count):
    cnt = lib().orc_lex_triples(n, start, count, None)
    out = np.zeros(3 * max(cnt, 1), np.int32)
    lib().orc_lex_triples(n, start, count, out.ctypes.data_as(C.c_void_p))
    return out[:3 * cnt].reshape(cnt, 3)


def unrank_perm(k, rank):
    out = np.zeros(k, np.int32)
    lib().orc_unrank_perm(k, rank, out)
    return out


def next_permutation(arr):
    a = _i32(arr).copy()
    ok = lib().orc_next_permutation(a, len(a))
    return bool(ok), a


def apply_g3(m, m1, m2, m3, p):
    m = _i32(m).copy()
    lib().orc_apply_g3(m, m1, m2, m3, p)
    return m


def apply_perm(m, sset, perm):
    m = _i32(m).copy()
    lib().orc_apply_perm(m, _i32(sset), _i32(perm), len(sset))
    return m
