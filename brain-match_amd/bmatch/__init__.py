"""ctypes binding of the MI355X move-delta engine (include/bmatch.h, include/bm_host.h).

The engine is the C ABI in ``brain-match_amd/lib/libbmatch.so``; this module
is thin plumbing over it for the test suite, the benchmark and scripting.
Arrays cross the boundary as contiguous numpy int32 buffers.  There is no
fallback path: if the shared library is missing or a call fails, an
exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libbmatch.so")
BIN_DIR = os.path.join(_PKG, "bin")

BM_OK, BM_EINVAL, BM_ENOMEM, BM_EHIP, BM_ESTATE, BM_ENODEV, BM_ERANGE, BM_EIO = 0, -1, -2, -3, -4, -5, -6, -7
BM_ROW_FIRST_POSITIVE = 0
BM_ROW_ARGMAX = 1
BM_ROW_GROUP_MAX = 16
BM_ROT3_ROTATIONS = 0
BM_ROT3_ALL = 1
BM_PERTURB_OMP = 0
BM_PERTURB_WEIGHTED = 1
BM_DELTA_LISTS = 0
BM_DELTA_DENSE = 1

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None


class BMError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: error {code}: {msg}")
        self.code = code


def lib():
    """Load libbmatch.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} not built: run `make -C brain-match_amd`")
    L = C.CDLL(LIB_PATH)
    sig = {
        "bm_strerror": (C.c_char_p, [C.c_int]),
        "bm_last_error": (C.c_char_p, []),
        "bm_abi_version": (C.c_int, []),
        "bm_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
        "bm_destroy": (C.c_int, [C.c_void_p]),
        "bm_upload_graphs": (C.c_int, [C.c_void_p, C.c_int32, _i32p, _i32p, _i32p, C.c_int64,
                                       _i32p, _i32p, _i32p, C.c_int64]),
        "bm_num_vertices": (C.c_int32, [C.c_void_p]),
        "bm_device_bytes": (C.c_int64, [C.c_void_p]),
        "bm_set_mapping": (C.c_int, [C.c_void_p, _i32p]),
        "bm_get_mapping": (C.c_int, [C.c_void_p, _i32p]),
        "bm_score": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
        "bm_eval_swaps": (C.c_int, [C.c_void_p, _i32p, _i32p, C.c_int64, _i32p]),
        "bm_eval_swaps_noalloc": (C.c_int, [C.c_void_p, _i32p, _i32p, C.c_int64, _i32p]),
        "bm_eval_swap_row": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int,
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "bm_eval_swap_rows": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int,
                                        _i32p, _i32p]),
        "bm_sweep_all_pairs": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32), C.c_void_p, C.c_void_p, C.c_void_p]),
        "bm_sweep_all_pairs_dense": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                               C.POINTER(C.c_int32), C.c_void_p, C.c_void_p, C.c_void_p]),
        "bm_eval_rot3": (C.c_int, [C.c_void_p, _i32p, C.c_int64, C.c_int, _i32p, _i32p]),
        "bm_eval_permset": (C.c_int, [C.c_void_p, C.c_int32, _i32p, C.c_int64, _i32p, _i32p]),
        "bm_perm_from_rank": (C.c_int, [C.c_int32, C.c_int32, _i32p]),
        "bm_eval_swaps_hillclimb": (C.c_int, [C.c_void_p, _i32p, _i32p, C.c_int64, _i32p]),
        "bm_eval_rows": (C.c_int, [C.c_void_p, _i32p, C.c_int32, C.c_int32, _i32p]),
        "bm_eval_swaps_dense": (C.c_int, [C.c_void_p, _i32p, _i32p, C.c_int64, _i32p]),
        "bm_perturb": (C.c_int, [C.c_void_p, C.c_int32, C.c_int, C.POINTER(C.c_uint64), C.c_void_p]),
        "bm_rng_next": (C.c_uint64, [C.POINTER(C.c_uint64)]),
        "bm_apply_swap": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
        "bm_apply_swap_scored": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int64)]),
        "bm_apply_rot3": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
        "bm_apply_perm": (C.c_int, [C.c_void_p, C.c_int32, _i32p, _i32p]),
        "bm_sweep_config": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "bm_moves_set": (C.c_int, [C.c_void_p, C.c_int32, _i32p, C.c_int64]),
        "bm_moves_eval": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32)]),
        "bm_moves_results": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
        "bm_refine_configure": (C.c_int, [C.c_void_p, C.c_void_p]),
        "bm_refine_run": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int32, C.c_void_p]),
        "bm_refine_history": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]),
        "bm_refine_run_many": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_int64, C.c_int32, C.c_void_p]),
        "bm_refine_log": (C.c_int, [C.c_void_p, C.c_int32]),
        "bm_refine_moves": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_void_p, C.c_int64]),
        "bm_refine_config": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "bm_job_footprint": (C.c_int64, [C.c_int32, C.c_int64, C.c_int64, C.c_void_p]),
        "bm_server_create": (C.c_int, [C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
        "bm_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]),
        "bm_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
        "bm_server_stats_get": (C.c_int, [C.c_void_p, C.c_void_p]),
        "bm_server_destroy": (C.c_int, [C.c_void_p]),
        "bm_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
        "bm_kernel_time": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
        "bm_reset_timing": (C.c_int, [C.c_void_p]),
        # host layer
        "bm_gen_powerlaw": (C.c_int, [C.c_uint64, C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_void_p]),
        "bm_gen_chunglu": (C.c_int, [C.c_uint64, C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_double,
                                     C.c_void_p]),
        "bm_instance_free": (None, [C.c_void_p]),
        "bm_gen_triples_2hop": (C.c_int, [C.c_uint64, C.c_void_p, C.c_int64, _i32p]),
        "bm_gen_sets": (C.c_int, [C.c_uint64, C.c_int32, C.c_int32, C.c_int64, _i32p]),
        "bm_edges_read_csv": (C.c_int, [C.c_char_p, C.c_int, C.c_void_p]),
        "bm_edges_free": (None, [C.c_void_p]),
        "bm_pairs_read_optimized": (C.c_int, [C.c_char_p, C.c_void_p]),
        "bm_pairs_read_permset": (C.c_int, [C.c_char_p, C.c_void_p]),
        "bm_pairs_read_plain": (C.c_int, [C.c_char_p, C.c_void_p]),
        "bm_pairs_free": (None, [C.c_void_p]),
        "bm_mapping_write_mf": (C.c_int, [C.c_char_p, _i32p, C.c_int32]),
        "bm_edges_write_csv": (C.c_int, [C.c_char_p, C.c_char_p, _i32p, _i32p, _i32p, C.c_int64]),
        "bm_drive_first_improvement": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_int, C.c_void_p,
                                                 C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
        "bm_drive_row_best": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int64, C.c_void_p]),
        "bm_drive_global_best": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_int, C.c_void_p, C.c_int64,
                                           C.c_void_p]),
        "bm_drive_row_best_restart": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "bm_drive_quick": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "bm_drive_q3": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "bm_q3_triples": (C.c_int64, [C.c_int32, _i32p, C.c_int32, C.c_void_p]),
        "bm_lex_triples": (C.c_int64, [C.c_int32, C.c_int64, C.c_int64, C.c_void_p]),
        "bm_drive_triples": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Every entry point of the C ABI that this binding declares."""
    lib()
    return sorted(n for n in dir(_lib) if n.startswith("bm_"))


def _check(rc: int, where: str):
    if rc != 0:
        L = lib()
        raise BMError(rc, where, L.bm_last_error().decode(errors="replace") or L.bm_strerror(rc).decode())


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


class _Instance(C.Structure):
    _fields_ = [("n", C.c_int32), ("e1", C.c_int64), ("e2", C.c_int64),
                ("s1", C.POINTER(C.c_int32)), ("d1", C.POINTER(C.c_int32)), ("w1", C.POINTER(C.c_int32)),
                ("s2", C.POINTER(C.c_int32)), ("d2", C.POINTER(C.c_int32)), ("w2", C.POINTER(C.c_int32)),
                ("pi", C.POINTER(C.c_int32)), ("p0", C.POINTER(C.c_int32))]


class _DriveStats(C.Structure):
    _fields_ = [("sweeps", C.c_int64), ("rows", C.c_int64), ("accepted", C.c_int64), ("evals_rows", C.c_int64),
                ("initial_score", C.c_int64), ("final_score", C.c_int64), ("mismatches", C.c_int64),
                ("rows_cleared", C.c_int64)]


_ROW_HOOK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.c_int32, C.c_int64, C.c_int64)


class RefineParams(C.Structure):
    """bm_refine_params of include/bmatch.h."""
    _fields_ = [("pairs", C.c_int32), ("triples", C.c_int32), ("sets", C.c_int32), ("set_k", C.c_int32),
                ("set_gen", C.c_int32), ("rounds", C.c_int32), ("steps_per_check", C.c_int32),
                ("seed", C.c_uint64)]


class _RefineStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("moves", C.c_int64), ("pair_moves", C.c_int64),
                ("triple_moves", C.c_int64), ("set_moves", C.c_int64), ("initial_score", C.c_int64),
                ("final_score", C.c_int64), ("idle_steps", C.c_int64), ("bad_step", C.c_int64),
                ("converged", C.c_int32), ("verified", C.c_int32)]


class _RestartParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("initial_swaps", C.c_int32), ("restart_swaps", C.c_int32),
                ("idle_passes", C.c_int32), ("verify", C.c_int32), ("max_passes", C.c_int64)]


class _RestartStats(C.Structure):
    _fields_ = [(f, C.c_int64) for f in ("passes", "restarts", "accepted", "evals_rows", "initial_score",
                                         "final_score", "best_score", "mismatches")]


class _QuickParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("max_evals", C.c_int64), ("zero_limit", C.c_int32),
                ("best_window", C.c_int32), ("gap", C.c_double), ("gap_norm", C.c_double),
                ("delta_kind", C.c_int32)]


class _QuickStats(C.Structure):
    _fields_ = [(f, C.c_int64) for f in ("evals", "perturbations", "moves", "positive_moves", "zero_moves",
                                         "initial_score", "final_score", "best_score", "pairs")]


class _Q3Params(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("max_evals", C.c_int64), ("zero_limit", C.c_int32),
                ("best_window", C.c_int32), ("gap", C.c_double), ("gap_norm", C.c_double)]


class _Q3Stats(C.Structure):
    _fields_ = [(f, C.c_int64) for f in ("evals", "perturbations", "moves", "positive_moves", "zero_moves",
                                         "initial_score", "final_score", "best_score", "triples")]


class _TriplesParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("max_evals", C.c_int64), ("batch", C.c_int32)]


class _TriplesStats(C.Structure):
    _fields_ = [(f, C.c_int64) for f in ("evals", "passes", "perturbations", "moves", "initial_score",
                                         "final_score", "best_score", "triples")]


class _RefineStep(C.Structure):
    _fields_ = [("score", C.c_int64), ("gain", C.c_int64), ("moves", C.c_int32), ("ok", C.c_int32),
                ("moved", C.c_int32), ("full", C.c_int32)]


class _ServerConfig(C.Structure):
    _fields_ = [("max_active", C.c_int32), ("prep_threads", C.c_int32), ("hbm_budget", C.c_int64)]


class _JobSpec(C.Structure):
    _fields_ = [("n", C.c_int32), ("g1_src", C.c_void_p), ("g1_dst", C.c_void_p), ("g1_w", C.c_void_p),
                ("e1", C.c_int64), ("g2_src", C.c_void_p), ("g2_dst", C.c_void_p), ("g2_w", C.c_void_p),
                ("e2", C.c_int64), ("mapping", C.c_void_p), ("move_budget", C.c_int64), ("max_steps", C.c_int64),
                ("patience", C.c_int32), ("params", RefineParams)]


class _JobResult(C.Structure):
    _fields_ = [("job_id", C.c_int64), ("status", C.c_int32), ("n", C.c_int32), ("initial_score", C.c_int64),
                ("final_score", C.c_int64), ("steps", C.c_int64), ("moves", C.c_int64),
                ("device_bytes", C.c_int64), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double), ("error", C.c_char * 160)]


class _ServerStats(C.Structure):
    _fields_ = [(f, C.c_int64) for f in ("submitted", "done", "failed", "queued", "active", "peak_active",
                                         "reserved_bytes", "peak_reserved_bytes", "hbm_budget")]


def refine_params(pairs=1, triples=0, sets=0, set_k=4, set_gen=0, rounds=2, steps_per_check=4, seed=1):
    return RefineParams(int(pairs), int(triples), int(sets), int(set_k), int(set_gen), int(rounds),
                        int(steps_per_check), int(seed))


class Instance:
    """A seeded synthetic graph pair (numpy copies of a bm_instance)."""

    def __init__(self, n, g1, g2, pi, p0):
        self.n = int(n)
        self.g1 = tuple(_i32(a) for a in g1)  # (src, dst, w)
        self.g2 = tuple(_i32(a) for a in g2)
        self.pi = _i32(pi)
        self.p0 = _i32(p0)

    @classmethod
    def _from_c(cls, inst: _Instance):
        n = inst.n

        def arr(p, m):
            return np.ctypeslib.as_array(p, shape=(m,)).copy() if m > 0 else np.zeros(0, np.int32)

        return cls(n, (arr(inst.s1, inst.e1), arr(inst.d1, inst.e1), arr(inst.w1, inst.e1)),
                   (arr(inst.s2, inst.e2), arr(inst.d2, inst.e2), arr(inst.w2, inst.e2)),
                   arr(inst.pi, n + 1), arr(inst.p0, n + 1))

    @classmethod
    def powerlaw(cls, seed, n, m, del_frac=0.10, shuffle_frac=1.0):
        """Power-law out-degree pair (the plain CPU-runnable refinement workload)."""
        L = lib()
        inst = _Instance()
        _check(L.bm_gen_powerlaw(seed, n, m, del_frac, shuffle_frac, C.byref(inst)), "bm_gen_powerlaw")
        try:
            return cls._from_c(inst)
        finally:
            L.bm_instance_free(C.byref(inst))

    @classmethod
    def chunglu(cls, seed, n, m, rewire_frac=0.15, reweight_frac=0.15, shuffle_frac=0.30):
        """Chung-Lu pair with log-normal weights (the full-size sweep workload)."""
        L = lib()
        inst = _Instance()
        _check(L.bm_gen_chunglu(seed, n, m, rewire_frac, reweight_frac, shuffle_frac, C.byref(inst)),
               "bm_gen_chunglu")
        try:
            return cls._from_c(inst)
        finally:
            L.bm_instance_free(C.byref(inst))

    def _as_c(self):
        inst = _Instance()
        inst.n = self.n
        inst.e1 = len(self.g1[0])
        inst.e2 = len(self.g2[0])
        keep = []
        for name, a in (("s1", self.g1[0]), ("d1", self.g1[1]), ("w1", self.g1[2]),
                        ("s2", self.g2[0]), ("d2", self.g2[1]), ("w2", self.g2[2]),
                        ("pi", self.pi), ("p0", self.p0)):
            keep.append(a)
            setattr(inst, name, a.ctypes.data_as(C.POINTER(C.c_int32)))
        return inst, keep

    def triples_2hop(self, seed, count):
        out = np.zeros(3 * count, np.int32)
        inst, keep = self._as_c()
        _check(lib().bm_gen_triples_2hop(seed, C.byref(inst), count, out), "bm_gen_triples_2hop")
        return out.reshape(count, 3)


class _Edges(C.Structure):
    _fields_ = [("src", C.POINTER(C.c_int32)), ("dst", C.POINTER(C.c_int32)), ("w", C.POINTER(C.c_int32)),
                ("m", C.c_int64), ("cap", C.c_int64), ("max_id", C.c_int32), ("malformed", C.c_int64)]


class _Pairs(C.Structure):
    _fields_ = [("male", C.POINTER(C.c_int32)), ("female", C.POINTER(C.c_int32)), ("m", C.c_int64),
                ("max_id", C.c_int32)]


CSV_SILENT, CSV_LOG_ERROR, CSV_WARN_LINE = 0, 1, 2


def read_edges_csv(path, report=CSV_SILENT):
    """bm_edges_read_csv: (src, dst, w, max_id, malformed) in file order."""
    L = lib()
    e = _Edges()
    _check(L.bm_edges_read_csv(os.fsencode(path), report, C.byref(e)), "bm_edges_read_csv")
    try:
        arr = lambda p: np.ctypeslib.as_array(p, shape=(e.m,)).copy() if e.m else np.zeros(0, np.int32)
        return arr(e.src), arr(e.dst), arr(e.w), e.max_id, e.malformed
    finally:
        L.bm_edges_free(C.byref(e))


def read_pairs(path, fmt="optimized"):
    """Mapping-file pairs as the named reference loader parses them."""
    L = lib()
    p = _Pairs()
    fn = {"optimized": L.bm_pairs_read_optimized, "permset": L.bm_pairs_read_permset,
          "plain": L.bm_pairs_read_plain}[fmt]
    _check(fn(os.fsencode(path), C.byref(p)), fn.__name__)
    try:
        arr = lambda q: np.ctypeslib.as_array(q, shape=(p.m,)).copy() if p.m else np.zeros(0, np.int32)
        return arr(p.male), arr(p.female)
    finally:
        L.bm_pairs_free(C.byref(p))


def write_edges_csv(path, g, header="Source,Target,Weight"):
    """An edge list (src, dst, w) as the reference CLIs read it."""
    s, d, w = (_i32(x) for x in g)
    _check(lib().bm_edges_write_csv(os.fsencode(path), header.encode(), s, d, w, len(s)), "bm_edges_write_csv")


def q3_triples(n, changed):
    """bm_q3_triples: the triples q3.cu's kernel enumerates for the sorted changed vertices, (count, 3)."""
    c = _i32(changed)
    cnt = lib().bm_q3_triples(n, c, len(c), None)
    if cnt < 0:
        raise BMError(int(cnt), "bm_q3_triples", "bad arguments")
    out = np.zeros(3 * cnt, np.int32)
    lib().bm_q3_triples(n, c, len(c), out.ctypes.data_as(C.c_void_p))
    return out.reshape(cnt, 3)


def lex_triples(n, start, count):
    """bm_lex_triples: lexicographic triples of 1..n from rank start, (count', 3)."""
    cnt = lib().bm_lex_triples(n, start, count, None)
    if cnt < 0:
        raise BMError(int(cnt), "bm_lex_triples", "bad arguments")
    out = np.zeros(3 * max(cnt, 1), np.int32)
    lib().bm_lex_triples(n, start, count, out.ctypes.data_as(C.c_void_p))
    return out[:3 * cnt].reshape(cnt, 3)


def write_mapping_mf(path, mapping):
    m = _i32(mapping)
    _check(lib().bm_mapping_write_mf(os.fsencode(path), m, len(m) - 1), "bm_mapping_write_mf")


def rng_next(state):
    """bm_rng_next (splitmix64): returns (value, advanced state)."""
    st = C.c_uint64(state)
    v = lib().bm_rng_next(C.byref(st))
    return v, st.value


def gen_sets(seed, n, k, count):
    out = np.zeros(k * count, np.int32)
    _check(lib().bm_gen_sets(seed, n, k, count, out), "bm_gen_sets")
    return out.reshape(count, k)


def perm_from_rank(k, rank):
    out = np.zeros(k, np.int32)
    _check(lib().bm_perm_from_rank(k, rank, out), "bm_perm_from_rank")
    return out


class Engine:
    """One device context holding one graph pair and its mapping."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        _check(self._L.bm_create(device, C.byref(h)), "bm_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.bm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def n(self):
        return int(self._L.bm_num_vertices(self._h))

    @property
    def device_bytes(self):
        return int(self._L.bm_device_bytes(self._h))

    def upload(self, n, g1, g2):
        s1, d1, w1 = (_i32(a) for a in g1)
        s2, d2, w2 = (_i32(a) for a in g2)
        _check(self._L.bm_upload_graphs(self._h, n, s1, d1, w1, len(s1), s2, d2, w2, len(s2)), "bm_upload_graphs")

    def set_mapping(self, mapping):
        _check(self._L.bm_set_mapping(self._h, _i32(mapping)), "bm_set_mapping")

    def mapping(self):
        out = np.zeros(self.n + 1, np.int32)
        _check(self._L.bm_get_mapping(self._h, out), "bm_get_mapping")
        return out

    def score(self):
        s = C.c_int64()
        _check(self._L.bm_score(self._h, C.byref(s)), "bm_score")
        return s.value

    def eval_swaps(self, a, b):
        a, b = _i32(a), _i32(b)
        out = np.zeros(len(a), np.int32)
        _check(self._L.bm_eval_swaps(self._h, a, b, len(a), out), "bm_eval_swaps")
        return out

    def eval_swaps_noalloc(self, a, b):
        """bm_eval_swaps_noalloc: bm_eval_swaps staged in the library's fixed device array (no allocation)."""
        a, b = _i32(a), _i32(b)
        out = np.zeros(len(a), np.int32)
        _check(self._L.bm_eval_swaps_noalloc(self._h, a, b, len(a), out), "bm_eval_swaps_noalloc")
        return out

    def eval_swaps_hillclimb(self, a, b):
        a, b = _i32(a), _i32(b)
        out = np.zeros(len(a), np.int32)
        _check(self._L.bm_eval_swaps_hillclimb(self._h, a, b, len(a), out), "bm_eval_swaps_hillclimb")
        return out

    def eval_swap_row(self, a, lo, hi, mode=BM_ROW_FIRST_POSITIVE):
        b, d = C.c_int32(), C.c_int32()
        _check(self._L.bm_eval_swap_row(self._h, a, lo, hi, mode, C.byref(b), C.byref(d)), "bm_eval_swap_row")
        return b.value, d.value

    def eval_swap_rows(self, a, count, lo, lo_next, mode=BM_ROW_FIRST_POSITIVE):
        """bm_eval_swap_rows: rows a..a+count-1 under one mapping -> (b[count], delta[count])."""
        b = np.zeros(count, np.int32)
        d = np.zeros(count, np.int32)
        _check(self._L.bm_eval_swap_rows(self._h, a, count, lo, lo_next, mode, b, d), "bm_eval_swap_rows")
        return b, d

    def sweep_all_pairs(self, rows=False, matrix=False, kind=BM_DELTA_LISTS):
        """bm_sweep_all_pairs (Delta_B) or, with kind=BM_DELTA_DENSE, bm_sweep_all_pairs_dense."""
        a, b, d = C.c_int32(), C.c_int32(), C.c_int32()
        n1 = self.n + 1
        rb = np.zeros(n1, np.int32) if rows else None
        rd = np.zeros(n1, np.int32) if rows else None
        dm = np.zeros((n1, n1), np.int32) if matrix else None
        ptr = lambda x: x.ctypes.data_as(C.c_void_p) if x is not None else None
        fn = self._L.bm_sweep_all_pairs_dense if kind == BM_DELTA_DENSE else self._L.bm_sweep_all_pairs
        _check(fn(self._h, C.byref(a), C.byref(b), C.byref(d), ptr(rb), ptr(rd), ptr(dm)), fn.__name__)
        return {"best": (a.value, b.value, d.value), "row_best_b": rb, "row_best_delta": rd, "matrix": dm}

    def eval_rot3(self, triples, mode=BM_ROT3_ROTATIONS):
        t = _i32(triples).reshape(-1)
        cnt = len(t) // 3
        d = np.zeros(cnt, np.int32)
        p = np.zeros(cnt, np.int32)
        _check(self._L.bm_eval_rot3(self._h, t, cnt, mode, d, p), "bm_eval_rot3")
        return d, p

    def eval_permset(self, sets):
        s = _i32(sets)
        k = s.shape[1]
        cnt = s.shape[0]
        d = np.zeros(cnt, np.int32)
        r = np.zeros(cnt, np.int32)
        _check(self._L.bm_eval_permset(self._h, k, s.reshape(-1), cnt, d, r), "bm_eval_permset")
        return d, r

    def eval_rows(self, rows, kind=BM_DELTA_LISTS):
        """bm_eval_rows: Delta(rows[i], b) for b = 0..n as a (len(rows), n+1) array; kind BM_DELTA_LISTS
        gives Delta_B, BM_DELTA_DENSE the dense-matrix delta of quick.cu's kernels."""
        r = _i32(rows)
        out = np.zeros((len(r), self.n + 1), np.int32)
        _check(self._L.bm_eval_rows(self._h, r, len(r), kind, out.reshape(-1)), "bm_eval_rows")
        return out

    def eval_swaps_dense(self, a, b):
        """bm_eval_swaps_dense: the dense-matrix delta of quick.cu:48-250 for pairs (a[i], b[i])."""
        a, b = _i32(a), _i32(b)
        out = np.zeros(len(a), np.int32)
        _check(self._L.bm_eval_swaps_dense(self._h, a, b, len(a), out), "bm_eval_swaps_dense")
        return out

    def perturb(self, k, mode, rng_state):
        """bm_perturb: returns (the 2k drawn vertices, the advanced generator state)."""
        st = C.c_uint64(rng_state)
        sel = np.zeros(2 * k, np.int32)
        _check(self._L.bm_perturb(self._h, k, mode, C.byref(st), sel.ctypes.data_as(C.c_void_p)), "bm_perturb")
        return sel, st.value

    def row_best_restart(self, seed=1, initial_swaps=100, restart_swaps=50, idle_passes=2, verify=True,
                         max_passes=4):
        prm = _RestartParams(seed, initial_swaps, restart_swaps, idle_passes, 1 if verify else 0, max_passes)
        best = np.zeros(self.n + 1, np.int32)
        st = _RestartStats()
        _check(self._L.bm_drive_row_best_restart(self._h, C.byref(prm), best.ctypes.data_as(C.c_void_p),
                                                 C.byref(st)), "bm_drive_row_best_restart")
        return best, {f: getattr(st, f) for f, _ in _RestartStats._fields_}

    def quick(self, seed=1, max_evals=1000, zero_limit=1000, best_window=500, gap=0.002, gap_norm=5154247.0,
              kind=BM_DELTA_DENSE):
        prm = _QuickParams(seed, max_evals, zero_limit, best_window, gap, gap_norm, kind)
        best = np.zeros(self.n + 1, np.int32)
        st = _QuickStats()
        _check(self._L.bm_drive_quick(self._h, C.byref(prm), best.ctypes.data_as(C.c_void_p), C.byref(st)),
               "bm_drive_quick")
        return best, {f: getattr(st, f) for f, _ in _QuickStats._fields_}

    def q3(self, seed=1, max_evals=100, zero_limit=1000, best_window=1500, gap=0.002, gap_norm=5154247.0):
        """bm_drive_q3, the changed-vertex triple search of q3.cu: (best mapping, stats)."""
        prm = _Q3Params(seed, max_evals, zero_limit, best_window, gap, gap_norm)
        best = np.zeros(self.n + 1, np.int32)
        st = _Q3Stats()
        _check(self._L.bm_drive_q3(self._h, C.byref(prm), best.ctypes.data_as(C.c_void_p), C.byref(st)),
               "bm_drive_q3")
        return best, {f: getattr(st, f) for f, _ in _Q3Stats._fields_}

    def triples(self, seed=1, max_evals=100, batch=270000):
        """bm_drive_triples, the batched lexicographic triple pass of triple.cu: (best mapping, stats)."""
        prm = _TriplesParams(seed, max_evals, batch)
        best = np.zeros(self.n + 1, np.int32)
        st = _TriplesStats()
        _check(self._L.bm_drive_triples(self._h, C.byref(prm), best.ctypes.data_as(C.c_void_p), C.byref(st)),
               "bm_drive_triples")
        return best, {f: getattr(st, f) for f, _ in _TriplesStats._fields_}

    def apply_swap(self, a, b):
        _check(self._L.bm_apply_swap(self._h, a, b), "bm_apply_swap")

    def apply_swap_scored(self, a, b):
        """bm_apply_swap_scored: applies the swap, returns the exact score change."""
        ch = C.c_int64()
        _check(self._L.bm_apply_swap_scored(self._h, a, b, C.byref(ch)), "bm_apply_swap_scored")
        return ch.value

    def apply_rot3(self, m1, m2, m3, p):
        _check(self._L.bm_apply_rot3(self._h, m1, m2, m3, p), "bm_apply_rot3")

    def apply_perm(self, sset, perm):
        s, p = _i32(sset), _i32(perm)
        _check(self._L.bm_apply_perm(self._h, len(s), s, p), "bm_apply_perm")

    def moves_set(self, k, verts):
        """Upload a resident candidate list: k=2 pairs, k=3 triples, 4..8 sets."""
        v = _i32(verts).reshape(-1)
        _check(self._L.bm_moves_set(self._h, k, v, len(v) // k), "bm_moves_set")

    def moves_eval(self, mode=0):
        """Score the resident list on the device: (best index or -1, delta, aux)."""
        i, d, x = C.c_int64(), C.c_int32(), C.c_int32()
        _check(self._L.bm_moves_eval(self._h, mode, C.byref(i), C.byref(d), C.byref(x)), "bm_moves_eval")
        return i.value, d.value, x.value

    def moves_results(self, count):
        d = np.zeros(count, np.int32)
        x = np.zeros(count, np.int32)
        _check(self._L.bm_moves_results(self._h, d.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p)),
               "bm_moves_results")
        return d, x

    def sweep_config(self):
        """(kernel, dense_cols): 1 = dense slab kernel, 2 = sparse kernel, 0 = not chosen yet."""
        k, cols = C.c_int(), C.c_int()
        _check(self._L.bm_sweep_config(self._h, C.byref(k), C.byref(cols)), "bm_sweep_config")
        return k.value, cols.value

    # --- batched refinement
    def refine_configure(self, params=None, **kw):
        """bm_refine_configure with a RefineParams (or the keyword fields of refine_params)."""
        prm = params if params is not None else refine_params(**kw)
        _check(self._L.bm_refine_configure(self._h, C.byref(prm)), "bm_refine_configure")

    def refine_run(self, max_steps=1 << 30, move_budget=0, patience=2):
        st = _RefineStats()
        _check(self._L.bm_refine_run(self._h, max_steps, move_budget, patience, C.byref(st)), "bm_refine_run")
        return {f: getattr(st, f) for f, _ in _RefineStats._fields_}

    def refine_history(self, first, count):
        out = (_RefineStep * max(count, 1))()
        _check(self._L.bm_refine_history(self._h, first, count, out), "bm_refine_history")
        return [(h.score, h.gain, h.moves, h.ok, h.moved, h.full) for h in out[:count]]

    def refine_log(self, capacity):
        _check(self._L.bm_refine_log(self._h, capacity), "bm_refine_log")

    def refine_moves(self, capacity=1 << 20):
        """Applied-move records of the last run: int32 array (count, 22) =
        step, kind, k, delta, round, base delta, vertices[8], new images[8]."""
        cnt = C.c_int64()
        _check(self._L.bm_refine_moves(self._h, C.byref(cnt), None, 0), "bm_refine_moves")
        m = min(cnt.value, capacity)
        rec = np.zeros((max(m, 1), 22), np.int32)
        _check(self._L.bm_refine_moves(self._h, C.byref(cnt), rec.ctypes.data_as(C.c_void_p), m), "bm_refine_moves")
        return rec[:m], cnt.value

    def refine_config(self):
        a, b = C.c_int32(), C.c_int32()
        _check(self._L.bm_refine_config(self._h, C.byref(a), C.byref(b)), "bm_refine_config")
        return a.value, b.value

    # --- timing
    def set_timing(self, on=True):
        _check(self._L.bm_set_timing(self._h, 1 if on else 0), "bm_set_timing")

    def reset_timing(self):
        _check(self._L.bm_reset_timing(self._h), "bm_reset_timing")

    def kernel_time(self, name):
        n, ms = C.c_int64(), C.c_double()
        _check(self._L.bm_kernel_time(self._h, name.encode(), C.byref(n), C.byref(ms)), "bm_kernel_time")
        return n.value, ms.value

    # --- drivers
    def _drive(self, fn, *args, acc_cap):
        acc = np.zeros(3 * max(acc_cap, 1), np.int32)
        st = _DriveStats()
        _check(fn(self._h, *args, acc.ctypes.data_as(C.c_void_p), acc_cap, C.byref(st)), fn.__name__)
        k = min(st.accepted, acc_cap)
        return acc[: 3 * k].reshape(k, 3), {f: getattr(st, f) for f, _ in _DriveStats._fields_}

    def first_improvement(self, rows_per_sweep=0, max_sweeps=0, verify=True, acc_cap=1 << 20, max_rows=0,
                          stop=None):
        """bm_drive_first_improvement; max_rows > 0 stops after that many rows, stop(rows) -> True
        stops after the current row (both through the row hook)."""
        hook = None
        if max_rows > 0 or stop is not None:
            hook = _ROW_HOOK(lambda user, rows, row, score, imp:
                             1 if (max_rows > 0 and rows > max_rows) or (stop is not None and stop(rows)) else 0)
        return self._drive(self._L.bm_drive_first_improvement, rows_per_sweep, max_sweeps, int(verify),
                           C.cast(hook, C.c_void_p) if hook else None, None, acc_cap=acc_cap)

    def row_best(self, max_sweeps=0, verify=True, acc_cap=1 << 20):
        return self._drive(self._L.bm_drive_row_best, max_sweeps, int(verify), None, None, acc_cap=acc_cap)

    def global_best(self, max_steps=0, verify=True, acc_cap=1 << 20, kind=BM_DELTA_LISTS):
        return self._drive(self._L.bm_drive_global_best, kind, max_steps, int(verify), acc_cap=acc_cap)


def refine_run_many(engines, max_steps=1 << 30, move_budget=0, patience=2):
    """bm_refine_run_many: the configured engines' runs packed into one graph launch per chunk."""
    L = lib()
    arr = (C.c_void_p * len(engines))(*[e._h.value for e in engines])
    st = (_RefineStats * len(engines))()
    _check(L.bm_refine_run_many(arr, len(engines), max_steps, move_budget, patience, st), "bm_refine_run_many")
    return [{f: getattr(x, f) for f, _ in _RefineStats._fields_} for x in st]


def job_footprint(n, e1, e2, params):
    """Device bytes the server reserves for a job of this shape."""
    return lib().bm_job_footprint(n, e1, e2, C.byref(params))


class Server:
    """bm_server: refinement jobs admitted against an HBM budget and run
    concurrently on one device (include/bmatch.h, serving loop)."""

    def __init__(self, device=0, max_active=8, prep_threads=4, hbm_budget=0):
        self._L = lib()
        h = C.c_void_p()
        cfg = _ServerConfig(max_active, prep_threads, hbm_budget)
        _check(self._L.bm_server_create(device, C.byref(cfg), C.byref(h)), "bm_server_create")
        self._h = h

    def submit(self, n, g1, g2, mapping, params, move_budget=0, max_steps=0, patience=2):
        """Queue a job (the arrays are copied before this returns); returns its id."""
        a = [_i32(x) for x in (*g1, *g2, mapping)]
        if len(a[6]) != n + 1:
            raise ValueError("mapping needs n+1 entries")
        ptr = [x.ctypes.data_as(C.c_void_p) for x in a]
        spec = _JobSpec(n, ptr[0], ptr[1], ptr[2], len(a[0]), ptr[3], ptr[4], ptr[5], len(a[3]), ptr[6],
                        move_budget, max_steps, patience, params)
        jid = C.c_int64()
        _check(self._L.bm_server_submit(self._h, C.byref(spec), C.byref(jid)), "bm_server_submit")
        return jid.value

    def wait(self, job_id, n=None):
        """Block until the job is done: (result dict, final mapping or None)."""
        res = _JobResult()
        mp = np.zeros(n + 1, np.int32) if n is not None else None
        _check(self._L.bm_server_wait(self._h, job_id, C.byref(res),
                                      mp.ctypes.data_as(C.c_void_p) if mp is not None else None), "bm_server_wait")
        out = {f: getattr(res, f) for f, _ in _JobResult._fields_}
        out["error"] = out["error"].decode(errors="replace")
        return out, (mp if out["status"] == 0 else None)

    def stats(self):
        st = _ServerStats()
        _check(self._L.bm_server_stats_get(self._h, C.byref(st)), "bm_server_stats_get")
        return {f: getattr(st, f) for f, _ in _ServerStats._fields_}

    def close(self):
        if self._h:
            self._L.bm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
