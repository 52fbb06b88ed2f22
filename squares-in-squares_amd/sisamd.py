"""Python host mirror of the PT-SA engine's C ABI (include/sis_abi.h).

The reference drives its search as a process (``packing seed.txt epochs steps
chains rng_seed out.json``, cuda/packing.cu:768-868) from Python scripts such
as cuda/cpp_loop.py:47-80 and cuda/demo.py:84-86.  This module offers the same
operations in-process on top of ``lib/libsisamd.so``: ``solve`` (plain form),
``solve_batch`` (batched HIP-graph form), the serving loop, the certifier and
the device parity probes.  Every call goes to the HIP library; there is no
CPU fallback, and a missing library raises ``LibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsisamd.so")
PACKING_BIN = os.path.join(HERE, "bin", "packing")

SIS_OK, SIS_PENDING, SIS_EINVAL, SIS_EDEVICE, SIS_ETIMEOUT = 0, 1, 2, 3, 4
# sis_params.flags: count every proposal's near pairs (a diagnostic, about 3 %
# of sweep throughput; without it Result.near_pairs is 0)
FLAG_NEAR_PAIRS = 1


class LibraryMissing(RuntimeError):
    pass


class SisError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sis error {code}: {msg}")
        self.code = code


class SisInstance(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("L", ctypes.c_double), ("q", ctypes.POINTER(ctypes.c_double))]


class SisParams(ctypes.Structure):
    _fields_ = [
        ("tmin", ctypes.c_double), ("tmax", ctypes.c_double),
        ("move_sigma", ctypes.c_double), ("rot_sigma", ctypes.c_double),
        ("target_accept", ctypes.c_double), ("shrink", ctypes.c_double),
        ("cell", ctypes.c_double), ("init_jitter", ctypes.c_double),
        ("sweeps", ctypes.c_int64), ("replicas", ctypes.c_int), ("ladder", ctypes.c_int),
        ("sweeps_per_exchange", ctypes.c_int), ("replicas_per_group", ctypes.c_int),
        ("polish_iters", ctypes.c_int), ("init_mode", ctypes.c_int), ("flags", ctypes.c_int),
        ("rng_seed", ctypes.c_uint64), ("collective_moves", ctypes.c_int), ("kick_every", ctypes.c_int),
        ("sweeps_per_shift", ctypes.c_int), ("improve", ctypes.c_int), ("cool_period", ctypes.c_int),
        ("certify_every", ctypes.c_int),
    ]


class SisResult(ctypes.Structure):
    _fields_ = [
        ("q_out", ctypes.POINTER(ctypes.c_double)),
        ("target_L", ctypes.c_double), ("L_cert", ctypes.c_double), ("lambda_", ctypes.c_double),
        ("e32", ctypes.c_double), ("e64_before", ctypes.c_double), ("e64_after", ctypes.c_double),
        ("e_gap", ctypes.c_double),
        ("proposals", ctypes.c_uint64), ("accepted", ctypes.c_uint64), ("sweeps", ctypes.c_uint64),
        ("exchanges_tried", ctypes.c_uint64), ("exchanges_accepted", ctypes.c_uint64),
        ("candidates", ctypes.c_uint64), ("near_pairs", ctypes.c_uint64),
        ("shrinks", ctypes.c_int), ("feasible", ctypes.c_int), ("replicas", ctypes.c_int),
        ("status", ctypes.c_int), ("seconds", ctypes.c_double), ("device_seconds", ctypes.c_double),
        ("fp64_rejects", ctypes.c_int), ("kicks", ctypes.c_int),
        ("collective_tried", ctypes.c_uint64), ("collective_accepted", ctypes.c_uint64),
        ("feasible_checkpoints", ctypes.c_uint64),
    ]


class SisServerCfg(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_double), ("max_instances", ctypes.c_int),
        ("intervals_per_round", ctypes.c_int), ("certifier_threads", ctypes.c_int),
        ("certify", ctypes.c_int), ("request_timeout_s", ctypes.c_double),
    ]


# Per-replica and per-slot records of the engine (sis_rep_state, sis_slot_state).
REP_DTYPE = np.dtype([("E", "<f4"), ("slot", "<i4"), ("acc", "<u4"), ("pad", "<u4"), ("Lc", "<f8"),
                      ("lam", "<f8"), ("minx", "<f8"), ("miny", "<f8"),
                      ("cand", "<u8"), ("near_pairs", "<u8"), ("coll_tried", "<u4"), ("coll_acc", "<u4"),
                      ("kicks", "<u4"), ("pad2", "<u4")])
SLOT_DTYPE = np.dtype([("inv_t_log2e", "<f4"), ("step_xy", "<f4"), ("step_th", "<f4"), ("L", "<f4"),
                       ("T", "<f4"), ("base_xy", "<f4"), ("base_th", "<f4"), ("pad", "<f4")])


class SisLadderDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("R", ctypes.c_int), ("ladder", ctypes.c_int), ("improve", ctypes.c_int),
        ("cool_period", ctypes.c_int), ("spe", ctypes.c_int),
        ("sweeps_total", ctypes.c_int64), ("sweeps_done", ctypes.c_int64),
        ("key0", ctypes.c_uint32), ("key1", ctypes.c_uint32), ("interval", ctypes.c_uint32),
        ("shrink", ctypes.c_float), ("target_acc", ctypes.c_float), ("scale", ctypes.c_float),
        ("Lf", ctypes.c_float),
        ("L", ctypes.c_double), ("L_arch", ctypes.c_double), ("L_box_arch", ctypes.c_double),
        ("arch_valid", ctypes.c_int), ("shrinks", ctypes.c_int), ("feasible", ctypes.c_int),
        ("fp64_rejects", ctypes.c_int),
        ("ex_tried", ctypes.c_uint64), ("ex_acc", ctypes.c_uint64), ("certs", ctypes.c_uint64),
    ]


def desc_dict(d) -> dict:
    return {name: getattr(d, name) for name, _ in d._fields_}


_lib = None


def lib() -> ctypes.CDLL:
    """Load libsisamd.so (built in-tree by ``make -C squares-in-squares_amd``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} not built: run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    dp, fp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    L.sis_init.argtypes = [ctypes.c_int]
    L.sis_last_error.restype = ctypes.c_char_p
    L.sis_last_error_code.restype = ctypes.c_int
    L.sis_device_info.argtypes = [ip, ctypes.c_char_p, ctypes.c_int]
    L.sis_default_params.argtypes = [ctypes.POINTER(SisParams)]
    L.sis_default_params.restype = None
    L.sis_portfolio.argtypes = [ctypes.POINTER(SisInstance), ctypes.POINTER(SisParams), ctypes.c_int, dp]
    L.sis_solve.argtypes = [ctypes.POINTER(SisInstance), ctypes.POINTER(SisParams), ctypes.POINTER(SisResult)]
    L.sis_solve_noalloc.argtypes = [ctypes.POINTER(SisInstance), ctypes.POINTER(SisParams),
                                    ctypes.POINTER(SisResult)]
    L.sis_solve_batch.argtypes = [ctypes.POINTER(SisInstance), ctypes.c_int, ctypes.POINTER(SisParams),
                                  ctypes.POINTER(SisResult)]
    L.sis_autotune.argtypes = [ctypes.POINTER(SisInstance), ctypes.c_int, ctypes.POINTER(SisParams), ctypes.c_int,
                               ctypes.c_void_p]
    L.sis_batch_create.argtypes = [ctypes.POINTER(SisInstance), ctypes.c_int, ctypes.POINTER(SisParams)]
    L.sis_batch_create.restype = ctypes.c_void_p
    L.sis_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, dp,
                                ctypes.POINTER(ctypes.c_int64), dp]
    L.sis_batch_finalize.argtypes = [ctypes.c_void_p, ctypes.POINTER(SisResult)]
    L.sis_batch_destroy.argtypes = [ctypes.c_void_p]
    L.sis_server_create.argtypes = [ctypes.POINTER(SisServerCfg)]
    L.sis_server_create.restype = ctypes.c_void_p
    L.sis_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(SisInstance), ctypes.POINTER(SisParams), u64p]
    L.sis_submit_timeout.argtypes = [ctypes.c_void_p, ctypes.POINTER(SisInstance), ctypes.POINTER(SisParams),
                                     ctypes.c_double, u64p]
    L.sis_server_stats.argtypes = [ctypes.c_void_p, dp, dp, dp, ip]
    L.sis_server_inject_lane_fault.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.sis_poll.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(SisResult)]
    L.sis_ticket_times.argtypes = [ctypes.c_void_p, ctypes.c_uint64, dp, dp, dp, ctypes.c_char_p, ctypes.c_int]
    L.sis_ticket_release.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    L.sis_server_destroy.argtypes = [ctypes.c_void_p]
    L.sis_certify_strings.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                      ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p),
                                      ctypes.c_char_p, ctypes.c_int]
    L.sis_export_certify.argtypes = [ctypes.c_int, dp, dp, ctypes.c_char_p, ctypes.c_int]
    L.sis_dev_depth32.argtypes = [ctypes.c_int, fp, fp, fp]
    L.sis_dev_cellcoord.argtypes = [ctypes.c_int, fp, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                    ctypes.c_int, ip]
    L.sis_dev_accept_exp2.argtypes = [ctypes.c_int, fp, fp]
    L.sis_dev_sweeps.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                 fp, fp, fp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                 ctypes.c_int, fp, ip, fp, u64p]
    L.sis_dev_epoch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                fp, fp, fp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, fp, ip, fp, u64p]
    L.sis_dev_polish.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp, dp, dp, dp, dp]
    vp = ctypes.c_void_p
    L.sis_dev_ladder.argtypes = [ctypes.POINTER(SisLadderDesc), vp, vp, ip, fp, fp]
    L.sis_dev_batch_state.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(SisLadderDesc), vp, vp, ip,
                                      fp, fp]
    L.sis_batch_status.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(SisLadderDesc)]
    L.sis_dev_launch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, fp, ctypes.c_float, ctypes.c_int,
                                 fp, fp, fp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int, fp, ip,
                                 fp, u64p, vp]
    L.sis_dev_zero64.argtypes = [ctypes.c_int, ctypes.c_double, fp, ip, dp]
    _lib = L
    return L


def _err(code: int):
    raise SisError(code, lib().sis_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def default_params(**kw) -> SisParams:
    p = SisParams()
    lib().sis_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def with_near_pairs(p: SisParams) -> SisParams:
    """A copy of p that also counts near pairs (FLAG_NEAR_PAIRS); the
    search itself is unchanged."""
    q = SisParams.from_buffer_copy(p)
    q.flags |= FLAG_NEAR_PAIRS
    return q


def params_dict(p: SisParams) -> dict:
    return {name: getattr(p, name) for name, _ in SisParams._fields_}


def device_info() -> tuple[int, str]:
    cus = ctypes.c_int(0)
    name = ctypes.create_string_buffer(256)
    rc = lib().sis_device_info(ctypes.byref(cus), name, 256)
    if rc:
        _err(rc)
    return cus.value, name.value.decode()


@dataclass
class Result:
    q: np.ndarray
    fields: dict = field(default_factory=dict)

    def __getattr__(self, k):
        try:
            return self.fields[k]
        except KeyError as e:
            raise AttributeError(k) from e


def _result(r: SisResult, q: np.ndarray) -> Result:
    f = {name: getattr(r, name) for name, _ in SisResult._fields_ if name != "q_out"}
    f["lambda"] = f.pop("lambda_")
    return Result(q=q.copy(), fields=f)


def _instance(n: int, L: float, q) -> tuple[SisInstance, np.ndarray | None]:
    if q is None:
        return SisInstance(n, L, None), None
    arr = np.ascontiguousarray(np.asarray(q, dtype=np.float64).reshape(n, 3))
    return SisInstance(n, L, _ptr(arr, ctypes.c_double)), arr


def portfolio(n: int, L: float, q=None, params: SisParams | None = None, jitter: bool = True, **kw) -> np.ndarray:
    """Starting configurations of init_mode 2 (host only): (replicas, n, 3)."""
    p = params if params is not None else default_params(**kw)
    inst, _keep = _instance(n, L, q)
    out = np.zeros((p.replicas, n, 3), dtype=np.float64)
    rc = lib().sis_portfolio(ctypes.byref(inst), ctypes.byref(p), int(jitter), _ptr(out, ctypes.c_double))
    if rc:
        _err(rc)
    return out


def solve(n: int, L: float, q=None, params: SisParams | None = None, **kw) -> Result:
    """Plain form: one instance, the host waits for the end of every exchange
    interval (the next interval already queued)."""
    p = params if params is not None else default_params(**kw)
    inst, keep = _instance(n, L, q)
    out = np.zeros((n, 3), dtype=np.float64)
    r = SisResult()
    r.q_out = _ptr(out, ctypes.c_double)
    rc = lib().sis_solve(ctypes.byref(inst), ctypes.byref(p), ctypes.byref(r))
    if rc:
        _err(rc)
    del keep
    return _result(r, out)


def solve_noalloc(n: int, L: float, q=None, params: SisParams | None = None, **kw) -> Result:
    """The plain form without device allocation (sis_solve_noalloc)."""
    p = params if params is not None else default_params(**kw)
    inst, keep = _instance(n, L, q)
    out = np.zeros((n, 3), dtype=np.float64)
    r = SisResult()
    r.q_out = _ptr(out, ctypes.c_double)
    rc = lib().sis_solve_noalloc(ctypes.byref(inst), ctypes.byref(p), ctypes.byref(r))
    if rc:
        _err(rc)
    del keep
    return _result(r, out)


def solve_batch(instances, params) -> list[Result]:
    """Batched form: instances = [(n, L, q|None)], params = SisParams or list."""
    k = len(instances)
    if isinstance(params, SisParams):
        params = [params] * k
    insts = (SisInstance * k)()
    ps = (SisParams * k)(*params)
    keep, outs = [], []
    res = (SisResult * k)()
    for j, (n, L, q) in enumerate(instances):
        insts[j], arr = _instance(n, L, q)
        keep.append(arr)
        o = np.zeros((n, 3), dtype=np.float64)
        outs.append(o)
        res[j].q_out = _ptr(o, ctypes.c_double)
    rc = lib().sis_solve_batch(insts, k, ps, res)
    if rc:
        _err(rc)
    return [_result(res[j], outs[j]) for j in range(k)]


class SisTuning(ctypes.Structure):
    _fields_ = [("replicas_per_group", ctypes.c_int), ("sweeps_per_exchange", ctypes.c_int),
                ("cell", ctypes.c_double), ("proposals_per_s", ctypes.c_double),
                ("candidates_tried", ctypes.c_int), ("trajectory_kept", ctypes.c_int),
                ("quality_checked", ctypes.c_int), ("quality_sweeps", ctypes.c_int),
                ("checked_sweeps_per_exchange", ctypes.c_int), ("checked_cell", ctypes.c_double),
                ("quality_delta", ctypes.c_double), ("quality_se", ctypes.c_double)]


def autotune(instances, params, probe_sweeps: int = 20) -> dict:
    """Fastest (replicas per workgroup, sweeps per exchange, cell width) for a
    batch whose search quality, where the trajectory changes, passes the
    paired quality probe against the caller's own settings."""
    k = len(instances)
    if isinstance(params, SisParams):
        params = [params] * k
    insts = (SisInstance * k)()
    keep = []
    for j, (n, L, q) in enumerate(instances):
        insts[j], arr = _instance(n, L, q)
        keep.append(arr)
    ps = (SisParams * k)(*params)
    out = SisTuning()
    rc = lib().sis_autotune(insts, k, ps, probe_sweeps, ctypes.byref(out))
    if rc:
        _err(rc)
    return {name: getattr(out, name) for name, _ in SisTuning._fields_}


class BatchSession:
    """Resident instances run step-wise (the epoch loop of cuda/packing.cu:823-853)."""

    def __init__(self, instances, params):
        k = len(instances)
        if isinstance(params, SisParams):
            params = [params] * k
        self.k = k
        self.ns = [n for n, _, _ in instances]
        insts = (SisInstance * k)()
        keep = []
        for j, (n, L, q) in enumerate(instances):
            insts[j], arr = _instance(n, L, q)
            keep.append(arr)
        ps = (SisParams * k)(*params)
        self.h = lib().sis_batch_create(insts, k, ps)
        if not self.h:
            _err(lib().sis_last_error_code())

    def run(self, intervals: int, mode: int = 0):
        """Returns (sweep_kernel_ms, sweep_launches, total_ms): in plain mode
        the HIP-event time and number of the sweep launches timed (every
        eighth launch of the call)."""
        sm, tm = ctypes.c_double(), ctypes.c_double()
        nl = ctypes.c_int64()
        rc = lib().sis_batch_run(self.h, intervals, mode, ctypes.byref(sm), ctypes.byref(nl), ctypes.byref(tm))
        if rc:
            _err(rc)
        return sm.value, nl.value, tm.value

    def finalize(self) -> list[Result]:
        res = (SisResult * self.k)()
        outs = []
        for j, n in enumerate(self.ns):
            o = np.zeros((n, 3), dtype=np.float64)
            outs.append(o)
            res[j].q_out = _ptr(o, ctypes.c_double)
        rc = lib().sis_batch_finalize(self.h, res)
        if rc:
            _err(rc)
        return [_result(res[j], outs[j]) for j in range(self.k)]

    def status(self, k: int = 0) -> dict:
        """Descriptor fields of instance k as the host last read them (no
        device copy): after a plain run, what the last exchange step wrote."""
        d = SisLadderDesc()
        rc = lib().sis_batch_status(self.h, k, ctypes.byref(d))
        if rc:
            _err(rc)
        return desc_dict(d)

    def state(self, k: int = 0) -> dict:
        """Device state of instance k: descriptor fields, replica and slot records,
        slot owners, replica states (R, n, 4) and archive (n, 4)."""
        d = SisLadderDesc()
        rc = lib().sis_dev_batch_state(self.h, k, ctypes.byref(d), None, None, None, None, None)
        if rc:
            _err(rc)
        R, n = d.R, d.n
        reps = np.zeros(R, REP_DTYPE)
        slots = np.zeros(R, SLOT_DTYPE)
        ros = np.zeros(R, np.int32)
        st = np.zeros((R, n, 4), np.float32)
        arch = np.zeros((n, 4), np.float32)
        rc = lib().sis_dev_batch_state(self.h, k, ctypes.byref(d), reps.ctypes.data, slots.ctypes.data,
                                       _ptr(ros, ctypes.c_int), _ptr(st, ctypes.c_float), _ptr(arch, ctypes.c_float))
        if rc:
            _err(rc)
        return dict(desc=desc_dict(d), reps=reps, slots=slots, ros=ros, state=st, arch=arch)

    def close(self):
        if self.h:
            lib().sis_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server (sis_server_*)."""

    def __init__(self, max_instances=256, intervals_per_round=10, certifier_threads=8, certify=True,
                 hbm_budget_bytes=4 * 2**30, request_timeout_s=0.0):
        cfg = SisServerCfg(hbm_budget_bytes, max_instances, intervals_per_round, certifier_threads, int(certify),
                           request_timeout_s)
        self.h = lib().sis_server_create(ctypes.byref(cfg))
        if not self.h:
            _err(lib().sis_last_error_code())
        self.ns = {}

    def submit(self, n, L, q=None, params: SisParams | None = None, timeout_s=None, **kw) -> int:
        """Ticket of a new request; timeout_s: its own deadline in seconds
        (0: none) instead of the server's request_timeout_s."""
        p = params if params is not None else default_params(**kw)
        inst, keep = _instance(n, L, q)
        t = ctypes.c_uint64()
        if timeout_s is None:
            rc = lib().sis_submit(self.h, ctypes.byref(inst), ctypes.byref(p), ctypes.byref(t))
        else:
            rc = lib().sis_submit_timeout(self.h, ctypes.byref(inst), ctypes.byref(p), float(timeout_s),
                                          ctypes.byref(t))
        if rc:
            _err(rc)
        del keep
        self.ns[t.value] = n
        return t.value

    def poll(self, ticket: int):
        """Result when finished, else None."""
        n = self.ns[ticket]
        out = np.zeros((n, 3), dtype=np.float64)
        r = SisResult()
        r.q_out = _ptr(out, ctypes.c_double)
        rc = lib().sis_poll(self.h, ticket, ctypes.byref(r))
        if rc == SIS_PENDING:
            return None
        if rc:
            _err(rc)
        return _result(r, out)

    def stats(self) -> dict:
        """Arena capacity and resident replica state (bytes), resident requests."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        k = ctypes.c_int()
        rc = lib().sis_server_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(k))
        if rc:
            _err(rc)
        return {"arena_bytes": a.value, "resident_bytes": b.value, "peak_resident_bytes": c.value,
                "resident_requests": k.value}

    def inject_lane_fault(self, lane: int):
        """Test hook: the next round of `lane` fails at its first launch."""
        rc = lib().sis_server_inject_lane_fault(self.h, lane)
        if rc:
            _err(rc)

    def times(self, ticket: int):
        """(queued_s, device_done_s, certified_s, L50 or rejection reason)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        buf = ctypes.create_string_buffer(128)
        rc = lib().sis_ticket_times(self.h, ticket, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), buf, 128)
        if rc not in (SIS_OK, SIS_PENDING):
            _err(rc)
        return a.value, b.value, c.value, buf.value.decode()

    def release(self, ticket: int) -> bool:
        """Forget a finished ticket; False (ticket kept) while it is pending."""
        rc = lib().sis_ticket_release(self.h, ticket)
        if rc == SIS_PENDING:
            return False
        if rc:
            _err(rc)
        self.ns.pop(ticket, None)
        return True

    def close(self):
        if self.h:
            lib().sis_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def certify_strings(rows) -> str:
    """Exact certificate of decimal-string rows [(x, y, angle)]: the 50-decimal
    L, or SisError (code 2) with the rejection reason."""
    n = len(rows)
    cols = [(ctypes.c_char_p * n)(*[str(r[k]).encode() for r in rows]) for k in range(3)]
    buf = ctypes.create_string_buffer(64)
    rc = lib().sis_certify_strings(n, cols[0], cols[1], cols[2], buf, 64)
    if rc:
        _err(rc)
    return buf.value.decode()


def export_certify(q: np.ndarray):
    """Dilation ladder + exact certificate; returns (dilation, L50) or raises."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    d = ctypes.c_double()
    buf = ctypes.create_string_buffer(64)
    rc = lib().sis_export_certify(q.shape[0], _ptr(q, ctypes.c_double), ctypes.byref(d), buf, 64)
    if rc:
        _err(rc)
    return d.value, buf.value.decode()


def dev_depth32(pairs: np.ndarray):
    """pairs: (N, 6) float32 rows dx, dy, ci, si, cj, sj -> (depth, depth^2)."""
    N = pairs.shape[0]
    buf = np.zeros((N, 8), dtype=np.float32)
    buf[:, :6] = pairs
    d = np.zeros(N, dtype=np.float32)
    d2 = np.zeros(N, dtype=np.float32)
    rc = lib().sis_dev_depth32(N, _ptr(buf, ctypes.c_float), _ptr(d, ctypes.c_float), _ptr(d2, ctypes.c_float))
    if rc:
        _err(rc)
    return d, d2


def dev_cellcoord(x: np.ndarray, shift: float, w: float, lo: int, hi: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(x.shape[0], dtype=np.int32)
    rc = lib().sis_dev_cellcoord(x.shape[0], _ptr(x, ctypes.c_float), shift, w, lo, hi, _ptr(out, ctypes.c_int))
    if rc:
        _err(rc)
    return out


def dev_accept_exp2(x: np.ndarray) -> np.ndarray:
    """2^x of the sweep's Metropolis test, computed on the device."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(x.shape[0], dtype=np.float32)
    rc = lib().sis_dev_accept_exp2(x.shape[0], _ptr(x, ctypes.c_float), _ptr(out, ctypes.c_float))
    if rc:
        _err(rc)
    return out


def dev_sweeps(xycs: np.ndarray, L: float, cell: float, inv_t_log2e, step_xy, step_th, key0: int,
               key1: int, sweep0: int, sweeps: int, group: int = 1, collective: int = 0, kick_every: int = 0,
               interval: int = 0, shift_every: int = 4):
    """xycs: (R, n, 4) float32 (x, y, cos, sin); returns (state, accepted, energy, counts).  The
    cell grid gets a new random shift every `shift_every` sweeps.  With collective / kick_every the
    launch ends with collective moves and a kick (counts then has six entries: proposals,
    candidates, near pairs, collective proposals, collective accepts, kicks)."""
    R, n, _ = xycs.shape
    st = np.ascontiguousarray(xycs, dtype=np.float32).copy()
    ik = np.ascontiguousarray(np.broadcast_to(np.asarray(inv_t_log2e, np.float32), (R,)))
    sx = np.ascontiguousarray(np.broadcast_to(np.asarray(step_xy, np.float32), (R,)))
    sth = np.ascontiguousarray(np.broadcast_to(np.asarray(step_th, np.float32), (R,)))
    acc = np.zeros(R, dtype=np.int32)
    en = np.zeros(R, dtype=np.float32)
    if collective or kick_every:
        counts = np.zeros(6, dtype=np.uint64)
        rc = lib().sis_dev_epoch(n, R, L, cell, group, _ptr(ik, ctypes.c_float), _ptr(sx, ctypes.c_float),
                                 _ptr(sth, ctypes.c_float), key0, key1, sweep0, sweeps, shift_every, collective,
                                 kick_every, interval, _ptr(st, ctypes.c_float), _ptr(acc, ctypes.c_int),
                                 _ptr(en, ctypes.c_float), _ptr(counts, ctypes.c_uint64))
    else:
        counts = np.zeros(3, dtype=np.uint64)
        rc = lib().sis_dev_sweeps(n, R, L, cell, group, _ptr(ik, ctypes.c_float), _ptr(sx, ctypes.c_float),
                                  _ptr(sth, ctypes.c_float), key0, key1, sweep0, sweeps, shift_every,
                                  _ptr(st, ctypes.c_float), _ptr(acc, ctypes.c_int), _ptr(en, ctypes.c_float),
                                  _ptr(counts, ctypes.c_uint64))
    if rc:
        _err(rc)
    return st, acc, en, counts


def dev_zero64(xycs: np.ndarray, L: float):
    """The exchange kernel's FP64 zero-overlap test of one configuration ((n, 4) float32 x, y,
    cos, sin) and the all-pairs FP64 energy of the same configuration: (zero, e64)."""
    q = np.ascontiguousarray(xycs, dtype=np.float32)
    n = q.shape[0]
    z, e = ctypes.c_int(0), ctypes.c_double(0.0)
    rc = lib().sis_dev_zero64(n, L, _ptr(q, ctypes.c_float), ctypes.byref(z), ctypes.byref(e))
    if rc:
        _err(rc)
    return bool(z.value), e.value


def dev_polish(q: np.ndarray, L: float, iters: int):
    """FP64 CG polish + certified side; returns dict with energies, lambda, L_cert, export."""
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3).copy()
    n = q.shape[0]
    eb, ea, lam, lc = (ctypes.c_double() for _ in range(4))
    out = np.zeros_like(q)
    rc = lib().sis_dev_polish(n, L, iters, _ptr(q, ctypes.c_double), ctypes.byref(eb), ctypes.byref(ea),
                              ctypes.byref(lam), ctypes.byref(lc), _ptr(out, ctypes.c_double))
    if rc:
        _err(rc)
    return dict(e_before=eb.value, e_after=ea.value, lam=lam.value, L_cert=lc.value, q=out)


def dev_ladder(desc: dict, reps: np.ndarray, slots: np.ndarray, ros: np.ndarray, xycs: np.ndarray,
               arch: np.ndarray):
    """One exchange step on the device (k_ladder) on copies of the given records;
    returns (desc, reps, slots, ros, xycs, arch) after it."""
    d = SisLadderDesc(**desc)
    reps = np.ascontiguousarray(reps, REP_DTYPE).copy()
    slots = np.ascontiguousarray(slots, SLOT_DTYPE).copy()
    ros = np.ascontiguousarray(ros, np.int32).copy()
    st = np.ascontiguousarray(xycs, np.float32).copy()
    ar = np.ascontiguousarray(arch, np.float32).copy()
    rc = lib().sis_dev_ladder(ctypes.byref(d), reps.ctypes.data, slots.ctypes.data, _ptr(ros, ctypes.c_int),
                              _ptr(st, ctypes.c_float), _ptr(ar, ctypes.c_float))
    if rc:
        _err(rc)
    return desc_dict(d), reps, slots, ros, st, ar


def dev_launch(xycs: np.ndarray, Lgrid: float, box, cell: float, inv_t_log2e, step_xy, step_th, key0: int,
               key1: int, sweep0: int, sweeps: int, group: int = 1, collective: int = 0, kick_every: int = 0,
               interval: int = 0, shift_every: int = 4, certify: bool = False):
    """One sweep launch with a box per replica; with certify, every replica's FP64
    certified side at the end.  Returns (state, accepted, energy, counts[6], reps)."""
    R, n, _ = xycs.shape
    st = np.ascontiguousarray(xycs, dtype=np.float32).copy()
    bx = np.ascontiguousarray(np.broadcast_to(np.asarray(box, np.float32), (R,)))
    ik = np.ascontiguousarray(np.broadcast_to(np.asarray(inv_t_log2e, np.float32), (R,)))
    sx = np.ascontiguousarray(np.broadcast_to(np.asarray(step_xy, np.float32), (R,)))
    sth = np.ascontiguousarray(np.broadcast_to(np.asarray(step_th, np.float32), (R,)))
    acc = np.zeros(R, dtype=np.int32)
    en = np.zeros(R, dtype=np.float32)
    counts = np.zeros(6, dtype=np.uint64)
    reps = np.zeros(R, REP_DTYPE)
    rc = lib().sis_dev_launch(n, R, Lgrid, _ptr(bx, ctypes.c_float), cell, group, _ptr(ik, ctypes.c_float),
                              _ptr(sx, ctypes.c_float), _ptr(sth, ctypes.c_float), key0, key1, sweep0, sweeps,
                              shift_every, collective, kick_every, interval, int(certify),
                              _ptr(st, ctypes.c_float), _ptr(acc, ctypes.c_int), _ptr(en, ctypes.c_float),
                              _ptr(counts, ctypes.c_uint64), reps.ctypes.data)
    if rc:
        _err(rc)
    return st, acc, en, counts, reps
