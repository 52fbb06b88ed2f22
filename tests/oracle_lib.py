"""ctypes binding of the CPU oracle (oracle/liboracle.so).  Test infrastructure:
only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it."""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
SO = ROOT / "oracle" / "liboracle.so"

f32, f64, i32, u32, u64 = ctypes.c_float, ctypes.c_double, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64
P = ctypes.POINTER


class PtsaCfg(ctypes.Structure):
    _fields_ = [("n", i32), ("L0", f64), ("init", P(f64)), ("replicas", i32), ("ladder", i32),
                ("tmin", f64), ("tmax", f64), ("move_sigma", f64), ("rot_sigma", f64),
                ("target_accept", f64), ("sweeps", ctypes.c_int64), ("swap_every", i32),
                ("shrink", f64), ("seed", u64), ("threads", i32)]


class PtsaOut(ctypes.Structure):
    _fields_ = [("L_final", f64), ("L_feasible", f64), ("proposals", u64), ("accepted", u64),
                ("exchanges_tried", u64), ("exchanges_accepted", u64), ("shrinks", i32),
                ("seconds", f64), ("best_q", P(f64)), ("e_min", f64)]


class SweepCfg(ctypes.Structure):
    _fields_ = [("n", i32), ("L", f32), ("w", f32), ("step_xy", f32), ("step_th", f32),
                ("inv_t_log2e", f32), ("key0", u32), ("key1", u32), ("replica", u32), ("sweep", u32),
                ("sweeps_per_shift", u32)]


def ptr(a, t):
    return a.ctypes.data_as(P(t))


class Oracle:
    def __init__(self, lib):
        self.lib = lib
        L = lib
        L.o_philox.argtypes = [P(u32), P(u32), P(u32)]
        L.o_u01f.argtypes = [u32]
        L.o_u01f.restype = f32
        L.o_ref_cell.argtypes = [f32, f32]
        L.o_ref_depth32.argtypes = [P(f32), P(f32)]
        L.o_ref_depth32.restype = f32
        L.o_ref_wall32.argtypes = [P(f32), f32]
        L.o_ref_wall32.restype = f32
        L.o_ref_local32.argtypes = [i32, P(f32), f32, P(f32), P(f32)]
        L.o_cellcoord.argtypes = [f32, f32, f32, i32, i32]
        for fn in ("o_depth32_cs", "o_depth2_32_cs"):
            getattr(L, fn).argtypes = [f32] * 6
            getattr(L, fn).restype = f32
        L.o_wall32_cs.argtypes = [f32] * 5
        L.o_wall32_cs.restype = f32
        L.o_rot32.argtypes = [f32, f32, f32, P(f32), P(f32)]
        L.o_accept_exp2.argtypes = [f32]
        L.o_accept_exp2.restype = f32
        L.o_pair64.argtypes = [P(f64), P(f64), P(f64)]
        L.o_pair64.restype = f64
        L.o_boundary64.argtypes = [P(f64), f64, P(f64)]
        L.o_boundary64.restype = f64
        L.o_sat_oracle64.argtypes = [P(f64), P(f64)]
        L.o_sat_oracle64.restype = f64
        L.o_sat_oracle_depth_f.argtypes = [P(f32), P(f32)]
        L.o_sat_oracle_depth_f.restype = f64
        L.o_total64.argtypes = [i32, P(f64), f64]
        L.o_total64.restype = f64
        L.o_total64_cs.argtypes = [i32, P(f64), P(f64), P(f64), P(f64), f64]
        L.o_total64_cs.restype = f64
        L.o_certified_side.argtypes = [i32, P(f64), f64, P(f64), P(f64)]
        L.o_certified_side.restype = f64
        L.o_polish64.argtypes = [i32, P(f64), f64, i32, P(f64)]
        L.o_wrap_angle.argtypes = [f64]
        L.o_wrap_angle.restype = f64
        L.o_ptsa_run.argtypes = [P(PtsaCfg), P(PtsaOut)]
        L.o_sweep32_model.argtypes = [P(SweepCfg), P(f32), P(f32), P(f32), P(f32), P(u64)]
        L.o_energy32_model.argtypes = [i32, f32, f32, P(f32), P(f32), P(f32), P(f32)]
        L.o_energy32_model.restype = f32
        L.o_collective32_model.argtypes = [P(SweepCfg), i32, P(f32), P(f32), P(f32), P(f32), P(f32)]
        L.o_kick32_model.argtypes = [P(SweepCfg), u32, i32, P(f32), P(f32), P(f32), P(f32), P(f32)]
        L.o_grid_packing.argtypes = [i32, P(f64)]

    # -- thin wrappers -------------------------------------------------------
    def philox(self, ctr, key):
        c = np.asarray(ctr, dtype=np.uint32)
        k = np.asarray(key, dtype=np.uint32)
        o = np.zeros(4, dtype=np.uint32)
        self.lib.o_philox(ptr(c, u32), ptr(k, u32), ptr(o, u32))
        return [int(v) for v in o]

    def pair64(self, p, q):
        p = np.asarray(p, np.float64); q = np.asarray(q, np.float64)
        g = np.zeros(3)
        e = self.lib.o_pair64(ptr(p, f64), ptr(q, f64), ptr(g, f64))
        return e, g

    def total64(self, q, L):
        q = np.ascontiguousarray(q, np.float64)
        return self.lib.o_total64(q.shape[0], ptr(q, f64), L)

    def certified_side(self, q, best=1e18):
        q = np.ascontiguousarray(q, np.float64)
        out = np.zeros_like(q)
        lam = ctypes.c_double()
        side = self.lib.o_certified_side(q.shape[0], ptr(q, f64), best, ptr(out, f64), ctypes.byref(lam))
        return side, lam.value, out

    def polish64(self, q, L, iters):
        q = np.ascontiguousarray(q, np.float64).copy()
        e = ctypes.c_double()
        it = self.lib.o_polish64(q.shape[0], ptr(q, f64), L, iters, ctypes.byref(e))
        return q, e.value, it

    def sweep_model(self, x, y, c, s, L, w, step_xy, step_th, inv_t_log2e, key0, key1, replica, sweep,
                    shift_every=1):
        cfg = SweepCfg(len(x), L, w, step_xy, step_th, inv_t_log2e, key0, key1, replica, sweep, shift_every)
        counts = np.zeros(3, dtype=np.uint64)
        acc = self.lib.o_sweep32_model(ctypes.byref(cfg), ptr(x, f32), ptr(y, f32), ptr(c, f32), ptr(s, f32),
                                       ptr(counts, u64))
        return acc, counts

    def energy32(self, x, y, c, s, L, w):
        return self.lib.o_energy32_model(len(x), L, w, ptr(x, f32), ptr(y, f32), ptr(c, f32), ptr(s, f32))

    def collective_model(self, x, y, c, s, L, w, step_xy, step_th, inv_t_log2e, key0, key1, replica, sweep_end,
                         moves):
        cfg = SweepCfg(len(x), L, w, step_xy, step_th, inv_t_log2e, key0, key1, replica, sweep_end)
        e = ctypes.c_float()
        acc = self.lib.o_collective32_model(ctypes.byref(cfg), moves, ptr(x, f32), ptr(y, f32), ptr(c, f32),
                                            ptr(s, f32), ctypes.byref(e))
        return acc, e.value

    def kick_model(self, x, y, c, s, L, w, key0, key1, replica, interval, relocate):
        cfg = SweepCfg(len(x), L, w, 0.0, 0.0, 0.0, key0, key1, replica, 0)
        e = ctypes.c_float()
        k = self.lib.o_kick32_model(ctypes.byref(cfg), interval, int(relocate), ptr(x, f32), ptr(y, f32),
                                    ptr(c, f32), ptr(s, f32), ctypes.byref(e))
        return k, e.value

    def ptsa(self, n, L0, replicas, tmin, tmax, sweeps, swap_every=10, shrink=1e-4, seed=42, threads=1,
             ladder=0, move_sigma=0.05, rot_sigma=0.05, target_accept=0.3, init=None):
        best = np.zeros((n, 3))
        cfg = PtsaCfg(n, L0, None if init is None else ptr(init, f64), replicas, ladder or replicas, tmin, tmax,
                      move_sigma, rot_sigma, target_accept, sweeps, swap_every, shrink, seed, threads)
        out = PtsaOut()
        out.best_q = ptr(best, f64)
        rc = self.lib.o_ptsa_run(ctypes.byref(cfg), ctypes.byref(out))
        if rc:
            raise RuntimeError(f"o_ptsa_run failed: {rc}")
        res = {name: getattr(out, name) for name, _ in PtsaOut._fields_ if name != "best_q"}
        res["best_q"] = best
        return res


_oracle = None


def load() -> Oracle:
    global _oracle
    if _oracle is None:
        if not SO.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle"), "oracle"], check=True)
        _oracle = Oracle(ctypes.CDLL(str(SO)))
    return _oracle
