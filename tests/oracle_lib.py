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
                ("sweeps_per_shift", u32), ("dE32", P(f32)), ("dE64", P(f64)), ("depth64", P(f64))]


# o_rep_state / o_slot_state / o_ladder_desc of sis_oracle.h
REP_DTYPE = np.dtype([("E", "<f4"), ("slot", "<i4"), ("acc", "<u4"), ("pad", "<u4"), ("Lc", "<f8"),
                      ("lam", "<f8"), ("minx", "<f8"), ("miny", "<f8"),
                      ("cand", "<u8"), ("near_pairs", "<u8"), ("coll_tried", "<u4"), ("coll_acc", "<u4"),
                      ("kicks", "<u4"), ("pad2", "<u4")])
SLOT_DTYPE = np.dtype([("inv_t_log2e", "<f4"), ("step_xy", "<f4"), ("step_th", "<f4"), ("L", "<f4"),
                       ("T", "<f4"), ("base_xy", "<f4"), ("base_th", "<f4"), ("pad", "<f4")])


class LadderDesc(ctypes.Structure):
    _fields_ = [("n", i32), ("R", i32), ("ladder", i32), ("improve", i32), ("cool_period", i32), ("spe", i32),
                ("sweeps_total", ctypes.c_int64), ("sweeps_done", ctypes.c_int64),
                ("key0", u32), ("key1", u32), ("interval", u32),
                ("shrink", f32), ("target_acc", f32), ("scale", f32), ("Lf", f32),
                ("L", f64), ("L_arch", f64), ("L_box_arch", f64),
                ("arch_valid", i32), ("shrinks", i32), ("feasible", i32), ("fp64_rejects", i32),
                ("ex_tried", u64), ("ex_acc", u64), ("certs", u64)]


def ptr(a, t):
    return a.ctypes.data_as(P(t))


class Oracle:
    def __init__(self, lib):
        self.lib = lib
        L = lib
        L.o_philox.argtypes = [P(u32), P(u32), P(u32)]
        L.o_u01f.argtypes = [u32]
        L.o_u01f.restype = f32
        L.o_u01d.argtypes = [u32, u32]
        L.o_u01d.restype = f64
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
        L.o_certify_xycs.argtypes = [i32, P(f32), P(f64)]
        L.o_ladder_model.argtypes = [P(LadderDesc), ctypes.c_void_p, ctypes.c_void_p, P(i32), P(f32), P(f32)]
        L.o_exp2_fixed64.argtypes = [f64]
        L.o_exp2_fixed64.restype = f64
        L.o_cool_factor.argtypes = [u32, i32, P(f64), P(f64)]

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
                    shift_every=1, per_move=False):
        """One sweep in place.  With per_move, also returns the FP32 energy change of
        every square's proposal, its FP64 recompute over all pairs and the summed
        overlap depths (NaN where the proposal was rejected before the stencil)."""
        cfg = SweepCfg(len(x), L, w, step_xy, step_th, inv_t_log2e, key0, key1, replica, sweep, shift_every)
        moves = None
        if per_move:
            n = len(x)
            moves = (np.zeros(n, np.float32), np.zeros(n, np.float64), np.zeros(n, np.float64))
            cfg.dE32, cfg.dE64, cfg.depth64 = ptr(moves[0], f32), ptr(moves[1], f64), ptr(moves[2], f64)
        counts = np.zeros(3, dtype=np.uint64)
        acc = self.lib.o_sweep32_model(ctypes.byref(cfg), ptr(x, f32), ptr(y, f32), ptr(c, f32), ptr(s, f32),
                                       ptr(counts, u64))
        if per_move:
            return acc, counts, moves
        return acc, counts

    def certify_xycs(self, q):
        """FP64 certified side of an (n, 4) float32 state: (side, lam, minx, miny)."""
        q = np.ascontiguousarray(q, np.float32)
        out = np.zeros(4)
        self.lib.o_certify_xycs(q.shape[0], ptr(q, f32), ptr(out, f64))
        return tuple(float(v) for v in out)

    def ladder(self, desc: dict, reps, slots, ros, xycs, arch):
        """One exchange step (o_ladder_model) on copies; returns the updated
        (desc, reps, slots, ros, xycs, arch) and whether it improved."""
        d = LadderDesc(**desc)
        reps = np.ascontiguousarray(reps, REP_DTYPE).copy()
        slots = np.ascontiguousarray(slots, SLOT_DTYPE).copy()
        ros = np.ascontiguousarray(ros, np.int32).copy()
        st = np.ascontiguousarray(xycs, np.float32).copy()
        ar = np.ascontiguousarray(arch, np.float32).copy()
        imp = self.lib.o_ladder_model(ctypes.byref(d), reps.ctypes.data, slots.ctypes.data, ptr(ros, i32),
                                      ptr(st, f32), ptr(ar, f32))
        out = {name: getattr(d, name) for name, _ in LadderDesc._fields_}
        return out, reps, slots, ros, st, ar, bool(imp)

    def cool_factor(self, k, P):
        c, r = ctypes.c_double(), ctypes.c_double()
        self.lib.o_cool_factor(k, P, ctypes.byref(c), ctypes.byref(r))
        return c.value, r.value

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
