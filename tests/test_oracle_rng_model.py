"""Philox4x32-10 known answers, rocRAND cross-check, and the CPU sweep model's
invariants (no GPU)."""
from __future__ import annotations

import math
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle_lib import f32, f64, ptr

# Random123 kat_vectors, philox4x32 with 10 rounds: (ctr, key) -> out
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(oracle, ctr, key, out):
    assert tuple(oracle.philox(ctr, key)) == out


ROCRAND_CHECK = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
struct Open : rocrand_device::philox4x32_10_engine {
  using rocrand_device::philox4x32_10_engine::ten_rounds;
};
int main() {
  Open e;
  unsigned s = 12345u;
  for (int k = 0; k < 64; ++k) {
    uint4 c; uint2 key;
    s = s * 1664525u + 1013904223u; c.x = s; s = s * 1664525u + 1013904223u; c.y = s;
    s = s * 1664525u + 1013904223u; c.z = s; s = s * 1664525u + 1013904223u; c.w = s;
    s = s * 1664525u + 1013904223u; key.x = s; s = s * 1664525u + 1013904223u; key.y = s;
    uint4 o = e.ten_rounds(c, key);
    printf("%u %u %u %u %u %u %u %u %u %u\n", c.x, c.y, c.z, c.w, key.x, key.y, o.x, o.y, o.z, o.w);
  }
}
"""


def test_philox_matches_rocrand_open_header(oracle, tmp_path):
    """rocRAND's philox4x32_10_engine::ten_rounds (open source header) is the
    same bijection as the oracle's restatement."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    src = tmp_path / "chk.cpp"
    src.write_text(ROCRAND_CHECK)
    exe = tmp_path / "chk"
    r = subprocess.run([hipcc, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("cannot build the rocRAND host check: " + r.stderr[-300:])
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    n = 0
    for line in out:
        if not line.strip():
            continue
        v = [int(t) for t in line.split()]
        assert oracle.philox(v[0:4], v[4:6]) == v[6:10]
        n += 1
    assert n == 64


def test_uniform_conversion_exact_and_symmetric(oracle):
    for r in (0, 1, 511, 512, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF):
        u = oracle.lib.o_u01f(r)
        k = r >> 9
        assert u == np.float32((k + 0.5) / 8388608.0)
        assert 0.0 < u < 1.0
    # 2u - 1 is symmetric: u(r) and u(~r) mirror
    for r in (0, 12345 << 9, 0xABCDEF00):
        assert np.float32(2 * oracle.lib.o_u01f(r) - 1) == -np.float32(2 * oracle.lib.o_u01f(0xFFFFFFFF ^ r) - 1)


def _random_state(rng, n, L, R=1):
    x = rng.uniform(0.3, L - 0.3, (R, n)).astype(np.float32)
    y = rng.uniform(0.3, L - 0.3, (R, n)).astype(np.float32)
    a = rng.uniform(0, math.pi / 2, (R, n))
    return x, y, np.cos(a).astype(np.float32), np.sin(a).astype(np.float32)


def _e64(x, y, c, s, L):
    xd, yd, cd, sd = (np.ascontiguousarray(v, np.float64) for v in (x, y, c, s))
    h = 1.0 / np.sqrt(cd * cd + sd * sd)
    cd, sd = np.ascontiguousarray(cd * h), np.ascontiguousarray(sd * h)
    from oracle_lib import load
    o = load()
    return o.lib.o_total64_cs(len(x), ptr(xd, f64), ptr(yd, f64), ptr(cd, f64), ptr(sd, f64), L)


def test_model_energy_matches_fp64_and_all_pairs(oracle):
    """Stencil energy (5x5 cells of width >= sqrt(2)/2) finds every pair:
    FP32 canonical energy within 1e-5 relative of the FP64 all-pairs recompute."""
    rng = np.random.default_rng(21)
    for n, L in ((11, 5.0), (50, 8.0), (200, 15.0)):
        for _ in range(4):
            x, y, c, s = (v[0] for v in _random_state(rng, n, L))
            e32 = oracle.energy32(x, y, c, s, L, 0.7075)
            e64 = _e64(x, y, c, s, L)
            assert abs(e32 - e64) <= 1e-5 * max(1.0, abs(e64)), (n, e32, e64)


def test_model_sweep_delta_energy_vs_fp64(oracle):
    """Sum of accepted FP32 energy changes over a sweep == FP64 energy change of
    the configuration (1e-5 relative): every move's neighbours were found and
    no two simultaneous movers interacted."""
    rng = np.random.default_rng(5)
    for n, L in ((20, 5.2), (64, 9.0)):
        x, y, c, s = (np.ascontiguousarray(v[0]) for v in _random_state(rng, n, L))
        for sw in range(6):
            e_before = _e64(x, y, c, s, L)
            e32_before = oracle.energy32(x, y, c, s, L, 0.7075)
            acc, counts = oracle.sweep_model(x, y, c, s, L, 0.7075, 0.15, 0.1, 1.4426950408889634 / 0.05,
                                             7, 0, 3, sw)
            e_after = _e64(x, y, c, s, L)
            e32_after = oracle.energy32(x, y, c, s, L, 0.7075)
            assert counts[0] == n
            assert 0 < acc <= n
            assert abs((e32_after - e32_before) - (e_after - e_before)) <= 1e-5 * max(1.0, e_before)
            assert np.all((x >= 0) & (x <= L) & (y >= 0) & (y <= L))
            assert np.allclose(c * c + s * s, 1.0, atol=2e-6)


def test_model_is_independent_of_group_order(oracle):
    """Members of one colour group cannot interact: the sweep equals the same
    sweep with every square processed strictly one after another in group
    order (the model already is sequential); two identical calls agree."""
    rng = np.random.default_rng(8)
    x, y, c, s = (np.ascontiguousarray(v[0]) for v in _random_state(rng, 40, 7.0))
    a = [v.copy() for v in (x, y, c, s)]
    b = [v.copy() for v in (x, y, c, s)]
    oracle.sweep_model(*a, 7.0, 0.7075, 0.1, 0.1, 50.0, 1, 2, 0, 9)
    oracle.sweep_model(*b, 7.0, 0.7075, 0.1, 0.1, 50.0, 1, 2, 0, 9)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_ptsa_engine_bookkeeping(oracle):
    """CPU PT-SA (the reported CPU baseline): proposal count = R*n*sweeps,
    exchanges attempted every swap_every sweeps on alternating parities,
    shrink only on zero-overlap configurations that it keeps."""
    res = oracle.ptsa(n=11, L0=5.0, replicas=8, tmin=1e-4, tmax=1e-1, sweeps=400, swap_every=10, shrink=1e-4,
                      seed=42, threads=2)
    assert res["proposals"] == 8 * 11 * 400
    # 40 exchange points alternating 4 and 3 adjacent pairs in an 8-ladder
    assert res["exchanges_tried"] == 20 * 4 + 20 * 3
    assert 0 < res["exchanges_accepted"] <= res["exchanges_tried"]
    assert res["L_final"] <= 5.0
    if res["shrinks"]:
        q = res["best_q"]
        assert oracle.total64(q, res["L_feasible"]) == 0.0


def test_ptsa_equal_temperatures_accept_every_exchange(oracle):
    """Forced-exchange invariant (cuda/tests/globaltest.cu:39-46 analogue): with
    equal temperatures the exchange criterion is log u < 0, always true."""
    res = oracle.ptsa(n=6, L0=4.0, replicas=4, tmin=1e-2, tmax=1e-2, sweeps=50, swap_every=5, shrink=0.0,
                      seed=3, threads=1)
    assert res["exchanges_tried"] > 0
    assert res["exchanges_accepted"] == res["exchanges_tried"]
