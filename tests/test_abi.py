"""The C-ABI library and the `packing` CLI without a GPU: the library loads,
exports every entry point include/sis_abi.h declares, validates input with the
reference's error codes, and fails loudly (SIS_EDEVICE) when no device exists."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from conftest import PKG, ROOT

HEADER = ROOT / "include" / "sis_abi.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sis_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(sis):
    names = declared_functions()
    assert len(names) >= 15
    lib = ctypes.CDLL(str(PKG / "lib" / "libsisamd.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_layouts_match_header(sis, tmp_path):
    """The Python mirror's structs have the header's size and field offsets
    (compiled with the host C compiler)."""
    structs = {"sis_instance": sis.SisInstance, "sis_params": sis.SisParams, "sis_result": sis.SisResult,
               "sis_server_cfg": sis.SisServerCfg, "sis_tuning": sis.SisTuning}
    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            cf = "lambda" if fname == "lambda_" else fname
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {cf}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for ln in filter(None, out):
        cname, field, val = ln.split()
        py = structs[cname]
        got = ctypes.sizeof(py) if field == "size" else getattr(py, field).offset
        assert got == int(val), (cname, field, got, val)


def test_record_dtypes_match_header(sis, tmp_path):
    """The numpy records the probes exchange (replica, temperature slot) and
    the exchange-step descriptor have the header's size and field offsets,
    and the oracle's mirrors have the same layout."""
    import oracle_lib
    recs = {"sis_rep_state": sis.REP_DTYPE, "sis_slot_state": sis.SLOT_DTYPE}
    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, dt in recs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname in dt.names:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append('printf("sis_ladder_desc size %zu\\n", sizeof(sis_ladder_desc));')
    for fname, _ in sis.SisLadderDesc._fields_:
        lines.append(f'printf("sis_ladder_desc {fname} %zu\\n", offsetof(sis_ladder_desc, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "records.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "records"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for ln in filter(None, out):
        cname, field, val = ln.split()
        if cname == "sis_ladder_desc":
            py = sis.SisLadderDesc
            got = ctypes.sizeof(py) if field == "size" else getattr(py, field).offset
        else:
            dt = recs[cname]
            got = dt.itemsize if field == "size" else dt.fields[field][1]
        assert got == int(val), (cname, field, got, val)
        seen += 1
    assert seen > 30
    assert oracle_lib.REP_DTYPE == sis.REP_DTYPE and oracle_lib.SLOT_DTYPE == sis.SLOT_DTYPE


def test_default_params(sis):
    p = sis.default_params()
    assert p.tmin == 1e-5 and p.tmax == 5e-2
    assert p.move_sigma == 0.05 and p.rot_sigma == 0.05 and p.target_accept == 0.3
    assert p.shrink == 1e-4 and p.sweeps_per_exchange == 10 and p.polish_iters == 200
    assert p.improve == 0 and p.cool_period == 0 and p.certify_every == 1
    assert abs(p.cell - 0.7075) < 1e-12


@pytest.mark.parametrize("n,L,kw", [
    (0, 5.0, {}), (1001, 40.0, {}), (5, 0.0, {}), (5, 65.0, {}), (5, float("nan"), {}),
    (5, 5.0, {"tmin": 0.0}), (5, 5.0, {"tmin": 1e-2, "tmax": 1e-3}), (5, 5.0, {"replicas": 0}),
    (5, 5.0, {"replicas": 12, "ladder": 5}), (5, 5.0, {"cell": 0.5}), (5, 5.0, {"sweeps_per_exchange": 0}),
])
def test_invalid_input_is_einval(sis, n, L, kw):
    with pytest.raises(sis.SisError) as e:
        sis.solve(n, L, None, **kw)
    assert e.value.code == sis.SIS_EINVAL


def test_nonfinite_seed_is_einval(sis):
    q = [[0.5, 0.5, 0.0], [float("nan"), 1.5, 0.0]]
    with pytest.raises(sis.SisError) as e:
        sis.solve(2, 3.0, q)
    assert e.value.code == sis.SIS_EINVAL


def _has_device(sis) -> bool:
    return sis.lib().sis_init(0) == 0


def test_no_device_fails_loudly(sis):
    if _has_device(sis):
        pytest.skip("a HIP device is present; covered by the GPU tests")
    with pytest.raises(sis.SisError) as e:
        sis.solve(5, 3.0, None, sweeps=10, replicas=8)
    assert e.value.code == sis.SIS_EDEVICE


# -- the `packing` CLI keeps the reference's process contract -----------------
PACKING = PKG / "bin" / "packing"


def run_packing(args, env=None, cwd=None):
    e = dict(os.environ)
    for k in list(e):
        if k.startswith("PACK_"):
            del e[k]
    if env:
        e.update(env)
    return subprocess.run([str(PACKING), *args], capture_output=True, text=True, env=e, cwd=cwd, timeout=120)


def test_cli_usage_exit_2():
    r = run_packing(["only", "three", "args"])
    assert r.returncode == 2 and "packing seed.txt" in r.stderr


@pytest.mark.parametrize("args", [
    ["-1", "0", "0"], ["100001", "0", "0"], ["0", "4097", "0"], ["0", "0", "12"], ["0", "0", "4104"],
])
def test_cli_argument_ranges(tmp_path, args):
    seed = tmp_path / "seed.txt"
    seed.write_text("2 3\n0.5 0.5 0\n1.5 0.5 0\n")
    r = run_packing([str(seed), *args, "42", str(tmp_path / "o.json")])
    assert r.returncode == 2


@pytest.mark.parametrize("text", ["", "0 3\n", "1001 3\n", "2 0\n", "2 65\n", "2 3\n0.5 0.5 0\n1.5 0.5\n",
                                  "2 3\n0.5 nan 0\n1.5 0.5 0\n"])
def test_cli_seed_file_validation(tmp_path, text):
    seed = tmp_path / "seed.txt"
    seed.write_text(text)
    r = run_packing([str(seed), "0", "0", "0", "42", str(tmp_path / "o.json")])
    assert r.returncode == 2


@pytest.mark.parametrize("name,value", [("PACK_TMIN", "0"), ("PACK_TMAX", "2"), ("PACK_MOVE", "0.5"),
                                        ("PACK_ROTATION", "abc"), ("PACK_SHRINK", "1"),
                                        ("PACK_AUTOTUNE", "3"), ("PACK_MAX_CHAINS", "4"),
                                        ("PACK_SWEEP_STEPS", "2")])
def test_cli_env_knobs_range_checked(tmp_path, name, value):
    seed = tmp_path / "seed.txt"
    seed.write_text("2 3\n0.5 0.5 0\n1.5 0.5 0\n")
    r = run_packing([str(seed), "0", "0", "0", "42", str(tmp_path / "o.json")], env={name: value})
    assert r.returncode == 2
    assert f"invalid {name}" in r.stderr


def test_cli_tmax_below_tmin(tmp_path):
    seed = tmp_path / "seed.txt"
    seed.write_text("2 3\n0.5 0.5 0\n1.5 0.5 0\n")
    r = run_packing([str(seed), "0", "0", "0", "42", str(tmp_path / "o.json")],
                    env={"PACK_TMIN": "1e-3", "PACK_TMAX": "1e-4"})
    assert r.returncode == 2


# -- the `seeded` CLI keeps cuda/seeded.cu:832-969 ------------------------------
SEEDED = PKG / "bin" / "seeded"


def run_seeded(args, env=None):
    e = {k: v for k, v in os.environ.items() if not k.startswith("SEED_")}
    if env:
        e.update(env)
    return subprocess.run([str(SEEDED), *args], capture_output=True, text=True, env=e, timeout=120)


def test_seeded_usage_exit_2():
    r = run_seeded(["x"])
    assert r.returncode == 2 and "seeded seed.txt" in r.stderr


@pytest.mark.parametrize("name,value", [("SEED_SHRINK", "0"), ("SEED_TMAX", "0.5"), ("SEED_MOVE", "1"),
                                        ("SEED_COLLECTIVE", "2"), ("SEED_POLISH", "5000"),
                                        ("SEED_STRUCTURED", "-1"), ("SEED_AUTOTUNE", "x"),
                                        ("SEED_MAX_CHAINS", "7"), ("SEED_SWEEP_STEPS", "-1")])
def test_seeded_env_knobs_range_checked(tmp_path, name, value):
    seed = tmp_path / "seed.txt"
    seed.write_text("2 3\n0.5 0.5 0\n1.5 0.5 0\n")
    r = run_seeded([str(seed), "0", "0", "0", "42", str(tmp_path / "o.json")], env={name: value})
    assert r.returncode == 2
    assert f"invalid {name}" in r.stderr


def test_seeded_argument_and_seed_validation(tmp_path):
    seed = tmp_path / "seed.txt"
    seed.write_text("2 3\n0.5 0.5 0\n1.5 0.5 0\n")
    assert run_seeded([str(seed), "0", "0", "12", "42", str(tmp_path / "o.json")]).returncode == 2
    seed.write_text("2 3\n0.5 0.5 0\n")
    assert run_seeded([str(seed), "0", "0", "0", "42", str(tmp_path / "o.json")]).returncode == 2


def test_server_without_device_fails_loudly(sis):
    if _has_device(sis):
        pytest.skip("a HIP device is present; covered by the GPU tests")
    with pytest.raises(sis.SisError) as e:
        sis.Server(max_instances=4)
    assert e.value.code == sis.SIS_EDEVICE


def test_handle_constructors_report_the_failure_code(sis):
    """Entry points that return a handle give NULL on failure and the code
    through sis_last_error_code: invalid arguments are SIS_EINVAL (checked
    before any device work), not a device error."""
    with pytest.raises(sis.SisError) as e:
        sis.Server(max_instances=0)
    assert e.value.code == sis.SIS_EINVAL
    with pytest.raises(sis.SisError) as e:
        sis.Server(intervals_per_round=0)
    assert e.value.code == sis.SIS_EINVAL
    for bad in (dict(request_timeout_s=-1.0), dict(request_timeout_s=float("nan")), dict(hbm_budget_bytes=-1.0)):
        with pytest.raises(sis.SisError) as e:
            sis.Server(**bad)
        assert e.value.code == sis.SIS_EINVAL, bad
    p = sis.default_params(replicas=8)
    with pytest.raises(sis.SisError) as e:
        sis.BatchSession([(0, 5.0, None)], p)
    assert e.value.code == sis.SIS_EINVAL
    with pytest.raises(sis.SisError) as e:
        sis.BatchSession([(10, 70.0, None)], p)
    assert e.value.code == sis.SIS_EINVAL

