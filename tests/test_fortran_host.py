"""The Fortran host program (marfa_amd/fortran/marfa_gpu), drop-in for app/main.f90.

CPU tests: command-line and input errors end the program with the reference's stop codes
(src/EnvironmentSetup.f90:43-54 stop 1, :86-89 stop 2, app/main.f90:70-73 stop 8,
src/Atmsosphere.f90:37-56 stop 6/7), and without a HIP device it fails loudly with the
library's MARFA_E_NODEVICE (23) instead of computing anything on the CPU.
GPU test: the .ptbin files it writes hold, record for record, exactly what the C ABI
returns for the same job through marfa_pt_tables, and agree with the FP64 oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from marfa_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EXE = os.path.join(ROOT, "marfa_amd", "fortran", "marfa_gpu")
NT = 20481

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built (amdflang)")


def small_job(seed=31, n=3000, nlev=3, chi=0, target=0):
    lines = synth.uniform_lines(n, 3970.0, 4060.0, seed)
    atm = synth.log_levels(nlev, 50.0, 1e-3, 500.0, 190.0, 0.965)
    return synth.Job(lines, atm, 4000.0, 4030.0, 25.0, mol_code=2, chi_id=chi, target=target)


def workdir(tmp_path, job, code="02"):
    for sub in ("data/databases", "data/Atmospheres", "data/TIPS"):
        os.makedirs(tmp_path / sub, exist_ok=True)
    with open(synth.data_path("TIPS.dat"), "rb") as f:
        (tmp_path / "data/TIPS/TIPS.dat").write_bytes(f.read())
    synth.write_db(str(tmp_path / f"data/databases/SYN.{code}"), job.lines)
    synth.write_atm(str(tmp_path / "data/Atmospheres/atm.dat"), job.atm)
    return tmp_path


def run(tmp_path, *args, timeout=300):
    return subprocess.run([EXE, *map(str, args)], cwd=tmp_path, capture_output=True, text=True,
                          timeout=timeout)


def has_gpu_device() -> bool:
    return os.path.exists("/dev/kfd")


@needs_exe
def test_usage_stop_1(tmp_path):
    r = run(tmp_path, "CO2", "4000", "4030")
    assert r.returncode == 1, r.stdout + r.stderr
    assert "Insufficient number of arguments" in r.stdout


@needs_exe
def test_bad_target_stop_2(tmp_path):
    workdir(tmp_path, small_job())
    r = run(tmp_path, "CO2", "4000", "4030", "SYN", "25", "XYZ", "atm.dat")
    assert r.returncode == 2, r.stdout + r.stderr
    assert "Invalid targetValue" in r.stdout


@needs_exe
def test_bad_molecule_stop_8(tmp_path):
    workdir(tmp_path, small_job())
    r = run(tmp_path, "XX9", "4000", "4030", "SYN", "25", "ACS", "atm.dat")
    assert r.returncode == 8, r.stdout + r.stderr
    assert "unsupported molecule" in r.stdout


@needs_exe
def test_missing_atmosphere_stop_7(tmp_path):
    workdir(tmp_path, small_job())
    r = run(tmp_path, "CO2", "4000", "4030", "SYN", "25", "ACS", "nope.dat")
    assert r.returncode == 7, r.stdout + r.stderr


@needs_exe
def test_too_many_levels_stop_6(tmp_path):
    job = small_job()
    workdir(tmp_path, job)
    txt = (tmp_path / "data/Atmospheres/atm.dat").read_text().splitlines()
    txt[1] = "201"
    (tmp_path / "data/Atmospheres/atm.dat").write_text("\n".join(txt) + "\n")
    r = run(tmp_path, "CO2", "4000", "4030", "SYN", "25", "ACS", "atm.dat")
    assert r.returncode == 6, r.stdout + r.stderr


@needs_exe
@pytest.mark.skipif(has_gpu_device(), reason="checks the no-device error path")
def test_no_device_fails_loudly(tmp_path):
    workdir(tmp_path, small_job())
    r = run(tmp_path, "CO2", "4000", "4030", "SYN", "25", "ACS", "atm.dat")
    assert r.returncode == 23, r.stdout + r.stderr
    assert not any(p.suffix == ".ptbin" for p in tmp_path.rglob("*.ptbin"))


def read_output(tmp_path, nlev, recnums):
    base = tmp_path / "output/ptTables"
    sub = (base / "latest_run.txt").read_text().strip()
    out = []
    for lev in range(1, nlev + 1):
        lab = str(lev) + "_" * (3 - len(str(lev)))
        raw = np.fromfile(base / sub / f"{lab}.ptbin", "<f4")
        assert raw.size == max(recnums) * NT
        out.append(np.stack([raw[(r - 1) * NT: r * NT] for r in recnums]))
    return np.stack(out)


@needs_exe
@pytest.mark.gpu
@pytest.mark.parametrize("target,chi,prec", [(0, 0, "FP64"), (1, 2, "FP32K")])
def test_host_matches_c_abi_and_oracle(tmp_path, target, chi, prec):
    from marfa_amd import lbl
    job = small_job(chi=chi, target=target)
    workdir(tmp_path, job)
    args = ["CO2", "4000", "4030", "SYN", "25", ["ACS", "VAC"][target], "atm.dat", "VOIGT",
            ["NONE", "TONKOV", "POLLACK", "PERRIN"][chi], prec]
    print("fortran host", args, flush=True)
    r = run(tmp_path, *args)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Congratulations" in r.stdout
    recnums = [int((a + 1.0) / 10.0) for a in (4000.0, 4010.0, 4020.0)]
    host = read_output(tmp_path, len(job.atm), recnums)

    tips = synth.load_tips()
    ctx = lbl.LBLContext(0)
    try:
        ctx.load_job(job, tips)
        spec = lbl.JobSpec.from_job(job, lbl.FP64 if prec == "FP64" else lbl.FP32K)
        ref = ctx.pt_tables(spec)
    finally:
        ctx.close()
    assert np.array_equal(host, ref.reshape(host.shape))
    # the Python writer (marfa_amd/pttable.py) produces the same files from the same records
    from marfa_amd import pttable
    base = tmp_path / "output/ptTables"
    sub = base / (base / "latest_run.txt").read_text().strip()
    for p in pttable.write_pt_tables(tmp_path / "py", ref.reshape(host.shape), recnums):
        assert p.read_bytes() == (sub / p.name).read_bytes(), p.name

    oj = po.OracleJob(job, tips)
    o, _ = oj.pt_tables(po.F64)
    scale = np.abs(o).max(axis=-1, keepdims=True)
    err = np.abs(host - o) / np.maximum(np.abs(o), 1e-30 * scale)
    tol = 1e-6 if prec == "FP64" else 1e-4
    assert err.max() < tol, err.max()
