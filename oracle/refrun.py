"""Run the reference's own program (oracle/_ref/marfa_ref) or the driver over its modules
(oracle/_ref/ref_driver) on a synthetic job in a scratch directory.

TEST INFRASTRUCTURE ONLY: used by tests/golden/gen_golden.py to produce fixtures and by
bench.py's cpu_baseline leg to time the reference on the host.  The reference hard-codes
data/databases/<slug>.<code>, data/Atmospheres/<file>, data/TIPS/TIPS.dat and
output/ptTables/ relative to its working directory (app/main.f90:208,
src/Atmsosphere.f90:31, src/EnvironmentSetup.f90:179,238); the scratch directory
provides exactly that layout.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
from marfa_amd import synth  # noqa: E402

REF_DIR = os.path.join(HERE, "_ref")
MARFA_REF = os.path.join(REF_DIR, "marfa_ref")
REF_DRIVER = os.path.join(REF_DIR, "ref_driver")
REF_PROBE = os.path.join(REF_DIR, "ref_probe")
NT = 20481
SPECIES = {1: ("H2O", "01"), 2: ("CO2", "02"), 3: ("O3", "03"), 5: ("CO", "05"), 6: ("CH4", "06")}


def available() -> bool:
    return all(os.path.exists(p) for p in (MARFA_REF, REF_DRIVER, REF_PROBE))


def scratch(job, root=None) -> str:
    d = tempfile.mkdtemp(prefix="marfa_ref_", dir=root)
    for sub in ("data/databases", "data/Atmospheres", "data/TIPS", "output/ptTables"):
        os.makedirs(os.path.join(d, sub))
    shutil.copy(synth.data_path("TIPS.dat"), os.path.join(d, "data/TIPS/TIPS.dat"))
    code = SPECIES.get(job.mol_code, ("CO2", "02"))[1]
    # sentinel beyond the last window: the reference's LBL_LOOP exits on it
    # (src/LBL.f90:57) instead of reading past the end of the file
    synth.write_db(os.path.join(d, f"data/databases/SYN.{code}"), job.lines,
                   sentinel_nu=job.end_wv + job.cutoff + 11.0)
    synth.write_atm(os.path.join(d, "data/Atmospheres/atm.dat"), job.atm)
    return d


def read_level_records(path, recnums):
    raw = np.fromfile(path, "<f4")
    return np.stack([raw[(r - 1) * NT: r * NT] for r in recnums])


def run(job, use_driver=None, max_levels=None, keep=False, root=None):
    """Returns (records[nlev][nrec][NT] float32, wall seconds of the program, stdout).

    The reference's own `marfa` program is used when the job needs neither a chi factor
    nor a non-default shape (those are compile-time choices of the reference); otherwise
    the driver over the same modules."""
    if use_driver is None:
        use_driver = job.chi_id != 0 or job.shape_id != 0 or max_levels is not None
    d = scratch(job, root)
    tgt = "VAC" if job.target else "ACS"
    species, code = SPECIES.get(job.mol_code, ("CO2", "02"))
    try:
        t0 = time.perf_counter()
        if use_driver:
            os.makedirs(os.path.join(d, "out"))
            args = [REF_DRIVER, repr(float(job.start_wv)), repr(float(job.end_wv)),
                    f"data/databases/SYN.{code}", repr(float(job.cutoff)), tgt,
                    "data/Atmospheres/atm.dat", str(job.chi_id), str(job.mol_code), "out",
                    str(max_levels or len(job.atm)), str(job.shape_id)]
            r = subprocess.run(args, cwd=d, capture_output=True, text=True)
            outdir = os.path.join(d, "out")
        else:
            args = [MARFA_REF, species, repr(float(job.start_wv)), repr(float(job.end_wv)), "SYN",
                    repr(float(job.cutoff)), tgt, "atm.dat"]
            r = subprocess.run(args, cwd=d, capture_output=True, text=True)
            subs = [s for s in os.listdir(os.path.join(d, "output/ptTables")) if s != "latest_run.txt"]
            outdir = os.path.join(d, "output/ptTables", subs[0]) if subs else None
        wall = time.perf_counter() - t0
        if r.returncode != 0 or outdir is None:
            raise RuntimeError(f"reference failed rc={r.returncode}: {r.stdout[-500:]} {r.stderr[-500:]}")
        recnums = []
        a = float(job.start_wv)
        while a < job.end_wv:
            recnums.append(int((a + 1.0) / 10.0))
            a = a + 10.0
        nl = min(len(job.atm), max_levels or len(job.atm))
        recs = []
        for lev in range(1, nl + 1):
            lab = str(lev)
            lab = lab + "_" * (3 - len(lab))
            recs.append(read_level_records(os.path.join(outdir, f"{lab}.ptbin"), recnums))
        return np.stack(recs), wall, r.stdout
    finally:
        if not keep:
            shutil.rmtree(d, ignore_errors=True)


def run_concurrent(jobs, root=None):
    """Run the driver over the reference's modules on every job at once, one process each
    (the reference's levels are independent, README.md:476), to time the reference on
    several host cores.  Returns (per-job LBL seconds of process CPU time, wall seconds from
    the first launch to the last exit)."""
    dirs, procs = [], []
    try:
        cmds = []
        for job in jobs:
            d = scratch(job, root)
            dirs.append(d)
            os.makedirs(os.path.join(d, "out"))
            code = SPECIES.get(job.mol_code, ("CO2", "02"))[1]
            cmds.append(([REF_DRIVER, repr(float(job.start_wv)), repr(float(job.end_wv)), f"data/databases/SYN.{code}",
                          repr(float(job.cutoff)), "VAC" if job.target else "ACS", "data/Atmospheres/atm.dat",
                          str(job.chi_id), str(job.mol_code), "out", str(len(job.atm)), str(job.shape_id)], d))
        t0 = time.perf_counter()
        for args, d in cmds:
            procs.append(subprocess.Popen(args, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
        outs = [p.communicate() for p in procs]
        wall = time.perf_counter() - t0
        secs = []
        for p, (out, err) in zip(procs, outs):
            if p.returncode != 0:
                raise RuntimeError(f"reference driver failed rc={p.returncode}: {out[-300:]} {err[-300:]}")
            secs.append(next(float(l.split()[1]) for l in out.splitlines() if l.startswith("LBL_SECONDS")))
        return secs, wall
    finally:
        for d in dirs:
            shutil.rmtree(d, ignore_errors=True)


def probe(commands: str, cwd=None) -> list[str]:
    """Run oracle/_ref/ref_probe on a command script; returns stdout lines."""
    d = tempfile.mkdtemp(prefix="marfa_probe_")
    try:
        os.makedirs(os.path.join(d, "data/TIPS"))
        shutil.copy(synth.data_path("TIPS.dat"), os.path.join(d, "data/TIPS/TIPS.dat"))
        r = subprocess.run([REF_PROBE], input=commands, capture_output=True, text=True, cwd=cwd or d)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-1000:])
        return r.stdout.splitlines()
    finally:
        shutil.rmtree(d, ignore_errors=True)
