"""Seeded synthetic line lists and atmospheres, and the reference's file formats.

The line-list record is the reference's 36-byte direct-access record
(app/processParFile.f90:140,176-177): f8 nu, f4 S, f4 gamma_air, f4 gamma_self,
f4 E'', f4 n_T, i4 jointMolIso (= isotopologue*100 + molecule), f4 delta.
The atmosphere file follows src/Atmsosphere.f90:29-60 (title line, level count, rows of
height [km], P [atm], T [K], number density [1/(cm^2 km)]).

Line lists are drawn with numpy's PCG64 generator; every list written for the
reference's own program gets a trailing sentinel line beyond the last window so that its
LBL_LOOP (src/LBL.f90:41-57) leaves through its `lineWV >= extEndDeltaWV` exit.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

BOLSI = np.float32(1.3806503e-23)   # src/Constants.f90:14
PATM = np.float32(101325.0)         # src/Constants.f90:16

RECORD_DTYPE = np.dtype([("nu", "<f8"), ("S", "<f4"), ("gair", "<f4"), ("gself", "<f4"),
                         ("elow", "<f4"), ("nT", "<f4"), ("molIso", "<i4"), ("delta", "<f4")])
assert RECORD_DTYPE.itemsize == 36

# nominal molar mass -> (jointMolIso, moleculeIntCode); WISO index = jointMolIso // 100
MOLIISO_FOR_MASS = {18: (101, 1), 44: (1002, 2), 16: (5506, 6), 28: (4605, 5), 48: (2303, 3)}

SEED_BASE = 20261203


@dataclass
class Lines:
    nu: np.ndarray
    S: np.ndarray
    gair: np.ndarray
    gself: np.ndarray
    elow: np.ndarray
    nT: np.ndarray
    delta: np.ndarray
    molIso: np.ndarray

    def __len__(self):
        return int(self.nu.shape[0])

    def sorted(self) -> "Lines":
        o = np.argsort(self.nu, kind="stable")
        return Lines(*(getattr(self, f)[o] for f in FIELDS))

    def take(self, idx) -> "Lines":
        return Lines(*(getattr(self, f)[idx] for f in FIELDS))

    def to_records(self) -> np.ndarray:
        rec = np.empty(len(self), RECORD_DTYPE)
        for f in FIELDS:
            rec[f] = getattr(self, f)
        return rec


FIELDS = ("nu", "S", "gair", "gself", "elow", "nT", "delta", "molIso")


@dataclass
class Atmosphere:
    P: np.ndarray      # [atm] float32
    T: np.ndarray      # [K] float32
    rho: np.ndarray    # [1/(cm^2 km)] float32
    h: np.ndarray = field(default=None)

    def __len__(self):
        return int(self.P.shape[0])


def draw_params(rng: np.random.Generator, n: int, s_log=(-30.0, -20.0), e_max=5000.0,
                molIso=1002) -> dict:
    """Per-line parameter laws shared by every configuration."""
    return dict(
        S=(10.0 ** rng.uniform(s_log[0], s_log[1], n)).astype(np.float32),
        gair=rng.uniform(0.05, 0.10, n).astype(np.float32),
        gself=rng.uniform(0.07, 0.12, n).astype(np.float32),
        nT=rng.uniform(0.5, 0.8, n).astype(np.float32),
        delta=rng.uniform(-0.01, 0.0, n).astype(np.float32),
        elow=rng.uniform(0.0, e_max, n).astype(np.float32),
        molIso=np.full(n, molIso, np.int32),
    )


def make_lines(nu: np.ndarray, rng: np.random.Generator, **kw) -> Lines:
    nu = np.sort(np.asarray(nu, np.float64))
    p = draw_params(rng, nu.shape[0], **kw)
    return Lines(nu=nu, **p)


def uniform_lines(n, lo, hi, seed, **kw) -> Lines:
    rng = np.random.Generator(np.random.PCG64(seed))
    nu = rng.uniform(lo, hi, n)
    return make_lines(nu, rng, **kw)


def banded_lines(n, lo, hi, nbands, width, seed, **kw) -> Lines:
    """Line density proportional to a sum of `nbands` equal Gaussian bands of std
    `width`, centres uniform in [lo, hi], truncated to [lo, hi]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(lo, hi, nbands)
    out = np.empty(0)
    while out.shape[0] < n:
        m = n - out.shape[0]
        c = centres[rng.integers(0, nbands, 2 * m)]
        x = c + width * rng.standard_normal(2 * m)
        x = x[(x >= lo) & (x <= hi)]
        out = np.concatenate([out, x[:m]])
    return make_lines(out, rng, **kw)


def vmr_density(P, T, vmr):
    """Number density [1/(cm^2 km)] whose pSelf (app/main.f90:92) equals vmr*P."""
    P = np.asarray(P, np.float64)
    T = np.asarray(T, np.float64)
    return (vmr * P * float(PATM) / (float(BOLSI) * T) / 10.0).astype(np.float32)


def log_levels(n, p0, p1, t0, t1, vmr) -> Atmosphere:
    i = np.arange(n, dtype=np.float64)
    den = max(n - 1, 1)
    P = (p0 * (p1 / p0) ** (i / den)).astype(np.float32)
    T = (t0 + (t1 - t0) * i / den).astype(np.float32)
    return Atmosphere(P=P, T=T, rho=vmr_density(P, T, vmr), h=i.astype(np.float32))


def p_self(atm: Atmosphere) -> np.ndarray:
    """pSelf exactly as app/main.f90:92 forms it (FP32, left to right)."""
    return (atm.rho * np.float32(10.0) * BOLSI * atm.T / PATM).astype(np.float32)


# --------------------------------------------------------------------------- configs
@dataclass
class Job:
    lines: Lines
    atm: Atmosphere
    start_wv: float
    end_wv: float
    cutoff: float
    mol_code: int = 2
    chi_id: int = 0
    shape_id: int = 0
    target: int = 0   # 0 ACS, 1 VAC

    @property
    def nrec(self) -> int:
        n, a = 0, float(self.start_wv)
        while a < self.end_wv:
            n += 1
            a = a + 10.0
        return n


def config1(seed=SEED_BASE + 1) -> Job:
    """20,000 CO2-like lines in 4000-4100 cm-1, 10 levels, cutoff 25, Voigt, no chi."""
    lines = uniform_lines(20000, 4000.0, 4100.0, seed)
    atm = log_levels(10, 90.0, 1e-3, 735.0, 170.0, 0.965)
    return Job(lines, atm, 4000.0, 4100.0, 25.0, mol_code=2, chi_id=0)


def config2(seed=SEED_BASE + 2, nlines=2_000_000, nlev=60) -> Job:
    """2,000,000 lines over 500-10,000 cm-1 in 20 Gaussian bands of width 50 cm-1,
    60 levels (P 90 -> 1e-4 atm log, T 735 -> 170 K), cutoff 125, Pollack chi factor."""
    lines = banded_lines(nlines, 500.0, 10000.0, 20, 50.0, seed)
    atm = log_levels(nlev, 90.0, 1e-4, 735.0, 170.0, 0.965)
    return Job(lines, atm, 500.0, 10000.0, 125.0, mol_code=2, chi_id=2)


def config4(seed=SEED_BASE + 4, nlines=200_000) -> Job:
    """Adjoint case: 200,000 lines in 1000-1100 cm-1, 30 levels (P 10 -> 1e-3, T 300 -> 180),
    cutoff 25."""
    lines = uniform_lines(nlines, 1000.0, 1100.0, seed)
    atm = log_levels(30, 10.0, 1e-3, 300.0, 180.0, 0.965)
    return Job(lines, atm, 1000.0, 1100.0, 25.0, mol_code=2, chi_id=0)


def config5(seed=SEED_BASE + 5, nlines=8_000_000, nlev=100) -> Job:
    """8,000,000 H2O-like lines over 10-20,000 cm-1, 100 levels (P 100 -> 1e-5 atm,
    T 750 -> 160 K), cutoff 250, no chi.  VMR is not fixed by the configuration; 0.01 is
    used."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nu = rng.uniform(10.0, 20000.0, nlines)
    lines = make_lines(nu, rng, s_log=(-32.0, -18.0), e_max=8000.0, molIso=101)
    atm = log_levels(nlev, 100.0, 1e-5, 750.0, 160.0, 0.01)
    return Job(lines, atm, 10.0, 20000.0, 250.0, mol_code=1, chi_id=0)


def config2_window(nlines=400_000, nlev=6, width=100.0, seed=SEED_BASE + 2) -> Job:
    """Configuration 2's laws at a reduced size: `nlines` banded lines over 500-10,000 cm-1,
    `nlev` levels from 90 to 1e-4 atm (735 -> 170 K), cutoff 125, Pollack chi, and the
    range cut to the `width` cm-1 (on the 10 cm-1 grid) that holds the most lines."""
    job = config2(seed=seed, nlines=nlines, nlev=nlev)
    h, _ = np.histogram(job.lines.nu, bins=np.arange(500.0, 10000.0 + 1e-9, 10.0))
    k = int(np.argmax(np.convolve(h, np.ones(int(width // 10)), "valid")))
    job.start_wv = 500.0 + 10.0 * k
    job.end_wv = job.start_wv + width
    return job


def config5_window(density=40.0, nlev=5, start=10000.0, width=100.0, seed=SEED_BASE + 5) -> Job:
    """Configuration 5's laws (H2O-161, S 10^U(-32,-18), E'' U(0, 8000), cutoff 250, no chi,
    P 100 -> 1e-5 atm, T 750 -> 160 K, VMR 0.01) on the range [start, start + width) with
    `density` lines per cm-1 over the range's reach (+-cutoff)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = start - 255.0, start + width + 255.0
    nu = rng.uniform(lo, hi, int(density * (hi - lo)))
    lines = make_lines(nu, rng, s_log=(-32.0, -18.0), e_max=8000.0, molIso=101)
    atm = log_levels(nlev, 100.0, 1e-5, 750.0, 160.0, 0.01)
    return Job(lines, atm, start, start + width, 250.0, mol_code=1, chi_id=0)


def config3_jobs(seed=SEED_BASE + 3, njobs=256, max_lines=1_000_000) -> list:
    """256 independent jobs: N ~ 10^U(3,6) lines, 10-100 levels, windows of 20-500 cm-1 at
    random positions in 10-20,000 cm-1, molar mass from {18,44,16,28,48}, cutoff from
    {25,125,250}.  Levels resample config 2's P/T law (90 -> 1e-4 atm, 735 -> 170 K)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    jobs = []
    for j in range(njobs):
        n = int(min(round(10 ** rng.uniform(3, 6)), max_lines))
        nlev = int(rng.integers(10, 101))
        w = 10.0 * round(rng.uniform(20, 500) / 10.0)
        a = 10.0 * np.floor(rng.uniform(10, 20000 - w) / 10.0)
        mass = int(rng.choice([18, 44, 16, 28, 48]))
        cutoff = float(rng.choice([25.0, 125.0, 250.0]))
        molIso, mol = MOLIISO_FOR_MASS[mass]
        sub = np.random.Generator(np.random.PCG64(seed * 1000 + j))
        nu = sub.uniform(max(a - cutoff, 1e-3), a + w + cutoff, n)
        lines = make_lines(nu, sub, molIso=molIso)
        atm = log_levels(nlev, 90.0, 1e-4, 735.0, 170.0, 0.965)
        jobs.append(Job(lines, atm, a, a + w, cutoff, mol_code=mol,
                        chi_id=2 if mol == 2 else 0))
    return jobs


# --------------------------------------------------------------------------- files
def write_db(path: str, lines: Lines, sentinel_nu: float | None = None) -> None:
    rec = lines.to_records()
    if sentinel_nu is not None:
        s = np.zeros(1, RECORD_DTYPE)
        s["nu"] = sentinel_nu
        s["molIso"] = lines.molIso[-1] if len(lines) else 1002
        s["gair"] = 0.07
        s["gself"] = 0.1
        s["nT"] = 0.7
        rec = np.concatenate([rec, s])
    rec.tofile(path)


def read_db(path: str) -> Lines:
    rec = np.fromfile(path, RECORD_DTYPE)
    return Lines(*(np.ascontiguousarray(rec[f]) for f in FIELDS))


def _f32s(x) -> str:
    """Shortest decimal that reads back as the same FP32 value."""
    return np.format_float_scientific(np.float32(x), unique=True)


def write_atm(path: str, atm: Atmosphere, title="synthetic atmosphere") -> None:
    with open(path, "w") as f:
        f.write(f"{title[:20]:<20}\n{len(atm)}\n")
        h = atm.h if atm.h is not None else np.arange(len(atm), dtype=np.float32)
        for k in range(len(atm)):
            f.write(f"{_f32s(h[k])} {_f32s(atm.P[k])} {_f32s(atm.T[k])} {_f32s(atm.rho[k])}\n")


def read_atm(path: str) -> Atmosphere:
    with open(path) as f:
        f.readline()
        n = int(f.readline().split()[0])
        rows = [f.readline().split() for _ in range(n)]
    arr = np.array([[np.float32(v) for v in r[:4]] for r in rows], np.float32)
    return Atmosphere(P=arr[:, 1].copy(), T=arr[:, 2].copy(), rho=arr[:, 3].copy(), h=arr[:, 0].copy())


def data_path(name: str) -> str:
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", name)


def load_tips(path: str | None = None) -> np.ndarray:
    """TIPS table (74 isotopologues x 501 temperatures from 20 K in 2 K steps) as read by
    src/EnvironmentSetup.f90:169-184: list-directed FP32 values."""
    path = path or data_path("TIPS.dat")
    with open(path) as f:
        head = f.readline().replace(",", " ").split()
        niso, nt = int(head[0]), int(head[1])
        toks = f.read().replace(",", " ").split()
    out = np.array(toks, dtype=np.float32)
    assert out.shape[0] == niso * nt
    return out.reshape(niso, nt)
