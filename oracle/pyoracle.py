"""ctypes binding of the CPU oracle (oracle/marfa_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker
and by bench.py's cpu_baseline leg -- never by the product path in marfa_amd/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmarfa_oracle.so")
NT = 20481
NGRID = 30690
R32, F64 = 0, 1


class _Lines(C.Structure):
    _fields_ = [("n", C.c_int64), ("nu", C.c_void_p), ("S", C.c_void_p), ("gair", C.c_void_p),
                ("gself", C.c_void_p), ("elow", C.c_void_p), ("nT", C.c_void_p),
                ("delta", C.c_void_p), ("molIso", C.c_void_p)]


class _Job(C.Structure):
    _fields_ = [("start_wv", C.c_double), ("end_wv", C.c_double), ("cutoff", C.c_float),
                ("shape_id", C.c_int), ("chi_id", C.c_int), ("mol_code", C.c_int),
                ("target", C.c_int)]


class Counts(C.Structure):
    _fields_ = [("lines_left", C.c_int64), ("lines_center", C.c_int64),
                ("lines_right", C.c_int64), ("lines_skipped", C.c_int64),
                ("samples", C.c_int64)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
        L.mo_voigt_r32.restype = C.c_float
        L.mo_voigt_r32.argtypes = [C.c_double, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                   C.c_int, C.c_float]
        L.mo_voigt_f64.restype = C.c_double
        L.mo_voigt_f64.argtypes = [C.c_double] * 4 + [C.c_int] * 3 + [C.c_double]
        L.mo_chi_r32.restype = C.c_float
        L.mo_chi_r32.argtypes = [C.c_int, C.c_double, C.c_int, C.c_float]
        L.mo_chi_f64.restype = C.c_double
        L.mo_chi_f64.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double]
        spec_args = [C.c_void_p, C.c_double] + [C.c_float] * 5 + [C.c_int32] + [C.c_float] * 4 + [dp]
        L.mo_spec_r32.argtypes = spec_args
        L.mo_spec_f64.argtypes = spec_args
        L.mo_cascade_r32.argtypes = [fp, fp]
        L.mo_cascade_f64.argtypes = [dp, dp]
        L.mo_wiso.restype = C.c_float
        L.mo_wiso.argtypes = [C.c_int]
        L.mo_constants.argtypes = [C.POINTER(C.c_uint32)]
        L.mo_windows.restype = C.c_int
        L.mo_windows.argtypes = [C.POINTER(_Lines), C.POINTER(_Job), C.c_int, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p]
        L.mo_record.restype = C.c_int
        L.mo_record.argtypes = [C.c_int, C.POINTER(_Lines), C.POINTER(_Job), C.c_void_p,
                                C.c_float, C.c_float, C.c_float, C.c_double, C.c_int64,
                                C.c_int64, dp, C.c_void_p, C.POINTER(Counts)]
        L.mo_pt_tables.restype = C.c_int
        L.mo_pt_tables.argtypes = [C.c_int, C.POINTER(_Lines), C.POINTER(_Job), C.c_void_p,
                                   C.c_int, fp, fp, fp, C.c_int, dp, C.POINTER(Counts)]
        L.mo_direct.restype = C.c_int
        L.mo_direct.argtypes = [C.POINTER(_Lines), C.POINTER(_Job), C.c_void_p, C.c_float,
                                C.c_float, C.c_float, C.c_double, C.c_int64, C.c_int64,
                                C.c_int, C.c_int, dp]
        L.mo_trace_begin.argtypes = []
        L.mo_trace_end.argtypes = [C.c_void_p]
        _lib = L
    return _lib


class Trace:
    """Decision trace of the F64 path (mo_trace_begin / mo_trace_end): `with Trace() as t:`
    around mo_record / mo_cascade_f64 calls; afterwards t.hashes = (selection, regions,
    limiter).  Equal hashes: the same branches were taken."""

    def __enter__(self):
        lib().mo_trace_begin()
        return self

    def __exit__(self, *exc):
        out = np.zeros(3, np.uint64)
        lib().mo_trace_end(_ptr(out))
        self.hashes = tuple(int(x) for x in out)
        return False


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleJob:
    """Holds contiguous arrays for one job and the ctypes views into them."""

    def __init__(self, job, tips):
        ln = job.lines
        self._keep = [np.ascontiguousarray(ln.nu, np.float64)]
        for f in ("S", "gair", "gself", "elow", "nT", "delta"):
            self._keep.append(np.ascontiguousarray(getattr(ln, f), np.float32))
        self._keep.append(np.ascontiguousarray(ln.molIso, np.int32))
        self.lines = _Lines(len(ln), *[a.ctypes.data for a in self._keep])
        self.job = _Job(float(job.start_wv), float(job.end_wv), float(job.cutoff), int(job.shape_id),
                        int(job.chi_id), int(job.mol_code), int(job.target))
        self.tips = np.ascontiguousarray(tips, np.float32)
        self.P = np.ascontiguousarray(job.atm.P, np.float32)
        self.T = np.ascontiguousarray(job.atm.T, np.float32)
        self.rho = np.ascontiguousarray(job.atm.rho, np.float32)
        self.nlev = len(job.atm)
        self.windows()

    def windows(self):
        cap = 1
        while True:
            cap *= 2
            lo = np.zeros(cap, np.int64)
            hi = np.zeros(cap, np.int64)
            rec = np.zeros(cap, np.int32)
            org = np.zeros(cap, np.float64)
            n = lib().mo_windows(C.byref(self.lines), C.byref(self.job), cap, _ptr(lo), _ptr(hi),
                                 _ptr(rec), _ptr(org))
            if n >= 0:
                break
        self.lo, self.hi, self.recnum, self.origin = lo[:n], hi[:n], rec[:n], org[:n]
        self.nrec = n
        return self.lo, self.hi, self.recnum, self.origin

    def record(self, precision, lev, r, grids=False):
        rk = np.zeros(NT, np.float64)
        g = np.zeros(NGRID + NT, np.float64) if grids else None
        cnt = Counts()
        lib().mo_record(precision, C.byref(self.lines), C.byref(self.job), _ptr(self.tips),
                        float(self.P[lev]), float(self.T[lev]), float(self.rho[lev]),
                        float(self.origin[r]), int(self.lo[r]), int(self.hi[r]),
                        rk.ctypes.data_as(C.POINTER(C.c_double)),
                        _ptr(g) if grids else None, C.byref(cnt))
        return (rk, g, cnt) if grids else (rk, cnt)

    def pt_tables(self, precision):
        out = np.zeros((self.nlev, self.nrec, NT), np.float64)
        cnt = Counts()
        rc = lib().mo_pt_tables(precision, C.byref(self.lines), C.byref(self.job), _ptr(self.tips),
                                self.nlev, self.P.ctypes.data_as(C.POINTER(C.c_float)),
                                self.T.ctypes.data_as(C.POINTER(C.c_float)),
                                self.rho.ctypes.data_as(C.POINTER(C.c_float)), self.nrec,
                                out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(cnt))
        assert rc == 0
        return out, cnt

    def line_grids(self, lev, r, i, precision=F64):
        """Pre-cascade grids (NGRID + NT) of record r at level lev holding only line i's
        deposits: the nodes line i samples and its value at each."""
        g = np.zeros(NGRID + NT, np.float64)
        rk = np.zeros(NT, np.float64)
        lib().mo_record(precision, C.byref(self.lines), C.byref(self.job), _ptr(self.tips),
                        float(self.P[lev]), float(self.T[lev]), float(self.rho[lev]),
                        float(self.origin[r]), int(i), int(i) + 1,
                        rk.ctypes.data_as(C.POINTER(C.c_double)), _ptr(g), None)
        return g

    def with_param(self, field, i, value) -> "OracleJob":
        """A copy with one line parameter (S, gair, nT, delta: field name of synth.Lines)
        replaced: line i's value."""
        import copy
        o = copy.copy(self)
        names = ("nu", "S", "gair", "gself", "elow", "nT", "delta", "molIso")
        k = names.index(field)
        o._keep = list(self._keep)
        o._keep[k] = self._keep[k].copy()
        o._keep[k][i] = value
        o.lines = _Lines(len(self._keep[0]), *[a.ctypes.data for a in o._keep])
        return o

    def with_level(self, lev, T=None, P=None) -> "OracleJob":
        """A copy with level lev's temperature and/or pressure replaced (density kept)."""
        import copy
        o = copy.copy(self)
        o.T, o.P = self.T.copy(), self.P.copy()
        if T is not None:
            o.T[lev] = T
        if P is not None:
            o.P[lev] = P
        return o

    def decisions_line(self, lev, i):
        """Selection and region hashes of line i's own deposits in every record whose window
        holds it, at level lev (the line alone: no cascade)."""
        recs = [r for r in range(self.nrec) if self.lo[r] <= i < self.hi[r]]
        with Trace() as t:
            for r in recs:
                g = np.zeros(NGRID + NT, np.float64)
                rk = np.zeros(NT, np.float64)
                lib().mo_record(F64, C.byref(self.lines), C.byref(self.job), _ptr(self.tips),
                                float(self.P[lev]), float(self.T[lev]), float(self.rho[lev]),
                                float(self.origin[r]), int(i), int(i) + 1,
                                rk.ctypes.data_as(C.POINTER(C.c_double)), _ptr(g), None)
        return t.hashes[:2], recs

    def decisions_level(self, lev):
        """Selection, region and limiter hashes of every record of level lev (all lines)."""
        with Trace() as t:
            for r in range(self.nrec):
                self.record(F64, lev, r)
        return t.hashes

    def direct(self, lev, r, i0=0, stride=1):
        npts = len(range(i0, NT, stride))
        out = np.zeros(npts, np.float64)
        lib().mo_direct(C.byref(self.lines), C.byref(self.job), _ptr(self.tips),
                        float(self.P[lev]), float(self.T[lev]), float(self.rho[lev]),
                        float(self.origin[r]), int(self.lo[r]), int(self.hi[r]), int(i0),
                        int(stride), out.ctypes.data_as(C.POINTER(C.c_double)))
        return out


def voigt_r32(X, gD, gL, S, shape_id=0, chi_id=0, mol=2, T=296.0):
    return lib().mo_voigt_r32(X, gD, gL, S, shape_id, chi_id, mol, T)


def voigt_f64(X, gD, gL, S, shape_id=0, chi_id=0, mol=2, T=296.0):
    return lib().mo_voigt_f64(X, gD, gL, S, shape_id, chi_id, mol, T)


def chi_r32(chi_id, X, mol, T):
    return lib().mo_chi_r32(chi_id, X, mol, T)


def chi_f64(chi_id, X, mol, T):
    return lib().mo_chi_f64(chi_id, X, mol, T)


def spec(precision, tips, nu, S, gair, gself, E, n, molIso, delta, P, T, rho):
    out = np.zeros(6, np.float64)
    f = lib().mo_spec_r32 if precision == R32 else lib().mo_spec_f64
    tips = np.ascontiguousarray(tips, np.float32)
    f(_ptr(tips), nu, S, gair, gself, E, n, molIso, delta, P, T, rho,
      out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def limiter_hash(grids_list):
    """Limiter-choice hash of the FP64 cascade of each pre-cascade grid set (NGRID + NT
    doubles each, e.g. the GPU's marfa_grid_sums output), in order."""
    with Trace() as t:
        for g in grids_list:
            cascade(F64, np.asarray(g[:NGRID], np.float64), np.asarray(g[NGRID:], np.float64))
    return t.hashes[2]


def cascade(precision, grids, rk):
    if precision == R32:
        g = np.ascontiguousarray(grids, np.float32)
        r = np.array(rk, np.float32)
        lib().mo_cascade_r32(g.ctypes.data_as(C.POINTER(C.c_float)),
                             r.ctypes.data_as(C.POINTER(C.c_float)))
    else:
        g = np.ascontiguousarray(grids, np.float64)
        r = np.array(rk, np.float64)
        lib().mo_cascade_f64(g.ctypes.data_as(C.POINTER(C.c_double)),
                             r.ctypes.data_as(C.POINTER(C.c_double)))
    return r


def constants():
    out = (C.c_uint32 * 7)()
    lib().mo_constants(out)
    return list(out)


def wiso(i):
    return lib().mo_wiso(i)
