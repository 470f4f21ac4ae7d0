"""Python host side of the MI355X LBL core: ctypes over the C ABI in include/marfa.h.

The names mirror the reference's Fortran entry points (lblScheme -> lbl_scheme, the
level x subinterval loops of app/main.f90 -> pt_tables).  Every call goes to the HIP
library marfa_amd/libmarfa.so; there is no CPU fallback: if the library or a GPU is
missing, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

NT = 20481
RECORD_BYTES = NT * 4
FP64, FP32K = 0, 1
ACS, VAC = 0, 1
SHAPE_VOIGT, SHAPE_LORENTZ = 0, 1
CHI_NONE, CHI_TONKOV, CHI_POLLACK, CHI_PERRIN = 0, 1, 2, 3

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmarfa.so")

ERRORS = {2: "unsupported molecule", 3: "unreadable .par field", 9: "line index outside the line list", 20: "invalid argument", 21: "inputs not set",
          22: "HIP runtime error", 23: "no HIP device", 24: "temperature outside the TIPS table"}


class MarfaError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"marfa error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class _Job(C.Structure):
    _fields_ = [("start_wv", C.c_double), ("end_wv", C.c_double), ("cutoff", C.c_float),
                ("shape_id", C.c_int32), ("chi_id", C.c_int32), ("mol_code", C.c_int32),
                ("target", C.c_int32), ("precision", C.c_int32)]


_lib = None


def library() -> C.CDLL:
    """Load libmarfa.so (built by __graft_entry__.build / marfa_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    try:
        # PyTorch-ROCm ships its own HIP runtime: load it first, so that the library binds to
        # the same runtime instead of loading a second one (two runtimes in one process do
        # not share devices, and torch then finds no GPU)
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i64p, i32p, dp, fp = C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_void_p, C.c_void_p
    L.marfa_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    L.marfa_ctx_destroy.argtypes = [vp]
    L.marfa_last_error.argtypes = [vp]
    L.marfa_last_error.restype = C.c_char_p
    L.marfa_set_tips.argtypes = [vp, C.c_int32, C.c_int32, fp]
    L.marfa_set_lines.argtypes = [vp, C.c_int64] + [vp] * 8
    L.marfa_set_levels.argtypes = [vp, C.c_int32, fp, fp, fp]
    L.marfa_windows.argtypes = [vp, C.POINTER(_Job), C.c_int32, i32p, vp, vp, vp, vp]
    L.marfa_lbl_scheme.argtypes = [vp, i64p, C.c_double, C.c_double, C.c_double, C.c_float,
                                   C.c_float, C.c_float, C.c_float, C.c_int32, C.c_int32,
                                   C.c_int32, C.c_int32, fp]
    L.marfa_pt_tables.argtypes = [vp, C.POINTER(_Job), vp, vp, vp, C.c_int32]
    L.marfa_pt_tables_adjoint.argtypes = [vp, C.POINTER(_Job), vp, vp, vp, vp]
    L.marfa_pt_tables_f64.argtypes = [vp, C.POINTER(_Job), vp, vp]
    L.marfa_pt_tables_noalloc.argtypes = [vp, C.POINTER(_Job), vp, vp]
    L.marfa_pt_tables_adjoint_dev.argtypes = [vp, C.POINTER(_Job), vp, vp, vp, vp]
    L.marfa_update_line_params.argtypes = [vp, vp, vp, vp, vp, vp]
    L.marfa_keep_adjoint_state.argtypes = [vp, C.c_int32]
    L.marfa_ctx_info.argtypes = [vp, i64p, i32p, i32p]
    L.marfa_batch_create.argtypes = [C.c_int, C.c_int32, C.c_double, C.POINTER(C.c_void_p)]
    L.marfa_batch_destroy.argtypes = [vp]
    L.marfa_batch_last_error.argtypes = [vp]
    L.marfa_batch_last_error.restype = C.c_char_p
    L.marfa_batch_set_tips.argtypes = [vp, C.c_int32, C.c_int32, fp]
    L.marfa_batch_add_job.argtypes = [vp, C.POINTER(_Job), C.c_int64] + [vp] * 8 + [C.c_int32, fp, fp, fp, i32p]
    L.marfa_batch_tune.argtypes = [vp, vp, vp]
    L.marfa_batch_capture.argtypes = [vp]
    L.marfa_batch_launch.argtypes = [vp, vp, C.c_int32]
    L.marfa_batch_info.argtypes = [vp, C.c_int32, i32p, i32p, C.POINTER(C.c_void_p), C.POINTER(C.c_float)]
    L.marfa_batch_records.argtypes = [vp, C.c_int32, fp]
    L.marfa_count_samples.argtypes = [vp, C.POINTER(_Job), i64p]
    L.marfa_grid_sums.argtypes = [vp, C.POINTER(_Job), C.c_int32, C.c_int32, dp, fp]
    L.marfa_last_timings.argtypes = [vp, fp]
    L.marfa_species_code.argtypes = [C.c_char_p, C.c_char_p]
    L.marfa_par_convert.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, i64p]
    _lib = L
    return L


# every symbol the header declares (tests check the library exports them)
EXPORTS = ("marfa_ctx_create", "marfa_ctx_destroy", "marfa_last_error", "marfa_version",
           "marfa_set_tips", "marfa_set_lines", "marfa_set_levels", "marfa_windows",
           "marfa_lbl_scheme", "marfa_pt_tables", "marfa_pt_tables_f64", "marfa_pt_tables_adjoint",
           "marfa_pt_tables_adjoint_dev", "marfa_update_line_params", "marfa_ctx_info", "marfa_count_samples",
           "marfa_pt_tables_noalloc", "marfa_keep_adjoint_state",
           "marfa_grid_sums", "marfa_last_timings", "marfa_batch_create", "marfa_batch_destroy",
           "marfa_batch_last_error", "marfa_batch_set_tips", "marfa_batch_add_job", "marfa_batch_tune",
           "marfa_batch_capture", "marfa_batch_launch", "marfa_batch_info", "marfa_batch_records",
           "marfa_species_code", "marfa_par_convert", "marfa_log_loss", "marfa_log_loss_grad",
           "marfa_log_loss_f64", "marfa_log_loss_grad_f64")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# ---- line-list ingest (host code of the library; no device needed)
# 36-byte direct-access line record (app/processParFile.f90:176-177, read at src/LBL.f90:43)
LINE_RECORD = np.dtype([("nu", "<f8"), ("S", "<f4"), ("gamma_air", "<f4"), ("gamma_self", "<f4"),
                        ("elow", "<f4"), ("n_T", "<f4"), ("mol_iso", "<i4"), ("delta", "<f4")])


def species_code(species: str) -> str:
    """getSpeciesCode (app/processParFile.f90:189-223): 'CO2' -> '02'; raises MarfaError(2)
    for a molecule the reference does not support."""
    buf = C.create_string_buffer(3)
    rc = library().marfa_species_code(species.encode(), buf)
    if rc:
        raise MarfaError(rc, f"unsupported molecule: {species.strip()}")
    return buf.value.decode()


def par_to_db(species: str, par_path: str, db_path: str) -> int:
    """Convert a HITRAN .par file into the 36-byte line records (the reference's program
    app/processParFile.f90); returns the number of records written.  Errors as the
    reference's stop codes: 2 unsupported molecule, 3 unreadable field."""
    n = C.c_int64(0)
    rc = library().marfa_par_convert(species.encode(), os.fsencode(par_path), os.fsencode(db_path), C.byref(n))
    if rc:
        raise MarfaError(rc, f"{par_path} -> {db_path}")
    return int(n.value)


def read_db(db_path: str) -> np.ndarray:
    """The 36-byte line records of a database file as a structured array (LINE_RECORD)."""
    return np.fromfile(db_path, dtype=LINE_RECORD)


@dataclass
class JobSpec:
    start_wv: float
    end_wv: float
    cutoff: float
    shape_id: int = SHAPE_VOIGT
    chi_id: int = CHI_NONE
    mol_code: int = 2
    target: int = ACS
    precision: int = FP64

    @staticmethod
    def from_job(job, precision=FP64) -> "JobSpec":
        return JobSpec(job.start_wv, job.end_wv, job.cutoff, job.shape_id, job.chi_id,
                       job.mol_code, job.target, precision)

    def _c(self):
        return _Job(self.start_wv, self.end_wv, self.cutoff, self.shape_id, self.chi_id,
                    self.mol_code, self.target, self.precision)


class LBLContext:
    """One device context: line list, TIPS and levels live on the GPU."""

    def __init__(self, device: int = 0):
        self.lib = library()
        h = C.c_void_p()
        rc = self.lib.marfa_ctx_create(device, C.byref(h))
        if rc:
            raise MarfaError(rc, "could not create a context (is a GPU visible?)")
        self.h = h
        self.nlev = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.marfa_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise MarfaError(rc, self.lib.marfa_last_error(self.h).decode())

    # ---- inputs
    def set_tips(self, tips: np.ndarray):
        t = np.ascontiguousarray(tips, np.float32)
        self._check(self.lib.marfa_set_tips(self.h, t.shape[0], t.shape[1], _p(t)))

    def set_lines(self, lines):
        arrs = [np.ascontiguousarray(lines.nu, np.float64)]
        for f in ("S", "gair", "gself", "elow", "nT", "delta"):
            arrs.append(np.ascontiguousarray(getattr(lines, f), np.float32))
        arrs.append(np.ascontiguousarray(lines.molIso, np.int32))
        self._check(self.lib.marfa_set_lines(self.h, len(arrs[0]), *[_p(a) for a in arrs]))
        self.nlines = len(arrs[0])

    def set_levels(self, atm):
        P = np.ascontiguousarray(atm.P, np.float32)
        T = np.ascontiguousarray(atm.T, np.float32)
        rho = np.ascontiguousarray(atm.rho, np.float32)
        self._check(self.lib.marfa_set_levels(self.h, P.shape[0], _p(P), _p(T), _p(rho)))
        self.nlev = P.shape[0]

    def load_job(self, job, tips):
        self.set_tips(tips)
        self.set_lines(job.lines)
        self.set_levels(job.atm)

    # ---- windows
    def windows(self, spec: JobSpec):
        n = C.c_int32()
        j = spec._c()
        self._check(self.lib.marfa_windows(self.h, C.byref(j), 0, C.byref(n), None, None, None, None))
        m = n.value
        lo, hi = np.zeros(m, np.int64), np.zeros(m, np.int64)
        rec, org = np.zeros(m, np.int32), np.zeros(m, np.float64)
        self._check(self.lib.marfa_windows(self.h, C.byref(j), m, C.byref(n), _p(lo), _p(hi), _p(rec), _p(org)))
        return lo, hi, rec, org

    # ---- the reference's lblScheme (src/LBL.f90:25) between reset and cascade
    def lbl_scheme(self, line_idx: int, cap_wv: float, start_dwv: float, ext_end_dwv: float,
                   cutoff: float, P: float, T: float, pself: float, shape_id=SHAPE_VOIGT,
                   chi_id=CHI_NONE, mol_code=2, precision=FP64):
        idx = C.c_int64(line_idx)
        rk = np.zeros(NT, np.float32)
        self._check(self.lib.marfa_lbl_scheme(self.h, C.byref(idx), cap_wv, start_dwv, ext_end_dwv,
                                              cutoff, P, T, pself, shape_id, chi_id, mol_code,
                                              precision, _p(rk)))
        return idx.value, rk

    # ---- all levels x subintervals
    def pt_tables(self, spec: JobSpec) -> np.ndarray:
        lo, _, _, _ = self.windows(spec)
        out = np.zeros((self.nlev, lo.shape[0], NT), np.float32)
        j = spec._c()
        self._check(self.lib.marfa_pt_tables(self.h, C.byref(j), None, _p(out), None, 0))
        return out

    def pt_tables_device(self, spec: JobSpec, out_ptr: int, stream: int = 0, sync: bool = True):
        """Write into an existing device buffer (e.g. a torch tensor's data_ptr())."""
        j = spec._c()
        self._check(self.lib.marfa_pt_tables(self.h, C.byref(j), C.c_void_p(out_ptr), None,
                                             C.c_void_p(stream) if stream else None, 0 if sync else 1))

    def pt_tables_noalloc(self, spec: JobSpec, out_ptr: int, stream: int = 0):
        """marfa_pt_tables_noalloc: the records into an existing device buffer, with the
        library's fixed scratch (no device allocation); enqueued on stream, not waited for."""
        j = spec._c()
        self._check(self.lib.marfa_pt_tables_noalloc(self.h, C.byref(j), C.c_void_p(out_ptr),
                                                     C.c_void_p(stream) if stream else None))

    def pt_tables_torch(self, spec: JobSpec):
        import torch
        lo, _, _, _ = self.windows(spec)
        out = torch.empty((self.nlev, lo.shape[0], NT), dtype=torch.float32, device="cuda")
        st = torch.cuda.current_stream().cuda_stream
        self.pt_tables_device(spec, out.data_ptr(), st, sync=True)
        return out

    def pt_tables_f64_torch(self, spec: JobSpec):
        """Records before their FP32 rounding, as an FP64 CUDA tensor."""
        import torch
        lo, _, _, _ = self.windows(spec)
        out = torch.empty((self.nlev, lo.shape[0], NT), dtype=torch.float64, device="cuda")
        j = spec._c()
        self._check(self.lib.marfa_pt_tables_f64(self.h, C.byref(j), C.c_void_p(out.data_ptr()),
                                                 C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return out

    def keep_adjoint_state(self, on: bool = True):
        """Forwards keep the cascade's limiter choices for a following adjoint
        (marfa_keep_adjoint_state)."""
        self._check(self.lib.marfa_keep_adjoint_state(self.h, 1 if on else 0))

    def pt_tables_adjoint(self, spec: JobSpec, grad_out_ptr: int, stream: int = 0):
        """dL/d(S, gamma_air, n_T, delta) per line ([4][n], float64) and dL/d(T, P) per level
        ([2][nlev]) from dL/d(records) in device memory (marfa_pt_tables' layout)."""
        gl = np.zeros((4, self.nlines), np.float64)
        gv = np.zeros((2, self.nlev), np.float64)
        j = spec._c()
        self._check(self.lib.marfa_pt_tables_adjoint(self.h, C.byref(j), C.c_void_p(grad_out_ptr),
                                                     C.c_void_p(stream) if stream else None, _p(gl), _p(gv)))
        return gl, gv

    def count_samples(self, spec: JobSpec, coarse: bool = False):
        n = np.zeros(2, np.int64)
        j = spec._c()
        self._check(self.lib.marfa_count_samples(self.h, C.byref(j), n.ctypes.data_as(C.POINTER(C.c_int64))))
        return (int(n[0]), int(n[1])) if coarse else int(n[0])

    def grid_sums(self, spec: JobSpec, level: int, rec: int):
        g = np.zeros(30690 + NT, np.float64)
        r = np.zeros(NT, np.float32)
        j = spec._c()
        self._check(self.lib.marfa_grid_sums(self.h, C.byref(j), level, rec, _p(g), _p(r)))
        return g[:30690], g[30690:], r

    def timings(self):
        t = np.zeros(5, np.float32)
        self._check(self.lib.marfa_last_timings(self.h, _p(t)))
        return dict(zip(("spectroscopy", "sort", "coarse", "fine", "total"), t.tolist()))


NVARIANTS = 6
VARIANTS = [(512, True), (1024, True), (2048, True), (512, False), (1024, False), (2048, False)]


class Batch:
    """Many independent jobs recorded into one hipGraph (marfa_batch_* in include/marfa.h)."""

    def __init__(self, device: int = 0, lanes: int = 4, lane_budget_gb: float = 0.0):
        self.lib = library()
        h = C.c_void_p()
        rc = self.lib.marfa_batch_create(device, lanes, lane_budget_gb, C.byref(h))
        if rc:
            raise MarfaError(rc, "could not create a batch (is a GPU visible?)")
        self.h = h
        self.shapes = []

    def _check(self, rc):
        if rc:
            raise MarfaError(rc, self.lib.marfa_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.marfa_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tips(self, tips):
        t = np.ascontiguousarray(tips, np.float32)
        self._check(self.lib.marfa_batch_set_tips(self.h, t.shape[0], t.shape[1], _p(t)))

    def add_job(self, job, precision=FP32K) -> int:
        L, A = job.lines, job.atm
        arrs = [np.ascontiguousarray(L.nu, np.float64)]
        for f in ("S", "gair", "gself", "elow", "nT", "delta"):
            arrs.append(np.ascontiguousarray(getattr(L, f), np.float32))
        arrs.append(np.ascontiguousarray(L.molIso, np.int32))
        lv = [np.ascontiguousarray(x, np.float32) for x in (A.P, A.T, A.rho)]
        jid = C.c_int32()
        j = JobSpec.from_job(job, precision)._c()
        self._check(self.lib.marfa_batch_add_job(self.h, C.byref(j), len(arrs[0]), *[_p(a) for a in arrs],
                                                 len(lv[0]), *[_p(x) for x in lv], C.byref(jid)))
        nlev, nrec = C.c_int32(), C.c_int32()
        self._check(self.lib.marfa_batch_info(self.h, jid.value, C.byref(nlev), C.byref(nrec), None, None))
        self.shapes.append((nlev.value, nrec.value, NT))
        return jid.value

    def tune(self):
        n = len(self.shapes)
        v = np.zeros(max(n, 1), np.int32)
        t = np.zeros((max(n, 1), NVARIANTS), np.float32)
        self._check(self.lib.marfa_batch_tune(self.h, _p(v), _p(t)))
        return v[:n], t[:n]

    def capture(self):
        self._check(self.lib.marfa_batch_capture(self.h))

    def launch(self, stream: int = 0, sync: bool = True) -> float:
        """One replay of the graph; returns its duration [ms] when sync."""
        self._check(self.lib.marfa_batch_launch(self.h, C.c_void_p(stream) if stream else None, 1 if sync else 0))
        ms = C.c_float()
        self._check(self.lib.marfa_batch_info(self.h, -1, None, None, None, C.byref(ms)))
        return ms.value

    def records(self, job: int) -> np.ndarray:
        out = np.zeros(self.shapes[job], np.float32)
        self._check(self.lib.marfa_batch_records(self.h, job, _p(out)))
        return out
