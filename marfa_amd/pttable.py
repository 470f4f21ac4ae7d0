"""PT-table files: the direct-access writer of the LBL run and the readers of the Python
handlers, over the records the HIP core produces.

The layout is the reference's (README.md:280-300): one file `<level label>.ptbin` per
atmospheric level (label '1__', '12_', '123': app/main.f90:99-108, handlers/constants.py:
15-34), records of NT = 20,481 little-endian FP32 values (81,924 bytes) covering 10 cm-1,
record number int((a + 1) / 10) for the subinterval starting at a (app/main.f90:126,
:163-168); records below the first written one are file holes (zeros).  Readers follow
handlers/parsers.py (base_parser: record int(v / 10), points at 10 * record + i * 10/20480)
and handlers/convert.py; torch versions return tensors (on a device if asked) for the
Python side of the pipeline.
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np

POINTS_PER_RECORD = 20481
RECORD_WV_SPAN = 10  # cm-1 per record
RECORD_SIZE = POINTS_PER_RECORD * 4
PT_EXTENSION = "ptbin"
INFO_FILENAME = "info.txt"
PT_DIRECTORY = "output/ptTables"        # runs, and latest_run.txt naming the last one
PROCESSED_DIRECTORY = "output/processedData"
PLOTS_DIRECTORY = "output/plots"


def get_pt_name(directory: Path, level: int) -> Path:
    """PT-table file of an atmospheric level (1-based): 1__.ptbin, 12_.ptbin, 123.ptbin
    (handlers/constants.py:15-34; the labels app/main.f90:99-108 writes): the level number
    padded with underscores to three characters."""
    return Path(directory) / f"{str(level).ljust(3, '_')}.{PT_EXTENSION}"


def record_number(a: float) -> int:
    """Record of the subinterval starting at a (app/main.f90:126)."""
    return int((a + 1.0) / 10.0)


def write_pt_tables(directory, records, recnums, first_level: int = 1, levels=None) -> list[Path]:
    """Write PT-table files as the LBL run does (app/main.f90:99-112, :163-168).

    records: [levels][subintervals][NT] FP32 (numpy or torch, host or device); recnums:
    record number of each subinterval (marfa_windows / record_number).  Level i of the
    array goes to file get_pt_name(directory, first_level + i); `levels` picks a subset.
    Records are written at (recnum - 1) * RECORD_SIZE; earlier records stay holes."""
    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    recnums = np.asarray(recnums, dtype=np.int64)
    if recnums.ndim != 1 or (recnums < 1).any():
        raise ValueError("record numbers must be a 1-d array of positive integers")
    nlev = records.shape[0]
    if records.shape[1] != recnums.shape[0] or records.shape[2] != POINTS_PER_RECORD:
        raise ValueError(f"records shape {tuple(records.shape)} does not match {recnums.shape[0]} records")
    out = []
    for i in (range(nlev) if levels is None else levels):
        lev = records[i]
        if hasattr(lev, "detach"):  # torch tensor, possibly on the GPU
            lev = lev.detach().to("cpu").numpy()
        lev = np.ascontiguousarray(lev, dtype="<f4")
        path = get_pt_name(directory, first_level + i)
        with open(path, "wb") as f:
            contiguous = bool(np.all(np.diff(recnums) == 1))
            if contiguous:  # one write from the first record on
                f.seek(int(recnums[0] - 1) * RECORD_SIZE)
                f.write(lev.tobytes())
            else:
                for r, rec in zip(recnums, lev):
                    f.seek(int(r - 1) * RECORD_SIZE)
                    f.write(rec.tobytes())
        out.append(path)
    return out


def _record_span(v1: float, v2: float) -> tuple[int, int]:
    """First and last record of the wavenumber range [v1, v2] as the handlers take it
    (handlers/parsers.py:23-25): int(v1 / 10) .. int((v2 - 1) / 10)."""
    return int(v1 / 10.0), int((v2 - 1) / 10.0)


def _wavenumbers(first: int, last: int) -> np.ndarray:
    """FP32 wavenumbers of the points of records first..last: 10 * record + i * 10/20480,
    formed in FP64 and rounded once, as the handlers form them."""
    step = RECORD_WV_SPAN / (POINTS_PER_RECORD - 1)
    i = np.arange(POINTS_PER_RECORD, dtype=np.float64)
    rec = np.arange(first, last + 1, dtype=np.int64)[:, None]
    return (RECORD_WV_SPAN * rec + i[None, :] * step).astype(np.float32)


def base_parser(pttable_file: Path, v1: float, v2: float) -> tuple[list[np.float32], list[np.float32]]:
    """Wavenumbers and values of records int(v1/10) .. int((v2-1)/10) of a PT-table file,
    as handlers/parsers.py:17-65 returns them (lists of np.float32): IndexError for a record
    past the end of the file; a missing file prints the reference's message and exits."""
    first, last = _record_span(v1, v2)
    try:
        size = os.path.getsize(pttable_file)
    except FileNotFoundError:
        print("Error: invalid number of atmospheric level")
        sys.exit()
    # records are read in one piece once every one of them is known to start inside the file
    past = [r for r in range(first, last + 1) if (r - 1) * RECORD_SIZE >= size]
    if past:
        raise IndexError(f"Record number {past[0]} exceeds a file size.")
    n = last - first + 1
    with open(pttable_file, "rb") as f:
        f.seek((first - 1) * RECORD_SIZE)
        raw = f.read(n * RECORD_SIZE)
    if len(raw) < n * RECORD_SIZE:  # the handlers index every point of every record
        raise IndexError(f"Record number {last} is cut short by the end of the file.")
    vals = np.frombuffer(raw, dtype="<f4")
    return list(_wavenumbers(first, last).reshape(-1)), list(vals)


def read_pt_range(pttable_file: Path, v1: float, v2: float, device=None):
    """base_parser's points as tensors (wavenumber, value), FP32, on `device` if given:
    one read of the whole record range, the wavenumbers formed as base_parser forms them."""
    import torch
    start, end = _record_span(v1, v2)
    n = end - start + 1
    size = os.path.getsize(pttable_file)
    if (end - 1) * RECORD_SIZE >= size:
        raise IndexError(f"Record number {end} exceeds a file size.")
    with open(pttable_file, "rb") as f:
        f.seek((start - 1) * RECORD_SIZE)
        vals = np.frombuffer(f.read(n * RECORD_SIZE), dtype="<f4").reshape(n, POINTS_PER_RECORD)
    wv_t = torch.from_numpy(_wavenumbers(start, end).reshape(-1))
    val_t = torch.from_numpy(vals.reshape(-1).copy())
    if device is not None:
        wv_t, val_t = wv_t.to(device), val_t.to(device)
    return wv_t, val_t


def read_info(directory: Path) -> dict:
    """The 'Key: value' lines of a run's info.txt (written by the LBL run, read by
    handlers/parsers.py:85-99 and handlers/convert.py:27-36)."""
    out = {}
    with open(Path(directory) / INFO_FILENAME) as info:
        for line in info.readlines():
            if ":" in line:
                k, v = line.split(":", 1)
                out[k.strip()] = v.strip()
    return out


def plot_parser(directory: Path, level: int, vl: float, vr: float):
    """handlers/parsers.py:68-110 without the plot: data of [vl, vr] from a level's file and
    the plot parameters (molecule, cutoff, y title, atmosphere file, level, file name);
    ValueError when [vl, vr] is outside the run's [Start, End] wavenumbers."""
    info = read_info(directory)
    v1, v2 = float(info["Start Wavenumber"]), float(info["End Wavenumber"])
    molecule = info["Input Molecule"]
    cutoff = float(info["Cut Off"])
    y_title = info["Target Value"]
    atm_filename = info["Atmospheric Profile File"]
    if not v1 <= vl < vr <= v2:
        raise ValueError(f"Left and right boundaries requested for plotting: {vl}, {vr} cm-1 are "
                         f"outside the spectre range: {v1}, {v2} cm-1.")
    x_data, y_data = base_parser(get_pt_name(directory, level), vl, vr)
    plot_filename = f"{molecule}_{int(vl)}-{int(vr)}_{y_title}.svg"
    return x_data, y_data, (molecule, cutoff, y_title, atm_filename, level, plot_filename)


def convert_pttable(directory: Path, level: int, out_dir: Path) -> Path:
    """handlers/convert.py:11-50: a level's whole run range as text, after a copy of info.txt:
    one line per point, f'{wavenumber:15.5f} {value:17.7e}'."""
    info = read_info(directory)
    v1, v2 = float(info["Start Wavenumber"]), float(info["End Wavenumber"])
    molecule = info["Input Molecule"]
    vw_data, absorption_data = base_parser(get_pt_name(directory, level), v1, v2)
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    output_file = out_dir / f"{molecule.strip()}_{int(v1)}-{int(v2)}_{level}level.dat"
    with open(Path(directory) / INFO_FILENAME) as src, open(output_file, "w") as out:
        out.write(src.read())
        for vw, a in zip(vw_data, absorption_data):
            out.write(f"{vw:15.5f} {a:17.7e}\n")
    return output_file


Y_AXIS_TITLES = {"ACS": "Absorption cross-section [cm$^{2}$ mol$^{-1}$]",
                 "VAC": "Volume absorption coefficient [km$^{-1}$]"}


def generate_plot(x_data, y_data, parameters, plots_dir=PLOTS_DIRECTORY) -> Path:
    """The spectrum of plot_parser's data as an SVG figure (handlers/plot.py:13-64): log10 of
    the values against wavenumber, titled with the molecule, level, atmosphere file and
    cutoff, the y axis named by the target value; saved under plots_dir as the file name
    plot_parser chose.  ValueError for a target other than ACS or VAC.  Drawn on its own
    Figure (no pyplot state), so repeated calls do not accumulate figures."""
    from matplotlib.backends.backend_svg import FigureCanvasSVG
    from matplotlib.figure import Figure
    molecule, cutoff, y_title, atm_filename, level, plot_filename = parameters
    if y_title not in Y_AXIS_TITLES:
        raise ValueError(f"Unknown target value {y_title}")
    x = np.asarray(x_data, dtype=np.float64)
    with np.errstate(divide="ignore"):
        y = np.log10(np.asarray(y_data, dtype=np.float64))
    fig = Figure(figsize=(12, 6))
    FigureCanvasSVG(fig)
    ax = fig.add_subplot()
    ax.plot(x, y, color="g", linestyle="-")
    ax.set_title(f"Absorption spectrum of {molecule} at {level} level of atmosphere {atm_filename}\n"
                 f"Cutoff is {cutoff} cm-1")
    ax.set_xlabel("Wavenumber [cm$^{-1}$]")
    ax.set_ylabel(Y_AXIS_TITLES[y_title])
    ax.grid(which="major", axis="both", color="gray", alpha=0.5)
    ax.minorticks_on()
    ax.tick_params(axis="y", which="both", right=True)
    out = Path(plots_dir)
    out.mkdir(parents=True, exist_ok=True)
    path = out / plot_filename
    fig.savefig(path, format="svg", bbox_inches="tight")
    print("Success!")
    print(f"Plot is saved as {path}")
    return path

