"""Command line of the PT-table handlers (the reference's manage.py): `python -m
marfa_amd.manage convert` writes a level of the latest run as text, `python -m
marfa_amd.manage plot` plots a wavenumber range of a level as SVG.  The run is the
subdirectory of output/ptTables named in its latest_run.txt (written by the LBL run); level
and range are asked for on standard input, as the reference asks for them."""
from __future__ import annotations

import sys
from pathlib import Path

from . import pttable as pt


def latest_run(base: Path) -> Path:
    """The run directory latest_run.txt names; exits with the reference's message when the
    file is empty."""
    name = (Path(base) / "latest_run.txt").read_text().strip()
    if not name:
        sys.exit("Error: latest_run.txt is empty.")
    return Path(base) / name


def main(argv=None, ask=input, base=pt.PT_DIRECTORY, processed=pt.PROCESSED_DIRECTORY, plots=pt.PLOTS_DIRECTORY):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("Usage: manage.py <command> [options]")
        sys.exit(1)
    command = argv[0]
    if command not in ("convert", "plot"):
        print(f"Unknown command: {command}")
        sys.exit(1)
    run = latest_run(Path(base))
    level = int(ask("Enter the level number: "))
    if command == "convert":
        return pt.convert_pttable(run, level, Path(processed))
    vl = float(ask("Enter the left boundary for plot (cm-1): "))
    vr = float(ask("Enter the right boundary for plot (cm-1): "))
    x_data, y_data, parameters = pt.plot_parser(run, level, vl, vr)
    return pt.generate_plot(x_data, y_data, parameters, plots)


if __name__ == "__main__":
    main()
