"""Generate tests/golden/par/cases.json from the reference's own .par preprocessing program.

Run in the container that holds the reference sources, after oracle/build_ref.sh has built
oracle/_ref/processParFile (app/processParFile.f90 compiled with amdflang):
    python tests/golden/gen_par_golden.py
Each case is a small .par text built from the reference's fixture line (test/testDB.txt)
with one aspect of Fortran's formatted input changed; the program is run on it and its
exit status and output records are stored (hex).  The records after the last input line
(the end-of-file pass, which the amdflang runtime fills with a copy of the last line and
the reference's fixture test/testDB.02 does not have) are stored separately as `eof_tail`;
`dat` holds the program's list-directed text copy of the records.
"""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PROG = os.path.join(ROOT, "oracle", "_ref", "processParFile")
BASE = open(os.path.join(HERE, "par", "testDB.txt")).read().split("\n")


def field(line: str, a: int, b: int, text: str) -> str:
    """line with columns [a, b) replaced by text (right-justified to the field width)"""
    assert len(text) <= b - a, (text, a, b)
    return line[:a] + text.rjust(b - a) + line[b:]


def cases() -> dict[str, tuple[str, list[str]]]:
    l0, l1, l2 = BASE[0], BASE[1], BASE[2]
    return {
        "fixture": ("CO2", BASE),
        "iso_letters": ("CO2", [l0[:2] + c + l0[3:] for c in "0ABCD"]),
        "unknown_iso": ("CO2", [l0, l1[:2] + "X" + l1[3:], l2]),
        "implied_decimals": ("CO2", [
            field(field(field(field(field(l0, 3, 15, "648888078"), 35, 40, "00705"), 45, 55, "2340834"),
                        55, 59, "0073"), 59, 67, "-0000709"),
            field(l1, 15, 25, "1384E-22"),
        ]),
        "exponent_forms": ("CO2", [field(l0, 15, 25, "1.023-019"), field(l1, 15, 25, "1.384D-19"),
                                   field(l2, 15, 25, "+1.377e-20"), field(l0, 15, 25, ".2E+2")]),
        "blank_fields": ("CO2", [field(field(l0, 59, 67, ""), 55, 59, ""), field(l1, 35, 40, "")]),
        "embedded_blanks": ("CO2", [field(l0, 45, 55, " 234. 0834"), field(l1, 3, 15, "2302 .649654")]),
        "short_line": ("CO2", [l0[:67], l1]),
        "molecules": ("H2O", [" 11" + l0[3:], " 12" + l1[3:], " 3G" + l2[3:], "12" + "2" + l0[3:]]),
        "read_error": ("CO2", [l0, field(l1, 35, 40, "0.0x6")]),
        "bad_species": ("XYZ", [l0]),
    }


def run_case(species: str, lines: list[str]) -> dict:
    with tempfile.TemporaryDirectory() as d:
        os.makedirs(os.path.join(d, "data", "databases"))
        src = os.path.join(d, "raw.par")
        with open(src, "w") as f:
            f.write("\n".join(lines) + "\n")
        r = subprocess.run([PROG, species, "case", "raw.par"], cwd=d, capture_output=True, text=True)
        out = b""
        dbs = os.listdir(os.path.join(d, "data", "databases"))
        if dbs:
            with open(os.path.join(d, "data", "databases", dbs[0]), "rb") as f:
                out = f.read()
        dat = [f for f in os.listdir(d) if f.endswith(".dat")]
        text = open(os.path.join(d, dat[0])).read().splitlines() if dat else []
    n = 36 * len(lines)
    # one text line per written record (lines with a known isotopologue); the end-of-file
    # pass may add one more
    kept = sum(1 for ln in lines if len(ln) > 2 and (ln[2] in "0123456789" or "A" <= ln[2] <= "I"))
    return {"species": species, "lines": lines, "status": r.returncode,
            "records": out[:n].hex(), "eof_tail": out[n:].hex(),
            "dat": text[:kept]}


def main():
    res = {name: run_case(sp, ls) for name, (sp, ls) in cases().items()}
    with open(os.path.join(HERE, "par", "cases.json"), "w") as f:
        json.dump(res, f, indent=1)
    for k, v in res.items():
        print(k, v["status"], len(v["records"]) // 72, "records", len(v["eof_tail"]) // 72, "tail")


if __name__ == "__main__":
    main()
