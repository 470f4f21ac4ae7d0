"""Line-list ingest (app/processParFile.f90): HITRAN .par text -> 36-byte line records.

Pinned by the reference's own fixture (test/testDB.txt -> test/testDB.02, byte for byte)
and by golden cases written by the reference program itself (tests/golden/par/cases.json,
from tests/golden/gen_par_golden.py): Fortran's formatted-input rules (implied decimals,
blank and embedded-blank fields, exponent forms, short lines), isotopologue letters,
unknown isotopologues (record left as a hole), other molecules, and the stop codes 2
(unsupported molecule) and 3 (unreadable field).  Host code only: no GPU."""
import json
import os

import numpy as np
import pytest

from marfa_amd import lbl

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "par")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))


def test_fixture_byte_exact(tmp_path):
    out = tmp_path / "testDB.02"
    n = lbl.par_to_db("CO2", os.path.join(GOLD, "testDB.txt"), str(out))
    assert n == 5
    assert out.read_bytes() == open(os.path.join(GOLD, "testDB.02"), "rb").read()
    recs = lbl.read_db(str(out))
    assert recs["mol_iso"].tolist() == [1002] * 5          # CO2-626: N_MOL_ISO(1, 2) = 10
    assert recs["nu"][0] == 648.888078


@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_cases(tmp_path, name):
    c = CASES[name]
    src = tmp_path / "raw.par"
    src.write_text("\n".join(c["lines"]) + "\n")
    out = tmp_path / "case.db"
    if c["status"]:
        with pytest.raises(lbl.MarfaError) as e:
            lbl.par_to_db(c["species"], str(src), str(out))
        assert e.value.code == c["status"]
    else:
        lbl.par_to_db(c["species"], str(src), str(out))
    got = out.read_bytes() if out.exists() else b""
    # the records the reference wrote for the input lines; nothing after the last line
    assert got.hex() == c["records"], name


def test_species_codes():
    names = ["H2O", "CO2", "O3", "N2O", "CO", "CH4", "O2", "NO", "SO2", "NO2", "NH3", "HNO3"]
    for i, m in enumerate(names):
        assert lbl.species_code(m) == f"{i + 1:02d}"
        assert lbl.species_code(f"  {m} ") == f"{i + 1:02d}"   # trim(adjustl(...))
    with pytest.raises(lbl.MarfaError) as e:
        lbl.species_code("OH")
    assert e.value.code == 2


def test_unknown_iso_leaves_a_hole(tmp_path):
    c = CASES["unknown_iso"]
    recs = np.frombuffer(bytes.fromhex(c["records"]), lbl.LINE_RECORD)
    assert recs["nu"][1] == 0 and recs["mol_iso"][1] == 0          # record 2 never written
    assert recs["nu"][2] > 0


def test_missing_input_is_an_error(tmp_path):
    with pytest.raises(lbl.MarfaError) as e:
        lbl.par_to_db("CO2", str(tmp_path / "nope.par"), str(tmp_path / "x.db"))
    assert e.value.code == 20


CLI = os.path.join(os.path.dirname(GOLD), "..", "..", "marfa_amd", "fortran", "marfa_par")
needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="marfa_par not built (amdflang)")


def run_cli(cwd, *args):
    import subprocess
    return subprocess.run([os.path.abspath(CLI), *args], cwd=cwd, capture_output=True, text=True, timeout=60)


@needs_cli
@pytest.mark.parametrize("name", ["fixture", "iso_letters", "unknown_iso", "molecules", "implied_decimals"])
def test_cli_drop_in(tmp_path, name):
    """marfa_par run as the reference's program is run: same records, same text copy."""
    c = CASES[name]
    (tmp_path / "data" / "databases").mkdir(parents=True)
    (tmp_path / "raw.par").write_text("\n".join(c["lines"]) + "\n")
    r = run_cli(tmp_path, c["species"], "slug", "raw.par")
    assert r.returncode == 0, r.stdout + r.stderr
    code = lbl.species_code(c["species"])
    assert (tmp_path / "data" / "databases" / f"slug.{code}").read_bytes().hex() == c["records"]
    assert (tmp_path / f"slug.{code}.dat").read_text().splitlines() == c["dat"]


@needs_cli
def test_cli_stop_codes(tmp_path):
    (tmp_path / "data" / "databases").mkdir(parents=True)
    assert run_cli(tmp_path, "CO2", "x").returncode == 1                       # argument count
    (tmp_path / "raw.par").write_text("\n".join(CASES["read_error"]["lines"]) + "\n")
    assert run_cli(tmp_path, "XYZ", "x", "raw.par").returncode == 2           # molecule
    assert run_cli(tmp_path, "CO2", "x", "raw.par").returncode == 3           # field
