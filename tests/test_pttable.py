"""PT-table files (marfa_amd/pttable.py): the LBL run's writer and the Python handlers'
readers (handlers/constants.py, handlers/parsers.py, handlers/convert.py), on the
reference's layout (README.md:280-300; app/main.f90:99-112, :126, :163-168).
CPU only: synthetic records stand for the device output."""
import os

import numpy as np
import pytest

from marfa_amd import pttable as pt

NT = pt.POINTS_PER_RECORD


def test_file_names_match_the_level_labels():
    # app/main.f90:99-108 writes '1__', '12_', '123'; handlers/constants.py:15-34 reads them
    assert pt.get_pt_name("d", 1).name == "1__.ptbin"
    assert pt.get_pt_name("d", 9).name == "9__.ptbin"
    assert pt.get_pt_name("d", 10).name == "10_.ptbin"
    assert pt.get_pt_name("d", 99).name == "99_.ptbin"
    assert pt.get_pt_name("d", 100).name == "100.ptbin"
    assert pt.RECORD_SIZE == 81924


def test_record_numbers():
    # int((a + 1) / 10): 4000 -> 400, 4003.5 -> 400, 9.0 -> 1
    assert [pt.record_number(a) for a in (4000.0, 4003.5, 4010.0, 9.0, 10.0)] == [400, 400, 401, 1, 1]


def records(nlev, nrec, seed=3):
    rng = np.random.default_rng(seed)
    return rng.random((nlev, nrec, NT), dtype=np.float32) * np.float32(1e-20)


def test_write_then_read_back(tmp_path):
    recs = records(2, 3)
    paths = pt.write_pt_tables(tmp_path, recs, [401, 402, 403])
    assert [p.name for p in paths] == ["1__.ptbin", "2__.ptbin"]
    assert os.path.getsize(paths[0]) == 403 * pt.RECORD_SIZE
    raw = np.fromfile(paths[1], dtype="<f4").reshape(-1, NT)
    assert not raw[:400].any()                        # records below the first are holes
    np.testing.assert_array_equal(raw[400:], recs[1])
    wv, val = pt.base_parser(paths[0], 4010.0, 4030.0)   # records 401, 402
    assert len(wv) == len(val) == 2 * NT
    np.testing.assert_array_equal(np.array(val, np.float32), recs[0, :2].reshape(-1))
    assert wv[0] == np.float32(4010.0) and wv[NT - 1] == np.float32(4020.0)
    assert wv[1] == np.float32(4010.0 + 10.0 / 20480)


def test_scattered_records_and_levels(tmp_path):
    recs = records(3, 2, seed=4)
    pt.write_pt_tables(tmp_path, recs, [5, 9], first_level=11, levels=[2])
    p = pt.get_pt_name(tmp_path, 13)
    raw = np.fromfile(p, dtype="<f4").reshape(-1, NT)
    assert raw.shape[0] == 9 and not raw[5:8].any()
    np.testing.assert_array_equal(raw[4], recs[2, 0])
    np.testing.assert_array_equal(raw[8], recs[2, 1])
    assert not pt.get_pt_name(tmp_path, 11).exists()
    with pytest.raises(ValueError):
        pt.write_pt_tables(tmp_path, recs, [5, 9, 10])


def test_tensor_reader_equals_base_parser(tmp_path):
    torch = pytest.importorskip("torch")
    recs = records(1, 4, seed=5)
    p = pt.write_pt_tables(tmp_path, torch.from_numpy(recs), [100, 101, 102, 103])[0]
    wv, val = pt.base_parser(p, 1005.0, 1031.0)
    twv, tval = pt.read_pt_range(p, 1005.0, 1031.0)
    assert twv.dtype == torch.float32 and tval.dtype == torch.float32
    np.testing.assert_array_equal(twv.numpy(), np.array(wv, np.float32))
    np.testing.assert_array_equal(tval.numpy(), np.array(val, np.float32))


def test_reader_errors(tmp_path):
    p = pt.write_pt_tables(tmp_path, records(1, 2), [1, 2])[0]
    with pytest.raises(IndexError):
        pt.base_parser(p, 10.0, 40.0)                 # record 3 is past the end
    with pytest.raises(SystemExit):
        pt.base_parser(tmp_path / "7__.ptbin", 10.0, 20.0)


def write_info(d):
    # the lines the LBL run writes (marfa_amd/fortran/marfa_main.f90, app/main.f90:99-112)
    (d / "info.txt").write_text("Command-Line Arguments:\nInput Molecule: CO2\nStart Wavenumber: 4000\n"
                                "End Wavenumber: 4030\nCut Off: 25\nTarget Value: ACS\n"
                                "Atmospheric Profile File: atm.dat\n")


def test_plot_parser_and_convert(tmp_path):
    recs = records(2, 3, seed=6)
    pt.write_pt_tables(tmp_path, recs, [400, 401, 402])
    write_info(tmp_path)
    x, y, params = pt.plot_parser(tmp_path, 2, 4010.0, 4020.0)
    assert params == ("CO2", 25.0, "ACS", "atm.dat", 2, "CO2_4010-4020_ACS.svg")
    np.testing.assert_array_equal(np.array(y, np.float32), recs[1, 1])
    with pytest.raises(ValueError):
        pt.plot_parser(tmp_path, 2, 3990.0, 4020.0)
    out = pt.convert_pttable(tmp_path, 1, tmp_path / "processed")
    assert out.name == "CO2_4000-4030_1level.dat"
    lines = out.read_text().splitlines()
    assert lines[0] == "Command-Line Arguments:" and len(lines) == 7 + 3 * NT
    assert lines[7] == f"{np.float32(4000.0):15.5f} {recs[0, 0, 0]:17.7e}"


def test_generate_plot_writes_svg(tmp_path):
    """handlers/plot.py:13-64: an SVG figure named as plot_parser names it, in the plots
    directory; an unknown target value is refused before anything is written."""
    import xml.etree.ElementTree as ET
    recs = records(2, 3, seed=7)
    pt.write_pt_tables(tmp_path, recs, [400, 401, 402])
    write_info(tmp_path)
    x, y, params = pt.plot_parser(tmp_path, 1, 4000.0, 4010.0)
    path = pt.generate_plot(x, y, params, tmp_path / "plots")
    assert path == tmp_path / "plots" / "CO2_4000-4010_ACS.svg"
    root = ET.parse(path).getroot()
    assert root.tag.endswith("svg")
    assert any(el.tag.endswith("path") for el in root.iter())
    with pytest.raises(ValueError):
        pt.generate_plot(x, y, params[:2] + ("XYZ",) + params[3:5] + ("bad.svg",), tmp_path / "plots")
    assert not (tmp_path / "plots" / "bad.svg").exists()


def test_manage_commands(tmp_path):
    """The reference's manage.py commands over the run latest_run.txt names."""
    from marfa_amd import manage
    base = tmp_path / "ptTables"
    run = base / "CO2_test"
    recs = records(2, 3, seed=8)
    pt.write_pt_tables(run, recs, [400, 401, 402])
    write_info(run)
    (base / "latest_run.txt").write_text("CO2_test\n")
    answers = iter(["1"])
    out = manage.main(["convert"], ask=lambda _: next(answers), base=base, processed=tmp_path / "proc",
                      plots=tmp_path / "plots")
    assert out == tmp_path / "proc" / "CO2_4000-4030_1level.dat"
    answers = iter(["2", "4010", "4030"])
    out = manage.main(["plot"], ask=lambda _: next(answers), base=base, processed=tmp_path / "proc",
                      plots=tmp_path / "plots")
    assert out == tmp_path / "plots" / "CO2_4010-4030_ACS.svg" and out.exists()
    with pytest.raises(SystemExit):
        manage.main(["unknown"], base=base)
    (base / "latest_run.txt").write_text("")
    with pytest.raises(SystemExit):
        manage.main(["convert"], ask=lambda _: "1", base=base)
