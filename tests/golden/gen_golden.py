"""Generate the golden fixtures in tests/golden/ from the reference itself.

Run in the container that holds the reference sources (after oracle/build_ref.sh):
    python tests/golden/gen_golden.py
It drives the reference's compiled code only -- oracle/_ref/ref_probe (unit functions of
the reference's modules), oracle/_ref/marfa_ref (the reference's own program,
app/main.f90) and oracle/_ref/ref_driver (its modules with a chi factor / shape chosen at
run time) -- and stores inputs with the raw bit patterns of the outputs:

  units.npz    constants, WISO, TIPS checksum, voigt / lorentzShape / chi / spectroscopy
               vectors, cascade outputs for seeded grid contents
  records.npz  end-to-end PT-table records of small synthetic jobs: SHA-256 of every
               record's bytes plus every 16th value (and the constructed line list of the
               boundary case)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from marfa_amd import synth  # noqa: E402
import refrun  # noqa: E402

SQLN2 = float(np.frombuffer(bytes.fromhex("3F55224D")[::-1], np.float32)[0])
STRIDE = 16


def f32s(x):
    return np.format_float_scientific(np.float32(x), unique=True)


def hexf(s):
    return np.uint32(int(s, 16))


# ----------------------------------------------------------------------------- units
def voigt_inputs(rng):
    rows = []
    vxs = np.concatenate([np.linspace(0, 30, 61), [1.0, 1.0 + 1e-6, np.sqrt(1.4), 1.1833, 5.0, 5.0 + 1e-5,
                                                   5.5, 14.999, 15.0, 15.001, 0.0]])
    vys = [1e-4, 1e-3, 0.005, 0.0199, 0.02, 0.0201, 0.1, 0.5, 1.0, 3.0, 6.0, 14.0, 20.0]
    for vx in vxs:
        for vy in vys:
            gD = float(np.float32(rng.uniform(0.002, 0.02)))
            gL = float(np.float32(vy * gD / SQLN2))
            X = vx * gD / SQLN2 * (1 if rng.random() < 0.5 else -1)
            chi = int(rng.integers(0, 4))
            rows.append((X, gD, gL, float(np.float32(10 ** rng.uniform(-24, -19))), 2, chi,
                         float(np.float32(rng.uniform(160, 750)))))
    for _ in range(400):  # far wings, chi active
        gD = float(np.float32(rng.uniform(0.002, 0.02)))
        gL = float(np.float32(rng.uniform(1e-4, 2.0)))
        X = rng.uniform(-300, 300)
        rows.append((X, gD, gL, float(np.float32(10 ** rng.uniform(-24, -19))), int(rng.integers(1, 3)),
                     int(rng.integers(0, 4)), float(np.float32(rng.uniform(160, 750)))))
    return np.array(rows, np.float64)


def gen_units(out):
    rng = np.random.Generator(np.random.PCG64(1203))
    const = [hexf(x) for x in refrun.probe("const\n")[0].split()]
    wiso = [hexf(x) for x in refrun.probe("wiso\n")]
    tips_bits = np.array([int(x, 16) for x in " ".join(refrun.probe("tips\n")).split()], np.uint32)
    tips_sha = hashlib.sha256(tips_bits.tobytes()).hexdigest()

    vin = voigt_inputs(rng)
    cmds = "".join(f"voigt {float(r[0])!r} {f32s(r[1])} {f32s(r[2])} {f32s(r[3])} {int(r[4])} {int(r[5])} {f32s(r[6])}\n"
                   for r in vin)
    vout = np.array([hexf(x) for x in refrun.probe(cmds)], np.uint32)

    lin = vin[:, [0, 2, 3, 4, 5, 6]]
    cmds = "".join(f"lorshape {float(r[0])!r} {f32s(r[1])} {f32s(r[2])} {int(r[3])} {int(r[4])} {f32s(r[5])}\n" for r in lin)
    lout = np.array([hexf(x) for x in refrun.probe(cmds)], np.uint32)

    cin = []
    for chi in range(4):
        for X in np.concatenate([np.linspace(-400, 400, 161), [3.0, -3.0, 10.0, 30.0, 120.0, 150.0, 300.0,
                                                                2.9999, 3.0001, 149.999, 150.001]]):
            for mol in (1, 2):
                cin.append((chi, X, mol, float(np.float32(rng.uniform(160, 750)))))
    cin = np.array(cin, np.float64)
    cmds = "".join(f"chi {int(r[0])} {float(r[1])!r} {int(r[2])} {f32s(r[3])}\n" for r in cin)
    cout = np.array([hexf(x) for x in refrun.probe(cmds)], np.uint32)

    sin = []
    for _ in range(600):
        iso = int(rng.choice([101, 1002, 5506, 4605, 2303, 1102, 6105]))
        P = float(np.float32(10 ** rng.uniform(-5, 2)))
        T = float(np.float32(rng.uniform(160, 1000)))
        rho = float(synth.vmr_density(np.array([P]), np.array([T]), rng.uniform(1e-3, 1.0))[0])
        sin.append((rng.uniform(10, 20000), float(np.float32(10 ** rng.uniform(-32, -18))),
                    float(np.float32(rng.uniform(0.05, 0.1))), float(np.float32(rng.uniform(0.07, 0.12))),
                    float(np.float32(rng.uniform(0, 8000))), float(np.float32(rng.uniform(0.5, 0.8))), iso,
                    float(np.float32(rng.uniform(-0.01, 0))), P, T, rho))
    sin.append((4050.123456, 1e-21, 0.07, 0.1, 1000.0, 0.7, 1002, -0.005, 10.0, 296.0, 1e26))  # T = Tref
    sin = np.array(sin, np.float64)
    cmds = "".join(f"spec {float(r[0])!r} {f32s(r[1])} {f32s(r[2])} {f32s(r[3])} {f32s(r[4])} {f32s(r[5])} "
                   f"{int(r[6])} {f32s(r[7])} {f32s(r[8])} {f32s(r[9])} {f32s(r[10])}\n" for r in sin)
    sout = []
    for line in refrun.probe(cmds):
        p = line.split()
        sout.append([int(p[0], 16), int(p[1], 16), int(p[2], 16), int(p[3], 16), int(p[4], 16), int(p[5], 16)])
    sout = np.array(sout, np.uint64)

    # cascade: seeded grid contents -> RK
    casc = {}
    for seed, kind in ((1, "uniform"), (2, "smooth")):
        g = cascade_input(seed, kind)
        with tempfile.NamedTemporaryFile("w", suffix=".hex", delete=False) as f:
            f.write("\n".join(f"{b:08X}" for b in g.view(np.uint32)))
            fn = f.name
        rk = np.array([hexf(x) for x in refrun.probe(f"cascade \"{fn}\"\n")], np.uint32)
        os.unlink(fn)
        casc[kind] = rk
    np.savez_compressed(out, const=np.array(const, np.uint32), wiso=np.array(wiso, np.uint32),
                        tips_sha=np.array(tips_sha), voigt_in=vin, voigt_out=vout, lor_in=lin, lor_out=lout,
                        chi_in=cin, chi_out=cout, spec_in=sin, spec_out=sout,
                        cascade_uniform=casc["uniform"], cascade_smooth=casc["smooth"])
    print("units:", len(vin), "voigt,", len(cin), "chi,", len(sin), "spec", flush=True)


def cascade_input(seed, kind):
    """51,171 FP32 values: grids 0..9 (P, mid, L blocks) then RK."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 30690 + 20481
    if kind == "uniform":
        return rng.uniform(0, 1, n).astype(np.float32)
    # smooth: sums of Lorentzians sampled at each grid's node positions
    centers = rng.uniform(0, 10, 40)
    widths = rng.uniform(0.001, 0.3, 40)
    out = []
    for k in range(10):
        h = 2.0 ** -k
        ncell = 10 * 2 ** k
        left = np.arange(ncell) * h
        for off in (0.0, 0.5 * h, h):
            x = left + off
            v = sum(w / ((x - c) ** 2 + w * w) for c, w in zip(centers, widths)) * 1e-3
            out.append(v.astype(np.float32))
    out.append(np.zeros(20481, np.float32))
    return np.concatenate(out)


# ----------------------------------------------------------------------------- records
def boundary_lines(seed=31):
    """Lines at the exact positions where the reference's selection changes branch
    (delta = 0, so the shifted centre equals the line centre), plus a random background."""
    H, H9 = 2.0 ** -11, 2.0 ** -9
    a = 4000.0
    special = [a, a + 10.0, a + 20.0, a - 25.0, a - 15.0, a + 45.0, a + 35.0, a - 1.0, a - 0.5, a - 0.25,
               a - H9, a - H9 / 2, a - H, a + 2.0, a + 1.0, a + H, a + H / 2, a + 9.999, a + 10.0 - H,
               a + 2 * H9, a + 5.5, a + 1.75, a + 11.0, a + 10.5, a + 10.25, a + 10.0 + H9, a + 10.0 + H9 / 2,
               a + 10.0 + H, a + 12.0, a + 19.0 + H, a + 20.0 - H9, a + 30.0 + H9, a + 31.0, a + 0.75 + H,
               a + 3.0 - H / 2, a + 7.0 + H9 / 3]
    rng = np.random.Generator(np.random.PCG64(seed))
    bg = rng.uniform(a - 30.0, a + 50.0, 400)
    nu = np.concatenate([np.array(special), bg])
    is_special = np.concatenate([np.ones(len(special), bool), np.zeros(len(bg), bool)])
    p = synth.draw_params(rng, nu.shape[0])
    p["delta"][is_special] = 0.0
    o = np.argsort(nu, kind="stable")
    lines = synth.Lines(nu=nu[o], **{k: v[o] for k, v in p.items()})
    return lines


def record_cases():
    lev3 = synth.log_levels(3, 90.0, 1e-3, 735.0, 170.0, 0.965)
    cases = {
        "tiny": dict(gen=("uniform", 2000, 3985.0, 4045.0, 7), atm=("log", 3, 90.0, 1e-3, 735.0, 170.0, 0.965),
                     a=4000.0, b=4030.0, cut=25.0),
        "carry": dict(gen=("uniform", 30000, 3900.0, 4200.0, 11), atm=("log", 4, 90.0, 1e-3, 735.0, 170.0, 0.965),
                      a=4000.0, b=4100.0, cut=25.0),
        "tonkov": dict(gen=("uniform", 5000, 3950.0, 4080.0, 12), atm=("log", 3, 90.0, 1e-3, 735.0, 170.0, 0.965),
                       a=4000.0, b=4030.0, cut=125.0, chi=1),
        "pollack_vac": dict(gen=("uniform", 5000, 3950.0, 4080.0, 13), atm=("log", 3, 90.0, 1e-3, 735.0, 170.0, 0.965),
                            a=4000.0, b=4030.0, cut=125.0, chi=2, target=1),
        "perrin": dict(gen=("uniform", 5000, 3700.0, 4330.0, 14), atm=("log", 3, 90.0, 1e-3, 735.0, 170.0, 0.965),
                       a=4000.0, b=4030.0, cut=250.0, chi=3),
        "lorentz_shape": dict(gen=("uniform", 3000, 3950.0, 4080.0, 15), atm=("log", 2, 10.0, 1e-2, 300.0, 200.0, 0.965),
                              a=4000.0, b=4030.0, cut=125.0, chi=2, shape=1),
        "boundaries": dict(gen=("boundary",), atm=("log", 3, 90.0, 1e-3, 735.0, 170.0, 0.965),
                           a=4000.0, b=4020.0, cut=25.0),
        "lowp": dict(gen=("uniform", 2000, 3985.0, 4045.0, 16), atm=("log", 3, 1e-2, 1e-5, 300.0, 160.0, 0.965),
                     a=4000.0, b=4030.0, cut=25.0),
        "h2o": dict(gen=("uniform", 3000, 1470.0, 1560.0, 17, 101), atm=("log", 3, 10.0, 1e-4, 300.0, 180.0, 0.01),
                    a=1500.0, b=1530.0, cut=25.0, mol=1),
    }
    del lev3
    return cases


def make_job(c):
    g = c["gen"]
    if g[0] == "uniform":
        kw = {"molIso": g[5]} if len(g) > 5 else {}
        lines = synth.uniform_lines(g[1], g[2], g[3], g[4], **kw)
    else:
        lines = boundary_lines()
    at = c["atm"]
    atm = synth.log_levels(at[1], at[2], at[3], at[4], at[5], at[6])
    return synth.Job(lines, atm, c["a"], c["b"], c["cut"], mol_code=c.get("mol", 2), chi_id=c.get("chi", 0),
                     shape_id=c.get("shape", 0), target=c.get("target", 0))


def lines_sha(lines):
    return hashlib.sha256(lines.to_records().tobytes()).hexdigest()


def gen_records(out):
    cases = record_cases()
    arrays, meta = {}, {}
    for name, c in cases.items():
        job = make_job(c)
        recs, wall, _ = refrun.run(job)
        if name == "tiny":  # the driver over the same modules must equal the reference's program
            drv, _, _ = refrun.run(job, use_driver=True)
            assert np.array_equal(drv.view(np.uint32), recs.view(np.uint32)), "driver differs from marfa_ref"
        shas = [[hashlib.sha256(recs[l, r].tobytes()).hexdigest() for r in range(recs.shape[1])]
                for l in range(recs.shape[0])]
        arrays[f"{name}__strided"] = recs[:, :, ::STRIDE].copy()
        if name == "boundaries":
            for f in synth.FIELDS:
                arrays[f"{name}__line_{f}"] = getattr(job.lines, f)
        meta[name] = dict(case=c, sha=shas, lines_sha=lines_sha(job.lines), wall=wall,
                          shape=list(recs.shape))
        print(f"record case {name}: {recs.shape} in {wall:.2f}s", flush=True)
    arrays["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(out, **arrays)


if __name__ == "__main__":
    assert refrun.available(), "build the reference first: oracle/build_ref.sh"
    gen_units(os.path.join(HERE, "units.npz"))
    gen_records(os.path.join(HERE, "records.npz"))
