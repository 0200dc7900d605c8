"""The reference's on-disk formats: particle bundles (MCC_jw ParticleBundle.f90:286-375),
the restart dump (Dump_2D.f90, Restart_2D.f90) and the Tecplot field file
(Output_To_Tecplot_IJK_2D.f90:144-161).

Pinned by files a Fortran compiler wrote (tests/golden/*_fixture.f90, run with the image's
flang by make_*_fixture.sh): the unformatted record framing of a bundle, and E15.6 text for
224 values including rounding carries, three-digit exponents, subnormals and -0."""
import struct
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def read_records(path):
    """Fortran unformatted sequential records: [len][payload][len], little-endian."""
    b = Path(path).read_bytes()
    out, o = [], 0
    while o < len(b):
        n = struct.unpack_from("<i", b, o)[0]
        out.append(b[o + 4:o + 4 + n])
        assert struct.unpack_from("<i", b, o + 4 + n)[0] == n
        o += 8 + n
    return out


def parse_bundle(path):
    rec = read_records(path)
    npar, nnorm = struct.unpack("<2i", rec[0])
    charge, mass, weight = struct.unpack("<3d", rec[1])
    assert len(rec) == 2 + npar
    parts = np.array([struct.unpack("<10d", r[:80]) for r in rec[2:]]).reshape(-1, 10)
    loc = np.array([struct.unpack("<i", r[80:84])[0] for r in rec[2:]], dtype=np.int64)
    assert all(len(r) == 84 for r in rec[2:])
    return npar, nnorm, (charge, mass, weight), parts, loc


# ------------------------------------------------------------------ CPU
def test_bundle_fixture_layout():
    """The flang-written bundle: 8-, 24- and 84-byte records (10 REAL(8) + one INTEGER, no
    padding), read back as written."""
    npar, nnorm, phys, parts, loc = parse_bundle(GOLDEN / "bundle_flang.dat")
    assert (npar, nnorm) == (3, 1000)
    assert phys == (-1.60217663e-19, 9.1093837e-31, 2.5e6)
    i = np.arange(1, 4)
    np.testing.assert_array_equal(parts[:, 0], 1.0e-6 * i)
    np.testing.assert_array_equal(parts[:, 3], 1.0e5 * i)
    np.testing.assert_array_equal(parts[:, 4], -2.0e5 * i)
    np.testing.assert_array_equal(loc, 96 * i + 1)


def test_fortran_e15_matches_flang(pkg):
    """The library's Ew.d writer against flang's E15.6 for every fixture value."""
    n = 0
    for line in (GOLDEN / "e15_flang.txt").read_text().splitlines():
        bits, text = int(line[:21]), line[22:]
        v = struct.unpack("<d", struct.pack("<q", bits))[0]
        assert pkg.format_fortran_e(v, 15, 6) == text, (v, text)
        n += 1
    assert n == 224


# ------------------------------------------------------------------ GPU
def _stepped(pl, pkg, case, bits):
    w = pkg.build_workload(pl, case, seed=7, ppc=4, mcc=False)
    pl.setup(w["domain"], bits=bits, solver=0, tol=1e-10)
    dom = w["domain"]
    nel = (dom.nnx - 1) * (dom.nny - 1)
    weight = pkg.species_weight(dom, 4)
    pl.set_species(0, pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], weight, 2 * pkg.EVTOK, nel * 8)
    pl.set_species(1, pkg.ARGON_ION["charge_C"], pkg.ARGON_ION["mass_kg"], weight, 0.026 * pkg.EVTOK, nel * 8)
    pl.load_maxwellian(0, 4, 2.0, 7)
    pl.load_maxwellian(1, 4, 0.026, 8)
    pl.set_run_params(0, 7, 0)
    pl.initial_field()
    pl.step(2)
    return dom


@pytest.mark.gpu
@pytest.mark.parametrize("case,bits", [("small24", 64), ("small24", 32), ("rz24", 64)])
def test_bundle_dump_and_load(gpu_lib, pkg, tmp_path, case, bits):
    """DumpParticleBundle writes X in metres, Y and Z as kept (Z = the azimuth on r-z, 0
    otherwise), velocities in m/s, the species' charge, mass and weight, and Location; the load
    restores X * (1/L_ref) and v * (1/v_ref) exactly as the reference's PosRes/VelRes do."""
    pl = gpu_lib
    print(f"case: bundle {case} bits={bits}", flush=True)
    dom = _stepped(pl, pkg, case, bits)
    norm = dom.normalisation()
    L, vref = norm["L_ref"], norm["v_ref"]
    p = pl.particles(0)
    th = pl.azimuth(0) if dom.delta == 1 else np.zeros(len(p["x"]))
    f = tmp_path / "10000e.dat"
    pl.dump_bundle(0, f)
    npar, nnorm, phys, parts, loc = parse_bundle(f)
    assert npar == len(p["x"]) and phys == (pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"],
                                            pkg.species_weight(dom, 4))
    np.testing.assert_array_equal(parts[:, 0], p["x"] * L)
    np.testing.assert_array_equal(parts[:, 1], p["y"])
    np.testing.assert_array_equal(parts[:, 2], th)
    for k, name in ((3, "vx"), (4, "vy"), (5, "vz")):
        np.testing.assert_array_equal(parts[:, k], p[name] * vref)
    assert np.all(parts[:, 6:] == 0) and np.array_equal(loc, p["loc"])
    # reload: the loader's arithmetic on the file's values, X * (1/L_ref) and v * (1/v_ref)
    assert pl.load_bundle_file(0, f) == npar
    q = pl.particles(0)
    ft = np.float64 if bits == 64 else np.float32
    as_stored = lambda a: a.astype(ft).astype(np.float64)
    np.testing.assert_array_equal(q["x"], as_stored(parts[:, 0] * (1.0 / L)))
    np.testing.assert_array_equal(q["y"], as_stored(parts[:, 1]))
    for k, name in ((3, "vx"), (4, "vy"), (5, "vz")):
        np.testing.assert_array_equal(q[name], as_stored(parts[:, k] * (1.0 / vref)))
    assert np.array_equal(q["loc"], loc)
    if dom.delta == 1:
        np.testing.assert_array_equal(pl.azimuth(0), as_stored(parts[:, 2]))
    # a bundle of another species is refused
    with pytest.raises(pkg.PicError):
        pl.load_bundle_file(1, f)


@pytest.mark.gpu
def test_bundle_from_flang_fixture(gpu_lib, pkg):
    """The flang-written bundle loads into a species with its charge, mass and weight."""
    pl = gpu_lib
    print("case: flang bundle fixture", flush=True)
    dom = pkg.workload("small24")["domain"]
    pl.setup(dom, bits=64, solver=0, tol=1e-10)
    pl.set_species(0, -1.60217663e-19, 9.1093837e-31, 2.5e6, 2 * pkg.EVTOK, 100)
    assert pl.load_bundle_file(0, GOLDEN / "bundle_flang.dat") == 3
    q = pl.particles(0)
    norm = dom.normalisation()
    i = np.arange(1, 4)
    np.testing.assert_allclose(q["x"], 1.0e-6 * i / norm["L_ref"], rtol=1e-12)
    np.testing.assert_array_equal(q["y"], 0.25 * i)
    np.testing.assert_allclose(q["vy"], -2.0e5 * i / norm["v_ref"], rtol=1e-12)
    # the Locations written are the elements holding the positions (the load checks them)
    np.testing.assert_array_equal(q["loc"], 96 * i + 1)
    np.testing.assert_array_equal(pl.locate(q["x"], q["y"]), q["loc"])


@pytest.mark.gpu
def test_restart_dump_round_trip(gpu_lib, pkg, tmp_path):
    """Dump_2D writes var<lap>dump (counts, mesh size) and phi<lap>dump (one record of Phi);
    Restart_2D reads Phi back bit for bit."""
    pl = gpu_lib
    print("case: restart dump", flush=True)
    dom = _stepped(pl, pkg, "small24", 64)
    phi = pl.field("PHI")
    pl.dump_2d(12, tmp_path)
    (rec,) = read_records(tmp_path / "phi0000012dump")
    np.testing.assert_array_equal(np.frombuffer(rec, dtype="<f8"), phi)
    words = (tmp_path / "var0000012dump").read_text().split()
    c0, c1 = pl.counts(0)["nlive"], pl.counts(1)["nlive"]
    assert [int(x) for x in words[:6]] == [12, c0 + c1, c0, c1, dom.nnx, dom.nny]
    pl.set_field("PHI", np.zeros_like(phi))
    pl.restart_2d(12, tmp_path)
    assert np.array_equal(pl.field("PHI"), phi)


@pytest.mark.gpu
def test_tecplot_field_file(gpu_lib, pkg, tmp_path):
    """field_IJ_<it>.dat: the header lines, one FORMAT 50 record per node (x, y, R, Phi, Rho,
    Rho_s of three species, efx, efy, Ek_s and Ek_tot of two, node_type) and one FORMAT 44
    record per element (HT 1..4)."""
    pl = gpu_lib
    print("case: tecplot field", flush=True)
    dom = _stepped(pl, pkg, "small24", 64)
    pl.tecplot_field(3, tmp_path)
    lines = (tmp_path / "field_IJ_000003.dat").read_text().split("\n")
    nn, ne = dom.nnx * dom.nny, (dom.nnx - 1) * (dom.nny - 1)
    assert lines[0] == ' TITLE = "Field Plot"'
    assert len(lines[1]) == 150 and lines[1].startswith('VARIABLES = "x" "y" "R" "Phi"')
    assert lines[2] == " ZONE N=" + f"{nn:10d}".ljust(50) + ", E=" + f"{ne:10d}".ljust(50) + ",F=FEPOINT,ET=QUADRILATERAL"
    node = lines[3:3 + nn]
    # column starts of FORMAT 50: 16 per E15.6 field plus a blank, none after fields 7 and 10
    starts, o = [], 0
    for k in range(14):
        starts.append(o)
        o += 15 + (0 if k in (6, 9) or k == 13 else 1)
    assert all(len(l) == o + 1 + 15 + 1 + 15 for l in node)

    def fortran_float(t):
        t = t.strip()
        if "E" not in t:                       # 0.dddddd+eee: the exponent without its E
            k = max(t.rfind("+"), t.rfind("-"))
            t = t[:k] + "E" + t[k:]
        return float(t)

    def col(k):
        return np.array([fortran_float(l[starts[k]:starts[k] + 15]) for l in node])
    hp = pl.setup_array("HP")
    np.testing.assert_allclose(col(0), hp[0::2], rtol=5.01e-6, atol=1e-30)
    np.testing.assert_allclose(col(3), pl.field("PHI"), rtol=5.01e-6, atol=1e-30)
    np.testing.assert_allclose(col(4), pl.field("RHO"), rtol=5.01e-6, atol=1e-30)
    np.testing.assert_allclose(col(8), pl.field("EFX"), rtol=5.01e-6, atol=1e-30)
    eks = pl.field("EK_S")
    np.testing.assert_allclose(col(10), eks[:nn], rtol=5.01e-6, atol=1e-30)
    # Ek_s is Ek_tot / rho before normalisation: the mean particle energy in eV, positive
    # wherever particles deposited
    rho_p = pl.field("RHO_P")[:nn]
    assert np.all(eks[:nn][rho_p > 0] > 0) and np.all(eks[:nn][rho_p == 0] == 0)
    nt = pl.setup_array("NODE_TYPE").reshape(-1, 2)
    assert [int(l[o + 1:o + 16]) for l in node[:5]] == list(nt[:5, 0])
    elem = lines[3 + nn:3 + nn + ne]
    first = [int(elem[0][k * 15:(k + 1) * 15]) for k in range(4)]
    assert first == [1, 1 + dom.nny, 2 + dom.nny, 2]
