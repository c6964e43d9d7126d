"""Frame output of the device library (host code, no GPU needed): the reference's
writers and reader (src/output.rs:23-122) and mesh writers (src/mesh.rs:356-456).

Pins: the reference's own tests (output.rs:132-184: VTK structure, bit-exact round
trip, XYZ byte layout, zero-padded frame names), the full VTK text of a two-point
frame, and the coordinate text against numpy's Dragon4 shortest positional formatting
(an independent implementation of the same rule as Rust's f64 Display)."""
import struct

import numpy as np
import pytest

import sdr_amd
from sdr_amd import _abi as A
from sdr_amd import output
from sdr_amd.surface import TriMesh


@pytest.fixture(scope="module")
def lib():
    return sdr_amd.load_library()


def test_vtk_has_expected_structure(lib, tmp_path):  # output.rs:132-143
    pts = np.array([[0.0, 1.0, 2.0], [-1.0, 0.5, 3.0]])
    p = tmp_path / "pts.vtk"
    output.write_points_vtk(p, pts, lib=lib)
    text = p.read_text()
    assert "DATASET UNSTRUCTURED_GRID" in text
    assert "POINTS 2 double" in text
    assert "0 1 2" in text
    assert "CELL_TYPES 2" in text
    assert text == ("# vtk DataFile Version 4.2\nsdr irrigation particles\nASCII\nDATASET UNSTRUCTURED_GRID\n"
                    "POINTS 2 double\n0 1 2\n-1 0.5 3\nCELLS 2 4\n1 0\n1 1\nCELL_TYPES 2\n1\n1\n")


def test_vtk_round_trips_positions(lib, tmp_path):  # output.rs:146-163
    pts = np.array([[0.0, 1.0, 2.0], [-1.5, 0.25, 3.5], [10.0, -20.0, 30.0]])
    p = tmp_path / "rt.vtk"
    output.write_points_vtk(p, pts, lib=lib)
    back = output.read_points_vtk(p, lib=lib)
    assert back.shape == pts.shape
    assert np.array_equal(back.view(np.uint64), pts.view(np.uint64))


def test_vtk_round_trip_is_bit_exact_over_magnitudes(lib, tmp_path):
    rng = np.random.default_rng(7)
    mant = rng.uniform(-1, 1, 3000)
    expo = rng.integers(-320, 300, 3000).astype(np.float64)
    vals = mant * 10.0 ** expo
    vals = np.concatenate([vals, [0.0, -0.0, 5e-324, -5e-324, 1.7976931348623157e308, 2.2250738585072014e-308,
                                  0.1, 1.0 / 3.0, 123456789.0]])
    vals = vals[: len(vals) // 3 * 3].reshape(-1, 3)
    p = tmp_path / "mag.vtk"
    output.write_points_vtk(p, vals, lib=lib)
    back = output.read_points_vtk(p, lib=lib)
    assert np.array_equal(back.view(np.uint64), vals.view(np.uint64))
    lines = p.read_text().splitlines()[5:5 + vals.shape[0]]
    for row, line in zip(vals, lines):
        want = " ".join(np.format_float_positional(x, unique=True, trim="-") for x in row)
        assert line == want


def test_vtk_positions_of_a_realistic_frame(lib, tmp_path):
    """A frame of the size the writer chunks over threads (> 2^18 points)."""
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.02, 0.03, (300_001, 3))
    p = tmp_path / "big.vtk"
    output.write_points_vtk(p, pts, lib=lib)
    back = output.read_points_vtk(p, lib=lib)
    assert np.array_equal(back.view(np.uint64), pts.view(np.uint64))
    text = p.read_text()
    assert text.count("\n") == 5 + pts.shape[0] + 1 + pts.shape[0] + 1 + pts.shape[0]


def test_xyz_is_binary_f32_triples(lib, tmp_path):  # output.rs:166-177
    pts = np.array([[0.0, 1.0, 2.0], [3.0, 4.0, 5.0]])
    p = tmp_path / "pts.xyz"
    output.write_points_xyz(p, pts, lib=lib)
    b = p.read_bytes()
    assert len(b) == 2 * 12
    assert struct.unpack("<f", b[0:4])[0] == 0.0
    assert struct.unpack("<f", b[16:20])[0] == 4.0
    rng = np.random.default_rng(1)
    pts = rng.normal(size=(1000, 3)) * 0.01
    output.write_points_xyz(p, pts, lib=lib)
    back = np.frombuffer(p.read_bytes(), "<f4").reshape(-1, 3)
    assert np.array_equal(back, pts.astype(np.float32))  # `as f32`: round to nearest even


def test_frame_path_is_zero_padded(lib):  # output.rs:180-184
    assert output.frame_path("/tmp/out", "particles", 42, "vtk", lib=lib).endswith("particles_000042.vtk")
    assert output.frame_path("/tmp/out/", "slice", 1234567, "png", lib=lib) == "/tmp/out/slice_1234567.png"
    assert output.frame_path("", "p", 0, "xyz", lib=lib) == "p_000000.xyz"


def test_read_points_vtk_errors(lib, tmp_path):
    with pytest.raises(A.AbiError) as e:
        output.read_points_vtk(tmp_path / "missing.vtk", lib=lib)
    assert e.value.code == A.SDR_E_IO
    (tmp_path / "a.vtk").write_text("# vtk\nno points here\n")
    with pytest.raises(A.AbiError) as e:
        output.read_points_vtk(tmp_path / "a.vtk", lib=lib)
    assert e.value.code == A.SDR_E_INVALID and "no POINTS block" in str(e.value)
    (tmp_path / "b.vtk").write_text("POINTS 2 double\n1 2 3\n4 5\n")
    with pytest.raises(A.AbiError) as e:
        output.read_points_vtk(tmp_path / "b.vtk", lib=lib)
    assert "truncated" in str(e.value)
    (tmp_path / "c.vtk").write_text("POINTS x double\n")
    with pytest.raises(A.AbiError) as e:
        output.read_points_vtk(tmp_path / "c.vtk", lib=lib)
    assert "malformed" in str(e.value)
    # any whitespace layout reads back (output.rs:82-84), signs and exponents parse
    (tmp_path / "d.vtk").write_text("junk\n  POINTS 2 float\n+1 -2.5e-3\n\t3\n4 5 6 CELLS\n")
    assert np.array_equal(output.read_points_vtk(tmp_path / "d.vtk", lib=lib),
                          np.array([[1.0, -2.5e-3, 3.0], [4.0, 5.0, 6.0]]))


def _unit_box():  # mesh.rs:471-500: side-2 cube centred at the origin, outward winding
    v = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    t = np.array([[0, 3, 2], [0, 2, 1], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [3, 7, 6], [3, 6, 2], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.uint32)
    return TriMesh(v, t)


def test_box_area_and_volume():  # mesh.rs:502-508
    m = _unit_box()
    assert m.surface_area() == pytest.approx(24.0, abs=1e-9)
    assert m.signed_volume() == pytest.approx(8.0, abs=1e-9)


def test_obj_round_trip(lib, tmp_path):  # mesh.rs:510-521, read back with a plain parser
    m = _unit_box()
    p = tmp_path / "sub" / "box.obj"  # save_obj creates the parent directory
    m.save_obj(p, lib=lib)
    lines = p.read_text().splitlines()
    assert lines[0] == "# sdr parametric sinus mesh"
    assert lines[1] == "# vertices: 8 faces: 12"
    assert lines[2] == "v -1 -1 -1" and lines[10] == "f 1 4 3" and len(lines) == 22
    v = np.array([[float(x) for x in ln.split()[1:]] for ln in lines if ln.startswith("v ")])
    f = np.array([[int(x) - 1 for x in ln.split()[1:]] for ln in lines if ln.startswith("f ")], np.uint32)
    m2 = TriMesh(v, f)
    assert len(m2.vertices) == 8 and m2.triangle_count() == 12
    assert m2.signed_volume() == pytest.approx(8.0, abs=1e-6)


def test_stl_round_trip(lib, tmp_path):  # mesh.rs:523-533, read back with struct
    m = _unit_box()
    s = tmp_path / "box.stl"
    m.save_stl(s, lib=lib)
    b = s.read_bytes()
    assert len(b) == 84 + 50 * 12 and b[:80] == bytes(80)
    assert struct.unpack("<I", b[80:84])[0] == 12
    recs = [struct.unpack("<12fH", b[84 + 50 * q:134 + 50 * q]) for q in range(12)]
    assert recs[0][:3] == (0.0, 0.0, -1.0) and recs[0][12] == 0  # -z face, unit normal
    v = np.array([r[3:12] for r in recs], np.float64).reshape(-1, 3)
    m2 = TriMesh(v, np.arange(36, dtype=np.uint32).reshape(12, 3))
    assert m2.triangle_count() == 12
    assert m2.surface_area() == pytest.approx(24.0, abs=1e-4)
    with pytest.raises(A.AbiError):
        TriMesh(m.vertices, np.array([[0, 1, 99]], np.uint32)).save_obj(tmp_path / "bad.obj", lib=lib)
