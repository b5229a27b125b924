"""The reference's on-disk formats (no GPU): field files of src/mod_io.f90,
calc_bands.py coefficient text files read by read_column_r8, and the yu.txt
faces of the reference's four example cases (tests/golden/)."""
import json
import os
import struct

import numpy as np
import pytest

import oracle as O
import dnsproj as D
from dnsproj import io as IO

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_field_file_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((5, 3, 7))  # (ny, nz, nx)
    p = tmp_path / "a.dat"
    IO.write_field(p, A)
    raw = p.read_bytes()
    # the format note of examples/*/array_ini_*.dat: 12-byte int32 header, then
    # FP64 array[i,j,k] with i fastest: (12 + 8*ni*nj*nk) bytes
    assert len(raw) == 12 + 8 * A.size
    assert struct.unpack("<3i", raw[:12]) == (7, 3, 5)
    assert np.array_equal(np.frombuffer(raw[12:], dtype="<f8"), A.ravel())
    assert IO.field_shape(p) == (7, 3, 5)
    assert np.array_equal(IO.read_field(p, 7, 3, 5), A)


def test_field_file_halo_interior_only(tmp_path):
    rng = np.random.default_rng(2)
    H = rng.standard_normal((6, 5, 9))  # halo array of (nx, nz, ny) = (7, 3, 4)
    p = tmp_path / "h.dat"
    IO.write_field(p, H, halo=True)
    assert IO.field_shape(p) == (7, 3, 4)
    assert np.array_equal(IO.read_field(p, 7, 3, 4), H[1:-1, 1:-1, 1:-1])
    back = IO.read_field(p, 7, 3, 4, halo=True)
    assert np.array_equal(back[1:-1, 1:-1, 1:-1], H[1:-1, 1:-1, 1:-1])
    assert np.all(back[0] == 0) and np.all(back[:, :, -1] == 0)


def test_field_file_errors(tmp_path):
    p = tmp_path / "a.dat"
    IO.write_field(p, np.zeros((2, 2, 2)))
    with pytest.raises(D.DnspError) as e:
        IO.read_field(p, 2, 2, 3)
    assert "Mismatch in grid params" in str(e.value)
    q = tmp_path / "short.dat"
    q.write_bytes(p.read_bytes()[:40])
    with pytest.raises(D.DnspError):
        IO.read_field(q, 2, 2, 2)
    with pytest.raises(D.DnspError):
        IO.field_shape(tmp_path / "missing.dat")


def python_writer(path, rows):
    """calc_bands.py:358-364: count line, then ' '.join(map(str, row))."""
    rows = [[x] if isinstance(x, (int, float)) else x for x in rows]
    with open(path, "w") as fw:
        fw.write(str(len(rows)) + "\n")
        for r in rows:
            fw.write(" ".join(map(str, r)) + "\n")


def test_read_column_follows_read_column_r8(tmp_path):
    rows = [[0.1 * i, 2.5e-05 * i, -3.0 + i, 7.0 / (i + 1)] for i in range(9)]
    p = tmp_path / "P_band_abc.dat"
    python_writer(p, rows)
    for icol in range(4):
        assert np.array_equal(IO.read_column(p, icol, 0, 9), np.array([r[icol] for r in rows]))
    assert np.array_equal(IO.read_column(p, 2, 3, 4), np.array([r[2] for r in rows[3:7]]))
    with pytest.raises(D.DnspError) as e:
        IO.read_column(p, 4, 0, 9)
    assert "icol" in str(e.value)
    with pytest.raises(D.DnspError):
        IO.read_column(p, 0, 5, 5)  # n_prev + n - 1 >= nb
    # Fortran exponents and a single column file
    q = tmp_path / "one.dat"
    q.write_text("3\n1.5d0\n-2.0D-3\n4\n")
    assert np.array_equal(IO.read_column(q, 0, 0, 3), np.array([1.5, -2e-3, 4.0]))


def test_write_table_round_trip_exact(tmp_path):
    rng = np.random.default_rng(3)
    rows = rng.standard_normal((11, 3)) * 10.0 ** rng.integers(-12, 12, size=(11, 3))
    p = tmp_path / "t.dat"
    IO.write_table(p, rows)
    for icol in range(3):
        assert np.array_equal(IO.read_column(p, icol, 0, 11), rows[:, icol])


def reference_grids():
    return json.load(open(os.path.join(GOLDEN, "reference_grids.json")))


@pytest.mark.parametrize("name", sorted(json.load(open(os.path.join(GOLDEN, "reference_grids.json")))))
def test_reference_example_faces(name):
    g = reference_grids()[name]
    path = os.path.join(GOLDEN, f"yu_{name}.txt")
    yu = IO.read_faces(path)
    assert len(yu) == g["ny"] + 1
    assert np.array_equal(yu, np.loadtxt(path))
    # the rough-wall cases extend below y = 0 and above Ly (the roughness region)
    assert np.all(np.diff(yu) > 0) and yu[-1] - yu[0] >= g["Ly"] * (1 - 1e-12)
    # the bands of these faces: the library's host routine equals the oracle's
    got = D.mk_P_abc(yu, False) + D.mk_P_abc(yu, True)
    ref = O.bands_from_yu(yu)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_faces_errors(tmp_path):
    p = tmp_path / "yu.txt"
    p.write_text("0\n1\n0.5\n")
    with pytest.raises(D.DnspError):
        IO.read_faces(p)
