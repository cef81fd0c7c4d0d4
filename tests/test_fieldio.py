"""Field files of the reference (int32 count + FP64 values of the type's index
range, column-major) and restart folders, without a GPU.  The expected bytes
are built here independently with numpy from the format's definition
(field.f90:419-476, setIndexesBounds field.f90:1746-1789)."""
import os

import numpy as np
import pytest

from tbfsolver_amd import fieldio as IO
from tbfsolver_amd._lib import TbfgError

N = (6, 4, 5)
rng = np.random.default_rng(5)


def expected_bytes(f_fortran):
    """int32 size, then the array in Fortran (column-major) order."""
    return np.int32(f_fortran.size).tobytes() + np.asfortranarray(f_fortran).tobytes(order="F")


def test_counts():
    nx, ny, nz = N
    assert IO.field_count("cl", N) == nx * ny * nz
    assert IO.field_count("sx", N) == (nx + 1) * ny * nz
    assert IO.field_count("sy", N) == nx * (ny + 1) * nz
    assert IO.field_count("sz", N) == nx * ny * (nz + 1)


def test_cell_field_bytes_and_round_trip(tmp_path):
    nx, ny, nz = N
    eng = rng.standard_normal((nz, ny, nx))
    p = str(tmp_path / "psi")
    IO.write_field(p, eng, "cl", N)
    fort = eng.transpose(2, 1, 0)  # f(i,j,k)
    assert open(p, "rb").read() == expected_bytes(fort)
    assert np.array_equal(IO.read_field(p, "cl", N), eng)


def test_staggered_x_holds_face_zero_as_periodic_image(tmp_path):
    nx, ny, nz = N
    eng = rng.standard_normal((nz, ny, nx))  # faces 1..nx
    p = str(tmp_path / "ux")
    IO.write_field(p, eng, "sx", N)
    fort = np.empty((nx + 1, ny, nz))
    fort[1:] = eng.transpose(2, 1, 0)
    fort[0] = fort[nx]  # u(0) = u(nx): periodic x
    assert open(p, "rb").read() == expected_bytes(fort)
    assert np.array_equal(IO.read_field(p, "sx", N), eng)


def test_staggered_y_and_z(tmp_path):
    nx, ny, nz = N
    uy = rng.standard_normal((nz, ny + 1, nx))  # faces 0..ny
    uz = rng.standard_normal((nz, ny, nx))      # faces 1..nz
    IO.write_field(str(tmp_path / "uy"), uy, "sy", N)
    IO.write_field(str(tmp_path / "uz"), uz, "sz", N)
    assert open(tmp_path / "uy", "rb").read() == expected_bytes(uy.transpose(2, 1, 0))
    fz = np.empty((nx, ny, nz + 1))
    fz[:, :, 1:] = uz.transpose(2, 1, 0)
    fz[:, :, 0] = fz[:, :, nz]
    assert open(tmp_path / "uz", "rb").read() == expected_bytes(fz)
    assert np.array_equal(IO.read_field(str(tmp_path / "uy"), "sy", N), uy)
    assert np.array_equal(IO.read_field(str(tmp_path / "uz"), "sz", N), uz)


def test_vector_field_names(tmp_path):
    nx, ny, nz = N
    u = (rng.standard_normal((nz, ny, nx)), rng.standard_normal((nz, ny + 1, nx)), rng.standard_normal((nz, ny, nx)))
    IO.write_vfield(str(tmp_path), "u", *u, N)
    assert sorted(os.listdir(tmp_path)) == ["ux", "uy", "uz"]
    back = IO.read_vfield(str(tmp_path), "u", N)
    assert all(np.array_equal(a, b) for a, b in zip(u, back))


def test_reading_a_file_of_another_grid_fails(tmp_path):
    nx, ny, nz = N
    p = str(tmp_path / "p")
    IO.write_field(p, np.zeros((nz, ny, nx)), "cl", N)
    with pytest.raises(TbfgError, match="not equal to mesh size"):
        IO.read_field(p, "cl", (nx, ny, nz + 1))
    with pytest.raises(TbfgError, match="TBFG_EIO"):
        IO.read_field(str(tmp_path / "absent"), "cl", N)
    with open(p, "r+b") as f:  # truncated values
        f.truncate(4 + 8 * 10)
    with pytest.raises(TbfgError, match="TBFG_EIO"):
        IO.read_field(p, "cl", N)
    with pytest.raises(ValueError):
        IO.write_field(p, np.zeros((nz, ny + 1, nx)), "cl", N)


def test_time_folder_and_info_restart(tmp_path):
    assert IO.folder_name(0) == "0" and IO.folder_name(12) == "12" and IO.folder_name(123456) == "123456"
    # ES25.15E3 and I6 as the Fortran formats print them
    assert IO._es25_15e3(0.0) == "    0.000000000000000E+000"[-25:]
    assert IO._es25_15e3(1.5) == "    1.500000000000000E+000"[-25:]
    assert IO._es25_15e3(-2.5e-7) == "   -2.500000000000000E-007"[-25:]
    IO.write_info_restart(str(tmp_path), 0.125, 18)
    assert open(tmp_path / "info_restart").read() == "    1.250000000000000E-001"[-25:] + "\n    18\n"
    assert IO.read_info_restart(str(tmp_path)) == (0.125, 18)
