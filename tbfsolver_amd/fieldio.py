"""The reference's on-disk fields and restart folders, so an instance can be
loaded from and saved to a TBFsolver case directory.

  write_field / read_field        one field file (field.f90:331-346,419-476), through libtbfgpu
  write_vfield / read_vfield      a vector field as its three components <name>x, <name>y, <name>z
                                  with types sx, sy, sz (vfield.f90:67-81)
  folder_name(n)                  the time folder of output n (time.f90:317, format I6 adjusted left)
  write_info_restart / read_...   <folder>/info_restart: time (ES25.15E3) and bubble count (I6)
                                  (time.f90:327-330, 424-426)

Arrays are in the engine layout: cell fields (nz, ny, nx), C order (the memory
order of the reference's f(i,j,k), x fastest); x and z faces hold faces 1..n
(nz, ny, nx); y faces hold 0..ny (nz, ny+1, nx).  torch tensors (host or
device) and numpy arrays are accepted; reads return numpy arrays.
"""
import ctypes as C
import os

import numpy as np

from ._lib import check, lib

CL, SX, SY, SZ = 0, 1, 2, 3
_TYPES = {"cl": CL, "sx": SX, "sy": SY, "sz": SZ}


def _type(t):
    return _TYPES[t] if isinstance(t, str) else int(t)


def _n3(n):
    return (C.c_int32 * 3)(*[int(v) for v in n])


def engine_shape(ftype, n):
    nx, ny, nz = n
    return (nz, ny + 1, nx) if _type(ftype) == SY else (nz, ny, nx)


def field_count(ftype, n):
    out = C.c_int64(0)
    check(lib().tbfg_field_count(_type(ftype), _n3(n), C.byref(out)))
    return int(out.value)


def _host(a):
    if hasattr(a, "detach"):  # torch tensor
        a = a.detach().to("cpu").numpy()
    return np.ascontiguousarray(a, dtype=np.float64)


def write_field(path, a, ftype, n):
    """writeField: the int32 count, then the FP64 values of the type's index range."""
    h = _host(a)
    if h.shape != engine_shape(ftype, n):
        raise ValueError("array of shape %s for a %s field of %s" % (h.shape, ftype, tuple(n)))
    check(lib().tbfg_field_write(os.fsencode(path), _type(ftype), _n3(n),
                                 h.ctypes.data_as(C.POINTER(C.c_double))))


def read_field(path, ftype, n):
    """readField: the engine array of the file (a count other than the grid's
    raises, as the reference aborts)."""
    out = np.empty(engine_shape(ftype, n), dtype=np.float64)
    check(lib().tbfg_field_read(os.fsencode(path), _type(ftype), _n3(n), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def write_vfield(folder, name, ux, uy, uz, n):
    for comp, a, t in (("x", ux, SX), ("y", uy, SY), ("z", uz, SZ)):
        write_field(os.path.join(folder, name + comp), a, t, n)


def read_vfield(folder, name, n):
    return tuple(read_field(os.path.join(folder, name + comp), t, n) for comp, t in (("x", SX), ("y", SY), ("z", SZ)))


def folder_name(n):
    buf = C.create_string_buffer(16)
    check(lib().tbfg_time_folder_name(int(n), buf, 16))
    return buf.value.decode()


def _es25_15e3(x):
    """Fortran ES25.15E3: d.ddddddddddddddd E+eee, right-adjusted in 25 columns."""
    m, e = ("%.15E" % float(x)).split("E")
    return ("%sE%s%03d" % (m, "-" if int(e) < 0 else "+", abs(int(e)))).rjust(25)


def write_info_restart(folder, t, nb):
    with open(os.path.join(folder, "info_restart"), "w") as f:
        f.write(_es25_15e3(t) + "\n")
        f.write("%6d\n" % int(nb))


def read_info_restart(folder):
    with open(os.path.join(folder, "info_restart")) as f:
        lines = f.read().split("\n")
    return float(lines[0].replace("E", "e")), int(lines[1])
