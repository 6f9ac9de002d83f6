"""Solution output: legacy-VTK structured grids of the solution points and state files in the
reference layout (checkpoint / resume of the Julia workflow's `u` arrays)."""
from __future__ import annotations

import numpy as np


def write_vtk(path: str, x, y, fields: dict, title: str = "frhip solution points") -> None:
    """Binary legacy VTK STRUCTURED_GRID of the point grid (x, y) (2D arrays of one shape, first
    index fastest in the file) with one scalar point field per entry of `fields` (arrays of the
    same shape).  `FREuler2D.primitives()` gives the grid and the fields of
    example/shock-vortex.jl:333-352."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if x.ndim != 2 or x.shape != y.shape:
        raise ValueError("write_vtk: x and y must be 2D arrays of the same shape")
    n0, n1 = x.shape
    npts = n0 * n1
    pts = np.zeros((npts, 3), dtype=">f8")
    pts[:, 0] = x.ravel(order="F")
    pts[:, 1] = y.ravel(order="F")
    title = title.replace("\n", " ")[:255]
    with open(path, "wb") as f:
        f.write(f"# vtk DataFile Version 3.0\n{title}\nBINARY\nDATASET STRUCTURED_GRID\n".encode())
        f.write(f"DIMENSIONS {n0} {n1} 1\nPOINTS {npts} double\n".encode())
        f.write(pts.tobytes())
        f.write(f"\nPOINT_DATA {npts}\n".encode())
        for name, v in fields.items():
            v = np.asarray(v, dtype=np.float64)
            if v.shape != x.shape:
                raise ValueError(f"write_vtk: field {name!r} has shape {v.shape}, grid is {x.shape}")
            if not name or any(c.isspace() for c in name):
                raise ValueError(f"write_vtk: field name {name!r} must be one word")
            f.write(f"SCALARS {name} double 1\nLOOKUP_TABLE default\n".encode())
            f.write(v.ravel(order="F").astype(">f8").tobytes())
            f.write(b"\n")


def read_vtk(path: str):
    """Reader of the files write_vtk makes: returns (x, y, fields)."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    if not line().startswith("# vtk DataFile"):
        raise ValueError("read_vtk: not a legacy VTK file")
    line()
    if line().strip() != "BINARY" or line().strip() != "DATASET STRUCTURED_GRID":
        raise ValueError("read_vtk: expected a binary structured grid")
    n0, n1, _ = (int(t) for t in line().split()[1:4])
    npts = int(line().split()[1])
    pts = np.frombuffer(data, dtype=">f8", count=3 * npts, offset=pos).reshape(npts, 3)
    pos += 24 * npts + 1
    x = pts[:, 0].reshape((n0, n1), order="F").astype(np.float64)
    y = pts[:, 1].reshape((n0, n1), order="F").astype(np.float64)
    line()  # POINT_DATA
    fields = {}
    while pos < len(data):
        head = line().split()
        if not head:
            continue
        name = head[1]
        line()  # LOOKUP_TABLE
        v = np.frombuffer(data, dtype=">f8", count=npts, offset=pos)
        pos += 8 * npts + 1
        fields[name] = v.reshape((n0, n1), order="F").astype(np.float64)
    return x, y, fields


def save_state(path: str, u, **meta) -> None:
    """State array in the reference layout (Julia column-major, ghost layer included) plus scalar
    metadata (numbers or strings: nx, ny, deg, gamma, t, ...) in an .npz file without pickles."""
    arrays = {"u": np.asfortranarray(np.asarray(u, dtype=np.float64))}
    for k, v in meta.items():
        if k == "u":
            raise ValueError("save_state: 'u' is reserved")
        arr = np.asarray(v)
        if arr.dtype == object:
            raise ValueError(f"save_state: metadata {k!r} must be a number or a string")
        arrays[k] = arr
    np.savez(path, **arrays)


def load_state(path: str):
    """Inverse of save_state: (u, meta) with u Fortran-ordered float64."""
    with np.load(path, allow_pickle=False) as z:
        u = np.asfortranarray(z["u"])
        meta = {k: z[k].item() for k in z.files if k != "u"}
    return u, meta
