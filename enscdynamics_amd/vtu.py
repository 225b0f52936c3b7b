"""VTK XML unstructured-grid files (.vtu) without VTK.

``write_vtu`` calls the library's writer (include/ensc_vtu.h), the
counterpart of DataOut::write_vtu (Explicit3D/source/dataOut.cpp:84-195).
``read_vtu`` parses the files that writer produces (appended raw arrays,
UInt64 headers) back into numpy arrays; it is a plain reader of that
layout, for checks and post-processing.
"""
from __future__ import annotations

import ctypes as C
import re
import xml.etree.ElementTree as ET
from typing import Dict, Mapping

import numpy as np

from . import engine as _eng


class VtuArray(C.Structure):
    _fields_ = [("name", C.c_char_p), ("n_comp", C.c_int32), ("pad_", C.c_int32),
                ("data", C.POINTER(C.c_double))]


def _arrays(d: Mapping[str, np.ndarray], n: int):
    keep, arr = [], (VtuArray * max(1, len(d)))()
    for i, (name, a) in enumerate(d.items()):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(n, -1))
        keep.append(a)
        arr[i].name = name.encode()
        arr[i].n_comp = a.shape[1]
        arr[i].data = a.ctypes.data_as(C.POINTER(C.c_double))
    return arr, keep


def write_vtu(path: str, points: np.ndarray, conn: np.ndarray, time: float,
              point_data: Mapping[str, np.ndarray] = {}, cell_data: Mapping[str, np.ndarray] = {}) -> None:
    """points (Nn, 3); conn (Ne, 8) 0-based; arrays (Nn|Ne, n_comp), n_comp in
    {1, 3, 6, 9}; written as Float32 like the reference's vtkFloatArray."""
    L = _eng.load_library()
    f = L.ensc_vtu_write
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_double,
                  C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    cn = np.ascontiguousarray(conn, dtype=np.uint32).reshape(-1, 8)
    pa, keep_p = _arrays(point_data, pts.shape[0])
    ca, keep_c = _arrays(cell_data, cn.shape[0])
    _eng._check(f(path.encode(), pts.shape[0], pts.ctypes.data, cn.shape[0], cn.ctypes.data, float(time),
                  C.cast(pa, C.c_void_p), len(point_data), C.cast(ca, C.c_void_p), len(cell_data)))


_TYPES = {"Float32": np.float32, "Float64": np.float64, "Int64": np.int64, "Int32": np.int32,
          "UInt8": np.uint8, "UInt32": np.uint32, "UInt64": np.uint64}


def read_vtu(path: str) -> Dict[str, object]:
    """Parse an appended-raw .vtu with UInt64 headers.  Returns a dict with
    'points', 'connectivity', 'offsets', 'types', 'time', 'point_data' and
    'cell_data' (name -> array of shape (n, n_comp))."""
    with open(path, "rb") as fh:
        blob = fh.read()
    m = re.search(rb'<AppendedData encoding="raw">\s*_', blob)
    if not m:
        raise ValueError("no appended raw section")
    raw0 = m.end()
    end = blob.rfind(b"</AppendedData>")
    xml = blob[:m.start()] + b"<AppendedData/></VTKFile>"
    root = ET.fromstring(xml)
    if root.get("type") != "UnstructuredGrid" or root.get("header_type") != "UInt64":
        raise ValueError("not an UnstructuredGrid with UInt64 headers")

    def arr(el) -> np.ndarray:
        off = raw0 + int(el.get("offset"))
        nbytes = int(np.frombuffer(blob, np.uint64, 1, off)[0])
        a = np.frombuffer(blob, _TYPES[el.get("type")], nbytes // np.dtype(_TYPES[el.get("type")]).itemsize,
                          off + 8)
        if off + 8 + nbytes > end:
            raise ValueError("array runs past the appended section")
        nc = int(el.get("NumberOfComponents", "1"))
        return a.reshape(-1, nc) if nc > 1 else a.copy()

    ug = root.find("UnstructuredGrid")
    piece = ug.find("Piece")
    out: Dict[str, object] = {
        "n_points": int(piece.get("NumberOfPoints")),
        "n_cells": int(piece.get("NumberOfCells")),
        "time": float(arr(ug.find("FieldData").find("DataArray"))[0]),
        "points": arr(piece.find("Points").find("DataArray")),
        "point_data": {e.get("Name"): arr(e) for e in piece.find("PointData")},
        "cell_data": {e.get("Name"): arr(e) for e in piece.find("CellData")},
    }
    for e in piece.find("Cells"):
        out[e.get("Name")] = arr(e)
    return out
