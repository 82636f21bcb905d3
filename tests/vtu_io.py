"""Minimal VTK XML UnstructuredGrid reader for the tests: ASCII arrays (the
reference's writer) and appended raw binary arrays with UInt64 headers
(pxg_write_vtu).  Returns the point-data arrays in file order, the points and
the cell arrays."""
import re

import numpy as np

_TYPES = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8, "Int32": np.int32}


def read_vtu(path):
    raw = open(path, "rb").read()
    k = raw.find(b"<AppendedData")
    head = raw if k < 0 else raw[:k]
    text = head.decode()
    data0 = None
    if k >= 0:
        data0 = raw.index(b"_", k) + 1
    def arrays(section):
        m = re.search(r"<%s[^>]*>(.*?)</%s>" % (section, section), text, re.S)
        out = []
        if not m:
            return out
        for a in re.finditer(r"<DataArray([^>]*?)(/>|>(.*?)</DataArray>)", m.group(1), re.S):
            attrs = dict(re.findall(r'(\w+)="([^"]*)"', a.group(1)))
            dt = _TYPES[attrs["type"]]
            nc = int(attrs.get("NumberOfComponents", "1"))
            if attrs.get("format") == "appended":
                o = data0 + int(attrs["offset"])
                nbytes = int(np.frombuffer(raw[o:o + 8], np.uint64)[0])
                v = np.frombuffer(raw[o + 8:o + 8 + nbytes], dt).copy()
            else:
                v = np.array(a.group(3).split(), dtype=dt)
            out.append((attrs.get("Name", ""), nc, v.reshape(-1, nc) if nc > 1 else v))
        return out
    piece = re.search(r'<Piece NumberOfPoints="(\d+)" NumberOfCells="(\d+)"', text)
    return {"npoints": int(piece.group(1)), "ncells": int(piece.group(2)),
            "point_data": arrays("PointData"), "points": arrays("Points")[0][2], "cells": arrays("Cells")}
