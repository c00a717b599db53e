"""Readers of the two VTU frame formats the tests compare: the library's
binary appended files (raw encoding, UInt64 block headers) and the ASCII
files the reference's writers print (core/solverBase.cpp:69-333, restated
by the oracle's vtu_ascii.cpp)."""
import re
import xml.etree.ElementTree as ET

import numpy as np

_DT = {"Float32": np.float32, "Int32": np.int32, "UInt8": np.uint8}


def read_vtu(path):
    """A binary appended VTU file: FieldData, NumberOfPoints and every
    DataArray by name ("points" for the points)."""
    raw = open(path, "rb").read()
    tag = b'<AppendedData encoding="raw">'
    at = raw.index(tag)
    head = raw[:at].decode()
    ET.fromstring(head + "</VTKFile>")  # the header is well-formed XML
    base = raw.index(b"_", at + len(tag)) + 1
    assert raw.rstrip().endswith(b"</VTKFile>")

    def block(off, dtype, comps):
        n = int(np.frombuffer(raw, np.uint64, 1, base + off)[0])
        a = np.frombuffer(raw, dtype, n // np.dtype(dtype).itemsize, base + off + 8)
        return a.reshape(-1, comps) if comps > 1 else a

    out = {"n": int(re.search(r'NumberOfPoints="(\d+)"', head).group(1))}
    for m in re.finditer(r"<DataArray ([^>]*)/>", head):
        attrs = dict(re.findall(r'(\w+)="([^"]*)"', m.group(1)))
        comps = int(attrs.get("NumberOfComponents", 1))
        name = attrs.get("Name", "points")
        out[name] = block(int(attrs["offset"]), _DT[attrs["type"]], comps)
    return out


def read_vtu_ascii(path):
    """An ASCII VTU file: the same dictionary with every value as printed
    (float64 for Float32 arrays, int64 for integer ones), plus "tokens": the
    text of every Float32 value and "order": the DataArray names in file
    order."""
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {"n": int(piece.get("NumberOfPoints")), "tokens": [], "order": []}
    for da in root.iter("DataArray"):
        name = da.get("Name", "points")
        words = (da.text or "").split()
        if da.get("type") == "Float32":
            out["tokens"] += words
            a = np.array([float(w) for w in words], np.float64)
        else:
            a = np.array([int(w) for w in words], np.int64)
        comps = int(da.get("NumberOfComponents", 1))
        out[name] = a.reshape(-1, comps) if comps > 1 else a
        out["order"].append(name)
    return out
