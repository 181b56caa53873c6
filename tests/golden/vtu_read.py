"""Minimal reader for VTK XML UnstructuredGrid files with appended, base64,
zlib-compressed arrays (the layout NLMech's rw::writer::VtkWriter produces,
src/rw/vtkWriter.cpp:20-298).  Used by the fixture scripts and the tests.

Appended block layout (VTK vtkZLibDataCompressor, header_type UInt32): at each
array's offset the base64 text holds a header [nblocks, block_size,
last_block_size, csize_1..csize_n] (UInt32) encoded on its own, followed by the
base64 of the concatenated zlib blocks.
"""
from __future__ import annotations

import base64
import re
import zlib

import numpy as np

_DT = {"Float64": np.float64, "Float32": np.float32, "Int32": np.int32,
       "Int64": np.int64, "UInt8": np.uint8, "UInt32": np.uint32,
       "UInt64": np.uint64}


def _b64len(nbytes: int) -> int:
    return (nbytes + 2) // 3 * 4


def read_vtu(path: str) -> dict:
    text = open(path, "rb").read().decode("latin-1")
    m = re.search(r'<AppendedData encoding="base64">\s*_', text)
    if not m:
        raise ValueError("only base64 appended data is supported")
    data = text[m.end():]
    end = data.find("</AppendedData>")
    data = re.sub(r"\s+", "", data[:end])
    compressed = "vtkZLibDataCompressor" in text
    out = {}
    for a in re.finditer(r"<DataArray([^>]*)>", text):
        attrs = dict(re.findall(r'(\w+)="([^"]*)"', a.group(1)))
        if attrs.get("format") != "appended":
            continue
        name = attrs.get("Name")
        dt = np.dtype(_DT[attrs["type"]])
        ncomp = int(attrs.get("NumberOfComponents", "1"))
        off = int(attrs["offset"].strip())
        s = data[off:]
        if compressed:
            first = base64.b64decode(s[:8])
            nb = int(np.frombuffer(first[:4], np.uint32)[0])
            hlen = _b64len(4 * (3 + nb))
            hdr = np.frombuffer(base64.b64decode(s[:hlen]), np.uint32)
            csizes = hdr[3:3 + nb]
            total = int(csizes.sum())
            raw = base64.b64decode(s[hlen:hlen + _b64len(total)])
            buf = b""
            pos = 0
            for cs in csizes:
                buf += zlib.decompress(raw[pos:pos + int(cs)])
                pos += int(cs)
        else:
            first = base64.b64decode(s[:8])
            nbytes = int(np.frombuffer(first[:4], np.uint32)[0])
            buf = base64.b64decode(s[:_b64len(4 + nbytes)])[4:4 + nbytes]
        arr = np.frombuffer(buf, dt).copy()
        if ncomp > 1:
            arr = arr.reshape(-1, ncomp)
        out[name] = arr
    return out
