"""PCD point cloud files (PCL's format, version 0.7) for the Python side, as
``include/ikd_pcd.hpp`` is for C++: the reference's demos load their map with
pcl::io::loadPCDFile (examples/ikd_Tree_Search_demo.cpp:64-76); here

    xyz = pcd.load_xyz("map.pcd")
    tree = ikd.KD_TREE(0.3, 0.6, 0.2)
    tree.Build(xyz)

Header keys VERSION, FIELDS, SIZE, TYPE, COUNT, WIDTH, HEIGHT, VIEWPOINT,
POINTS, DATA ('#' lines are comments); DATA ascii (one point per line),
binary (points packed, fields in header order) or binary_compressed (uint32
compressed size, uint32 raw size, an LZF stream of the fields one after
another).  TYPE F/U/I with SIZE 1, 2, 4 or 8; COUNT values per point.
"""
from __future__ import annotations

import numpy as np

_DTYPES = {("F", 4): "<f4", ("F", 8): "<f8", ("U", 1): "<u1", ("U", 2): "<u2", ("U", 4): "<u4", ("U", 8): "<u8",
           ("I", 1): "<i1", ("I", 2): "<i2", ("I", 4): "<i4", ("I", 8): "<i8"}


def lzf_decompress(src: bytes, out_len: int) -> bytes:
    """LZF (liblzf's format): control byte c < 32 copies c + 1 literal bytes;
    otherwise a back reference of length (c >> 5) + 2 (c >> 5 == 7: the next
    byte adds to the length) at distance ((c & 31) << 8 | next byte) + 1."""
    out = bytearray()
    i, n = 0, len(src)
    while i < n:
        c = src[i]
        i += 1
        if c < 32:
            if i + c + 1 > n:
                raise ValueError("LZF literal run past the end of the stream")
            out += src[i:i + c + 1]
            i += c + 1
            continue
        length = c >> 5
        if length == 7:
            if i >= n:
                raise ValueError("LZF stream ends inside a back reference")
            length += src[i]
            i += 1
        length += 2
        if i >= n:
            raise ValueError("LZF stream ends inside a back reference")
        dist = (((c & 31) << 8) | src[i]) + 1
        i += 1
        if dist > len(out):
            raise ValueError("LZF back reference before the start of the output")
        for _ in range(length):  # may overlap the bytes it writes
            out.append(out[-dist])
    if len(out) != out_len:
        raise ValueError(f"LZF stream gave {len(out)} bytes, expected {out_len}")
    return bytes(out)


def lzf_literal_stream(raw: bytes) -> bytes:
    """A valid LZF stream of literal runs only (what the writer stores)."""
    out = bytearray()
    for i in range(0, len(raw), 32):
        chunk = raw[i:i + 32]
        out.append(len(chunk) - 1)
        out += chunk
    return bytes(out)


def _parse_header(data: bytes):
    h = {"count": None, "height": 1, "points": None}
    pos = 0
    while True:
        eol = data.find(b"\n", pos)
        if eol < 0:
            raise ValueError("PCD header ends without DATA")
        line = data[pos:eol].decode("ascii").strip()
        pos = eol + 1
        if not line or line.startswith("#"):
            continue
        key, *vals = line.split()
        key = key.upper()
        if key in ("VERSION", "VIEWPOINT"):
            continue
        if key in ("FIELDS", "COLUMNS"):
            h["fields"] = vals
        elif key == "SIZE":
            h["size"] = [int(v) for v in vals]
        elif key == "TYPE":
            h["type"] = [v.upper() for v in vals]
        elif key == "COUNT":
            h["count"] = [int(v) for v in vals]
        elif key == "WIDTH":
            h["width"] = int(vals[0])
        elif key == "HEIGHT":
            h["height"] = int(vals[0])
        elif key == "POINTS":
            h["points"] = int(vals[0])
        elif key == "DATA":
            h["data"] = vals[0].lower()
            break
        else:
            raise ValueError(f"unknown PCD header line: {line}")
    nf = len(h["fields"])
    if h["count"] is None:
        h["count"] = [1] * nf
    if not (len(h["size"]) == len(h["type"]) == len(h["count"]) == nf):
        raise ValueError("PCD FIELDS/SIZE/TYPE/COUNT disagree")
    if h["points"] is None:
        h["points"] = h["width"] * h["height"]
    if h["points"] != h["width"] * h["height"]:
        raise ValueError("PCD POINTS != WIDTH * HEIGHT")
    for t, s in zip(h["type"], h["size"]):
        if (t, s) not in _DTYPES:
            raise ValueError(f"unsupported PCD field type {t}{s}")
    return h, pos


def read_pcd(path: str):
    """Returns (header, fields): fields maps each field name to an array of
    shape (points,) or (points, COUNT) in the file's own type."""
    with open(path, "rb") as f:
        data = f.read()
    h, pos = _parse_header(data)
    n = h["points"]
    dt = np.dtype([(f"f{i}", _DTYPES[(t, s)], (c,)) for i, (t, s, c) in
                   enumerate(zip(h["type"], h["size"], h["count"]))])
    body = data[pos:]
    if h["data"] == "ascii":
        rows = np.array(body.split(), dtype=object)
        width = sum(h["count"])
        if rows.size < n * width:
            raise ValueError("PCD ascii data ends early")
        vals = rows[: n * width].reshape(n, width)
        cols, j = [], 0
        for i, c in enumerate(h["count"]):
            typ = np.dtype(_DTYPES[(h["type"][i], h["size"][i])])
            cols.append(np.array([[float(v) for v in r] for r in vals[:, j:j + c]], np.float64).astype(typ))
            j += c
    elif h["data"] == "binary":
        if len(body) < n * dt.itemsize:
            raise ValueError("PCD binary data ends early")
        arr = np.frombuffer(body, dtype=dt, count=n)
        cols = [arr[f"f{i}"] for i in range(len(h["fields"]))]
    elif h["data"] == "binary_compressed":
        csize, rsize = np.frombuffer(body[:8], "<u4")
        if rsize != n * dt.itemsize:
            raise ValueError("PCD compressed size disagrees with the fields")
        raw = lzf_decompress(body[8:8 + int(csize)], int(rsize))
        cols, off = [], 0
        for i, c in enumerate(h["count"]):
            typ = np.dtype(_DTYPES[(h["type"][i], h["size"][i])])
            nb = n * c * typ.itemsize
            cols.append(np.frombuffer(raw[off:off + nb], typ).reshape(n, c))
            off += nb
    else:
        raise ValueError(f"unknown PCD DATA kind: {h['data']}")
    fields = {}
    for name, col in zip(h["fields"], cols):
        col = np.asarray(col).reshape(n, -1)
        fields[name] = col[:, 0].copy() if col.shape[1] == 1 else col.copy()
    return h, fields


def load_xyz(path: str, drop_nonfinite: bool = False) -> np.ndarray:
    """The x, y, z fields as an (N, 3) float32 array (non-finite points kept
    unless drop_nonfinite, as PCL keeps them)."""
    _, f = read_pcd(path)
    xyz = np.stack([f["x"], f["y"], f["z"]], axis=1).astype(np.float32)
    if drop_nonfinite:
        xyz = xyz[np.all(np.isfinite(xyz), axis=1)]
    return xyz


def write_pcd(path: str, fields: dict, mode: str = "binary"):
    """Writes fields (name -> array of shape (N,) or (N, COUNT); float32/64,
    (u)int8-64) in the given order; binary_compressed stores a literal-run LZF
    stream."""
    names = list(fields)
    cols = [np.asarray(fields[k]) for k in names]
    cols = [c.reshape(c.shape[0], -1) for c in cols]
    n = cols[0].shape[0]
    kinds = {"f": "F", "u": "U", "i": "I"}
    hdr = ["# .PCD v0.7 - Point Cloud Data file format", "VERSION 0.7", "FIELDS " + " ".join(names),
           "SIZE " + " ".join(str(c.dtype.itemsize) for c in cols),
           "TYPE " + " ".join(kinds[c.dtype.kind] for c in cols),
           "COUNT " + " ".join(str(c.shape[1]) for c in cols),
           f"WIDTH {n}", "HEIGHT 1", "VIEWPOINT 0 0 0 1 0 0 0", f"POINTS {n}", f"DATA {mode}"]
    out = bytearray(("\n".join(hdr) + "\n").encode("ascii"))
    if mode == "ascii":
        for p in range(n):
            out += (" ".join(repr(float(v)) if c.dtype.kind == "f" else str(int(v))
                             for c in cols for v in c[p]) + "\n").encode("ascii")
    elif mode == "binary":
        dt = np.dtype([(f"f{i}", c.dtype.newbyteorder("<"), (c.shape[1],)) for i, c in enumerate(cols)])
        arr = np.zeros(n, dt)
        for i, c in enumerate(cols):
            arr[f"f{i}"] = c
        out += arr.tobytes()
    elif mode == "binary_compressed":
        raw = b"".join(np.ascontiguousarray(c.astype(c.dtype.newbyteorder("<"))).tobytes() for c in cols)
        comp = lzf_literal_stream(raw)
        out += np.array([len(comp), len(raw)], "<u4").tobytes() + comp
    else:
        raise ValueError(f"unknown PCD DATA kind: {mode}")
    with open(path, "wb") as f:
        f.write(bytes(out))
