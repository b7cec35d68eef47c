"""MetaImage (.mha/.mhd) and NRRD (.nhdr) readers/writers, without ITK.

The reference reads and writes these through ITK (``ApplicationUtils::LoadImageITK``,
``SaveHFieldITK``; Code/Libraries/Base/ApplicationUtils.cxx).  Only what the
reference's test data and outputs use is supported: uncompressed little-endian
float32 scalar images and 3-channel vector fields (``ElementNumberOfChannels = 3``),
with ``ElementSpacing`` / ``Offset`` (MetaImage) or ``space directions`` /
``space origin`` (NRRD).

Arrays are returned indexed ``[z, y, x]`` (x fastest, the Array3D order), vector
fields ``[z, y, x, 3]``.
"""
import os

import numpy as np

_MET_TYPES = {
    "MET_FLOAT": np.float32,
    "MET_DOUBLE": np.float64,
    "MET_UCHAR": np.uint8,
    "MET_SHORT": np.int16,
    "MET_USHORT": np.uint16,
    "MET_INT": np.int32,
}


class Image:
    """An array plus its origin and spacing (the reference ``Image<T>``)."""

    def __init__(self, data, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
        self.data = data
        self.spacing = tuple(float(s) for s in spacing)
        self.origin = tuple(float(o) for o in origin)

    @property
    def size(self):
        """(nx, ny, nz) as the reference's ``getSize()``."""
        s = self.data.shape
        return (s[2], s[1], s[0])


def read_metaimage(path):
    with open(path, "rb") as f:
        raw = f.read()
    header = {}
    pos = 0
    while True:
        end = raw.index(b"\n", pos)
        line = raw[pos:end].decode("ascii").strip()
        pos = end + 1
        if "=" not in line:
            continue
        key, val = [s.strip() for s in line.split("=", 1)]
        header[key] = val
        if key == "ElementDataFile":
            break
    dims = [int(v) for v in header["DimSize"].split()]
    nch = int(header.get("ElementNumberOfChannels", "1"))
    dtype = np.dtype(_MET_TYPES[header["ElementType"]])
    if header.get("BinaryDataByteOrderMSB", "False").lower() == "true":
        dtype = dtype.newbyteorder(">")
    if header.get("CompressedData", "False").lower() == "true":
        raise ValueError("compressed MetaImage not supported")
    count = int(np.prod(dims)) * nch
    if header["ElementDataFile"] == "LOCAL":
        data = np.frombuffer(raw, dtype=dtype, count=count, offset=pos)
    else:
        other = os.path.join(os.path.dirname(path), header["ElementDataFile"])
        data = np.fromfile(other, dtype=dtype, count=count)
    shape = tuple(reversed(dims)) + ((nch,) if nch > 1 else ())
    arr = data.reshape(shape).astype(np.float32)
    sp = [float(v) for v in header.get("ElementSpacing", "1 1 1").split()]
    org = [float(v) for v in header.get("Offset", header.get("Origin", "0 0 0")).split()]
    return Image(arr, sp, org)


def write_metaimage(path, img):
    data = np.ascontiguousarray(img.data, dtype=np.float32)
    nch = data.shape[3] if data.ndim == 4 else 1
    nz, ny, nx = data.shape[:3]
    lines = [
        "ObjectType = Image",
        "NDims = 3",
        "BinaryData = True",
        "BinaryDataByteOrderMSB = False",
        "CompressedData = False",
        "TransformMatrix = 1 0 0 0 1 0 0 0 1",
        "Offset = %.9g %.9g %.9g" % img.origin,
        "CenterOfRotation = 0 0 0",
        "AnatomicalOrientation = RAI",
        "ElementSpacing = %.9g %.9g %.9g" % img.spacing,
        "DimSize = %d %d %d" % (nx, ny, nz),
    ]
    if nch > 1:
        lines.append("ElementNumberOfChannels = %d" % nch)
    lines += ["ElementType = MET_FLOAT", "ElementDataFile = LOCAL"]
    with open(path, "wb") as f:
        f.write(("\n".join(lines) + "\n").encode("ascii"))
        f.write(data.astype("<f4").tobytes())


def _fmt_num(v):
    return "%.9g" % float(v)


def write_nrrd(path, img):
    """Detached-header NRRD: ``path`` (.nhdr) plus its raw data file (the .nhdr
    name with .raw), the layout of the reference's test data and of the outputs its
    applications write with ``OutputSuffix nhdr``.  Scalar images have three
    domain axes; vector fields a leading ``vector`` axis of 3 (no direction)."""
    data = np.ascontiguousarray(img.data, dtype=np.float32)
    nz, ny, nx = data.shape[:3]
    vec = data.ndim == 4
    sx, sy, sz = img.spacing
    dirs = "(%s,0,0) (0,%s,0) (0,0,%s)" % (_fmt_num(sx), _fmt_num(sy), _fmt_num(sz))
    raw = os.path.splitext(os.path.basename(path))[0] + ".raw"
    lines = [
        "NRRD0004",
        "type: float",
        "dimension: %d" % (4 if vec else 3),
        "space: left-posterior-superior",
        "sizes: " + ("%d %d %d %d" % (data.shape[3], nx, ny, nz) if vec else "%d %d %d" % (nx, ny, nz)),
        "space directions: " + ("none " + dirs if vec else dirs),
        "kinds: " + ("vector domain domain domain" if vec else "domain domain domain"),
        "endian: little",
        "encoding: raw",
        "space origin: (%s)" % ",".join(_fmt_num(o) for o in img.origin),
        "data file: " + raw,
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    data.astype("<f4").tofile(os.path.join(os.path.dirname(path), raw))


def read_nrrd(path):
    """Detached-header NRRD (.nhdr + raw), as the BullseyeTestBlur fixtures use;
    a leading ``vector`` axis gives a ``[z, y, x, c]`` field."""
    header = {}
    with open(path, "r") as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#") or line.startswith("NRRD"):
                continue
            if ":" in line:
                k, v = line.split(":", 1)
                header[k.strip()] = v.strip()
    sizes = [int(v) for v in header["sizes"].split()]
    types = {"float": np.float32, "double": np.float64, "short": np.int16, "unsigned char": np.uint8}
    dtype = np.dtype(types[header["type"]])
    if header.get("endian", "little") == "big":
        dtype = dtype.newbyteorder(">")
    if header.get("encoding", "raw") != "raw":
        raise ValueError("only raw NRRD supported")
    datafile = header.get("data file", header.get("datafile"))
    data = np.fromfile(os.path.join(os.path.dirname(path), datafile), dtype=dtype,
                       count=int(np.prod(sizes)))
    kinds = header.get("kinds", "").split()
    vec = len(sizes) == 4 and kinds[:1] == ["vector"]
    arr = data.reshape(tuple(reversed(sizes))).astype(np.float32)
    if vec:
        arr = np.ascontiguousarray(arr)  # [z, y, x, c]: the vector axis is the fastest
    sp = [1.0, 1.0, 1.0]
    if "space directions" in header:
        vecs = header["space directions"].replace(")", "").split("(")[1:]
        for i, v in enumerate(vecs[:3]):
            comps = [float(c) for c in v.split(",")]
            sp[i] = float(np.sqrt(sum(c * c for c in comps)))
    org = [0.0, 0.0, 0.0]
    if "space origin" in header:
        org = [float(c) for c in header["space origin"].strip("()").split(",")]
    return Image(arr, sp, org)
