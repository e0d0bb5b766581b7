"""Minimal NIfTI-1 and part-file reader/writer for the host-driver tests,
written from the NIfTI-1 standard (348-byte header, "n+1" single file) and
the part-file layout of the reference's split/fit/merge steps
(split_parts.cc:19-34: int32 nvox, int32 nrows, int64 byte count, FP64
matrix).  Independent of the C++ implementation it checks."""
import gzip
import struct

import numpy as np

DT = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8, 512: np.uint16}
DT_CODE = {np.dtype(v): k for k, v in DT.items()}


def write_nifti(path, data, pixdim=(2.0, 2.0, 2.0), slope=0.0, inter=0.0, big_endian=False, sform=None):
    """data [nx, ny, nz(, nt)] in numpy axis order (x first); stored x fastest."""
    data = np.asarray(data)
    e = ">" if big_endian else "<"
    dims = [data.ndim] + list(data.shape) + [1] * (7 - data.ndim)
    h = bytearray(348)
    struct.pack_into(e + "i", h, 0, 348)
    struct.pack_into(e + "8h", h, 40, *dims)
    code = DT_CODE[data.dtype]
    struct.pack_into(e + "hh", h, 70, code, data.dtype.itemsize * 8)
    struct.pack_into(e + "8f", h, 76, 1.0, *pixdim, 1.0, 1.0, 1.0, 1.0)
    struct.pack_into(e + "f", h, 108, 352.0)
    struct.pack_into(e + "ff", h, 112, slope, inter)
    struct.pack_into(e + "B", h, 123, 10)  # mm, s
    if sform is not None:
        struct.pack_into(e + "h", h, 254, 1)
        struct.pack_into(e + "12f", h, 280, *np.asarray(sform, dtype=np.float32).ravel()[:12])
    h[344:348] = b"n+1\0"
    body = bytes(h) + b"\0\0\0\0" + data.astype(data.dtype.newbyteorder(e)).tobytes(order="F")
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wb") as f:
        f.write(body)


def read_nifti(path):
    """Returns (data [nx, ny, nz, nt] float, header dict)."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as f:
        b = f.read()
    e = "<" if struct.unpack_from("<i", b, 0)[0] == 348 else ">"
    dims = struct.unpack_from(e + "8h", b, 40)
    code, _ = struct.unpack_from(e + "hh", b, 70)
    pixdim = struct.unpack_from(e + "8f", b, 76)
    off = int(struct.unpack_from(e + "f", b, 108)[0])
    slope, inter = struct.unpack_from(e + "ff", b, 112)
    cal_max, cal_min = struct.unpack_from(e + "ff", b, 124)
    sform_code = struct.unpack_from(e + "h", b, 254)[0]
    srow = struct.unpack_from(e + "12f", b, 280)
    shape = tuple(dims[1:dims[0] + 1])
    dt = np.dtype(DT[code]).newbyteorder(e)
    n = int(np.prod(shape))
    x = np.frombuffer(b, dtype=dt, count=n, offset=off).reshape(shape, order="F").astype(np.float64)
    if slope not in (0.0,) and not (slope == 1.0 and inter == 0.0):
        x = x * slope + inter
    while x.ndim < 4:
        x = x[..., None]
    return x, dict(dims=dims, datatype=code, pixdim=pixdim, cal_max=cal_max, cal_min=cal_min, magic=b[344:348],
                   sform_code=sform_code, srow=srow)


def masked(vol4, mask3):
    """newimage's volume4D::matrix(mask): voxels with mask > 0 in x-fastest
    order, one row per volume -> returned voxel-major [nvox][nt]."""
    m = mask3.reshape(-1, order="F") > 0
    flat = vol4.reshape(-1, vol4.shape[3], order="F")
    return flat[m]


def write_part(path, mat, row_major=False):
    """mat [nvox][nrows] voxel-major."""
    mat = np.asarray(mat, dtype=np.float64)
    nvox, nrows = mat.shape
    raw = mat.T.copy() if row_major else mat
    with open(path, "wb") as f:
        f.write(struct.pack("<iiq", nvox, nrows, nvox * nrows * 8))
        f.write(raw.tobytes(order="C"))


def read_part(path, row_major=False):
    with open(path, "rb") as f:
        b = f.read()
    nvox, nrows, nbytes = struct.unpack_from("<iiq", b, 0)
    assert nbytes == nvox * nrows * 8 and len(b) == 16 + nbytes
    x = np.frombuffer(b, dtype="<f8", offset=16)
    return x.reshape(nrows, nvox).T.copy() if row_major else x.reshape(nvox, nrows).copy()
