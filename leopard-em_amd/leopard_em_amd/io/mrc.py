"""Minimal MRC2014 reader / writer for 2DTM inputs and result maps.

Counterpart of ``utils/data_io.py:13-201`` (which wraps the ``mrcfile`` package):
same function names, arguments and errors.  The file format is the published
MRC2014 layout (CCP-EM): a 1024-byte header of 256 little- or big-endian 32-bit
words, an optional extended header of NSYMBT bytes, then NZ sections of NY rows of
NX values.  Supported data modes: 0 int8, 1 int16, 2 float32, 4 complex64,
6 uint16, 12 float16.

Header words used (0-based word index):
  0-2 NX NY NZ, 3 MODE, 4-6 NXSTART.., 7-9 MX MY MZ, 10-12 cell lengths (A),
  13-15 cell angles, 16-18 MAPC MAPR MAPS, 19-21 DMIN DMAX DMEAN, 22 ISPG,
  23 NSYMBT, 26 EXTTYP, 27 NVERSION, 49-51 ORIGIN, 52 'MAP ', 53 MACHST, 54 RMS,
  55 NLABL, 56-255 ten 80-character labels.
"""

from __future__ import annotations

import os
from pathlib import Path
from typing import Any, Optional, Union

import numpy as np
import torch

HEADER_BYTES = 1024

_MODE_TO_DTYPE = {
    0: np.dtype(np.int8),
    1: np.dtype(np.int16),
    2: np.dtype(np.float32),
    4: np.dtype(np.complex64),
    6: np.dtype(np.uint16),
    12: np.dtype(np.float16),
}
_DTYPE_TO_MODE = {v: k for k, v in _MODE_TO_DTYPE.items()}


class MrcHeader(dict):
    """Parsed header fields (a plain dict with attribute access for the common ones)."""

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self["nz"], self["ny"], self["nx"])

    @property
    def voxel_size(self) -> tuple[float, float, float]:
        """(x, y, z) Angstrom per voxel; 0 where the sampling or the cell is unset."""
        out = []
        for cell, m in zip(self["cella"], (self["mx"], self["my"], self["mz"])):
            out.append(float(cell) / m if m > 0 else 0.0)
        return tuple(out)


def _byte_order(raw: bytes) -> str:
    stamp = raw[212:214]
    if stamp in (b"\x44\x44", b"\x44\x41"):
        return "<"
    if stamp == b"\x11\x11":
        return ">"
    # Files without a valid stamp: pick the order that gives a sane MODE / NX.
    le = np.frombuffer(raw[:16], dtype="<i4")
    return "<" if 0 <= le[3] <= 12 and 0 < le[0] < 1 << 20 else ">"


def read_mrc_header(mrc_path: str | os.PathLike | Path) -> MrcHeader:
    with open(mrc_path, "rb") as f:
        raw = f.read(HEADER_BYTES)
    return _parse_header(raw, str(mrc_path))


def _parse_header(raw: bytes, name: str) -> MrcHeader:
    if len(raw) < HEADER_BYTES:
        raise ValueError(f"{name}: file is shorter than an MRC header ({len(raw)} bytes)")
    bo = _byte_order(raw)
    iw = np.frombuffer(raw, dtype=bo + "i4", count=256)
    fw = np.frombuffer(raw, dtype=bo + "f4", count=256)
    if raw[208:212] not in (b"MAP ", b"MAP\x00", b"\x00\x00\x00\x00"):
        raise ValueError(f"{name}: missing 'MAP ' identifier; not an MRC2014 file")
    nlabl = int(iw[55])
    labels = [raw[224 + 80 * i: 224 + 80 * (i + 1)].rstrip(b"\x00 ").decode("ascii", "replace")
              for i in range(min(max(nlabl, 0), 10))]
    hdr = MrcHeader(
        nx=int(iw[0]), ny=int(iw[1]), nz=int(iw[2]), mode=int(iw[3]),
        nxstart=int(iw[4]), nystart=int(iw[5]), nzstart=int(iw[6]),
        mx=int(iw[7]), my=int(iw[8]), mz=int(iw[9]),
        cella=tuple(float(v) for v in fw[10:13]), cellb=tuple(float(v) for v in fw[13:16]),
        mapc=int(iw[16]), mapr=int(iw[17]), maps=int(iw[18]),
        dmin=float(fw[19]), dmax=float(fw[20]), dmean=float(fw[21]),
        ispg=int(iw[22]), nsymbt=int(iw[23]), exttyp=raw[104:108], nversion=int(iw[27]),
        origin=tuple(float(v) for v in fw[49:52]), rms=float(fw[54]), nlabl=nlabl, labels=labels,
        byteorder=bo,
    )
    if hdr["mode"] not in _MODE_TO_DTYPE:
        raise ValueError(f"{name}: unsupported MRC data mode {hdr['mode']}")
    if min(hdr["nx"], hdr["ny"], hdr["nz"]) <= 0:
        raise ValueError(f"{name}: invalid dimensions {hdr.shape[::-1]}")
    if (hdr["mapc"], hdr["mapr"], hdr["maps"]) not in ((1, 2, 3), (0, 0, 0)):
        raise ValueError(f"{name}: axis order (MAPC, MAPR, MAPS) = "
                         f"{(hdr['mapc'], hdr['mapr'], hdr['maps'])} is not supported")
    return hdr


def read_mrc(mrc_path: str | os.PathLike | Path) -> tuple[np.ndarray, MrcHeader]:
    """(data, header).  Data shape is (ny, nx) for single-section files, else (nz, ny, nx)."""
    with open(mrc_path, "rb") as f:
        raw = f.read(HEADER_BYTES)
        hdr = _parse_header(raw, str(mrc_path))
        f.seek(HEADER_BYTES + hdr["nsymbt"])
        dt = _MODE_TO_DTYPE[hdr["mode"]].newbyteorder(hdr["byteorder"])
        count = hdr["nx"] * hdr["ny"] * hdr["nz"]
        data = np.fromfile(f, dtype=dt, count=count)
    if data.size != count:
        raise ValueError(f"{mrc_path}: data block holds {data.size} of {count} values")
    data = data.astype(dt.newbyteorder("="), copy=False).reshape(hdr.shape)
    if hdr["nz"] == 1:
        data = data[0]
    return data, hdr


def read_mrc_to_numpy(mrc_path: str | os.PathLike | Path) -> np.ndarray:
    """``utils/data_io.py:13-27``: the data block as a (copied) numpy array."""
    data, _ = read_mrc(mrc_path)
    return np.array(data, copy=True)


def read_mrc_to_tensor(mrc_path: str | os.PathLike | Path) -> torch.Tensor:
    """``utils/data_io.py:30-47``: as a tensor, float16 widened to float32."""
    t = torch.from_numpy(read_mrc_to_numpy(mrc_path))
    if t.dtype == torch.float16:
        t = t.to(torch.float32)
    return t


def _header_bytes(data: np.ndarray, voxel_size, labels) -> bytes:
    nz, ny, nx = data.shape
    iw = np.zeros(256, dtype="<i4")
    fw = iw.view("<f4")
    iw[0:3] = (nx, ny, nz)
    iw[3] = _DTYPE_TO_MODE[data.dtype.newbyteorder("=")]
    iw[7:10] = (nx, ny, nz)
    vs = (voxel_size,) * 3 if np.isscalar(voxel_size) else tuple(voxel_size)
    fw[10:13] = (vs[0] * nx, vs[1] * ny, vs[2] * nz)
    fw[13:16] = (90.0, 90.0, 90.0)
    iw[16:19] = (1, 2, 3)
    if data.size:
        vals = np.abs(data) if np.iscomplexobj(data) else data.astype(np.float64)
        fw[19] = float(vals.min())
        fw[20] = float(vals.max())
        fw[21] = float(vals.mean(dtype=np.float64))
        fw[54] = float(vals.std(dtype=np.float64))
    else:
        fw[19], fw[20], fw[21], fw[54] = 0.0, -1.0, -2.0, -1.0
    # ISPG 1 marks a volume, 0 an image (or image stack).
    iw[22] = 1 if nz > 1 else 0
    iw[27] = 20141
    raw = bytearray(iw.tobytes())
    raw[208:212] = b"MAP "
    raw[212:216] = b"\x44\x44\x00\x00"
    labels = list(labels)[:10]
    raw[220:224] = np.int32(len(labels)).tobytes()
    for i, lab in enumerate(labels):
        b = lab.encode("ascii", "replace")[:80]
        raw[224 + 80 * i: 224 + 80 * i + len(b)] = b
    return bytes(raw)


def write_mrc_from_numpy(
    data: np.ndarray,
    mrc_path: str | os.PathLike | Path,
    mrc_header: Optional[dict] = None,
    overwrite: bool = False,
    voxel_size: float | tuple[float, float, float] = 0.0,
) -> None:
    """``utils/data_io.py:50-75``: write a 2D image or 3D volume (little-endian MRC2014).

    float64 data is stored as float32 (mode 2), complex128 as complex64 (mode 4).
    Raises ValueError if the file exists and ``overwrite`` is False (as ``mrcfile.new``
    does), and
    NotImplementedError for a ``mrc_header`` dict (as the reference does).
    """
    if mrc_header is not None:
        raise NotImplementedError("Setting header info is not yet implemented.")
    if os.path.exists(mrc_path) and not overwrite:
        raise ValueError(f"File '{mrc_path}' already exists; set overwrite=True to overwrite it")
    arr = np.asarray(data)
    if arr.dtype == np.float64:
        arr = arr.astype(np.float32)
    elif arr.dtype == np.complex128:
        arr = arr.astype(np.complex64)
    if arr.dtype.newbyteorder("=") not in _DTYPE_TO_MODE:
        raise ValueError(f"data type {arr.dtype} has no MRC mode")
    if arr.ndim == 2:
        arr = arr[None]
    if arr.ndim != 3:
        raise ValueError(f"MRC data must be 2D or 3D, got shape {arr.shape}")
    arr = np.ascontiguousarray(arr, dtype=arr.dtype.newbyteorder("<"))
    hdr = _header_bytes(arr, voxel_size, ["written by leopard_em_amd.io.mrc"])
    tmp = f"{mrc_path}.tmp{os.getpid()}"
    with open(tmp, "wb") as f:
        f.write(hdr)
        f.write(arr.tobytes())
    os.replace(tmp, mrc_path)


def write_mrc_from_tensor(
    data: torch.Tensor,
    mrc_path: str | os.PathLike | Path,
    mrc_header: Optional[dict] = None,
    overwrite: bool = False,
    voxel_size: float | tuple[float, float, float] = 0.0,
) -> None:
    """``utils/data_io.py:78-99``."""
    write_mrc_from_numpy(data.detach().cpu().numpy(), mrc_path, mrc_header, overwrite, voxel_size)


def load_mrc_image(file_path: str | os.PathLike | Path) -> torch.Tensor:
    """``utils/data_io.py:102-127``: a 2D image (squeezed), float32 for float16 files."""
    tensor = read_mrc_to_tensor(file_path).squeeze()
    if len(tensor.shape) != 2:
        raise ValueError(f"MRC file is not two-dimensional. Got shape: {tensor.shape}")
    return tensor


def load_mrc_volume(file_path: str | os.PathLike | Path) -> torch.Tensor:
    """``utils/data_io.py:130-157``: a 3D volume (squeezed)."""
    tensor = read_mrc_to_tensor(file_path).squeeze()
    if len(tensor.shape) != 3:
        raise ValueError(f"MRC file is not three-dimensional. Got shape: {tensor.shape}")
    return tensor


def load_template_tensor(
    template_volume: Optional[Union[torch.Tensor, Any]] = None,
    template_volume_path: Optional[Union[str, os.PathLike, Path]] = None,
) -> torch.Tensor:
    """``utils/data_io.py:160-201``: the given volume as a tensor, else the file's."""
    if template_volume is None:
        if template_volume_path is None:
            raise ValueError("template_volume or template_volume_path must be provided")
        template_volume = load_mrc_volume(template_volume_path)
    template = template_volume if isinstance(template_volume, torch.Tensor) else torch.from_numpy(template_volume)
    if template.dtype == torch.float16:
        template = template.to(torch.float32)
    return template
