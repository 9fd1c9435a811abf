"""The reference's on-disk formats for the level-set path.

All are Fortran unformatted sequential files (form='unformatted'): each
record is framed by a 4-byte little-endian length before and after it;
records longer than 2**31 - 1 bytes are split into subrecords whose leading
marker is negated when another subrecord follows (and whose trailing marker
is negated when one precedes), the convention of gfortran and flang.

* grid      per zone: x, y, z (and the wall distance when turbulence is on),
            one record each, img*jmg*kmg FP64 values, i fastest
            (init.F90:1863-1880)
* phi_ini   per zone: phi, then phi_n, one record each (init_LSM.F90:350-357)
* phi.NNN   restart of one rank: phi, then phi_n over the rank's owned nodes
            (main.F90:404-455)
* debugfiles/<name>.NN  a debug dump of one rank: one record of a field over
            the rank's nodes without ghost layers (global_debug.F90:15-99)

Arrays are returned shaped (nk, nj, ni) (C order, i fastest), the layout the
rest of this package uses.
"""
from __future__ import annotations

import struct
from typing import List, Sequence, Tuple

import numpy as np

_MAXREC = 2**31 - 1


def write_records(path: str, records: Sequence[bytes]) -> None:
    with open(path, "wb") as f:
        for rec in records:
            data = memoryview(rec).cast("B")
            n = len(data)
            if n <= _MAXREC:
                f.write(struct.pack("<i", n)); f.write(data); f.write(struct.pack("<i", n))
                continue
            pos = 0
            first = True
            while pos < n:
                m = min(_MAXREC, n - pos)
                last = pos + m == n
                head = m if last else -m
                tail = m if first else -m
                f.write(struct.pack("<i", head)); f.write(data[pos:pos + m]); f.write(struct.pack("<i", tail))
                pos += m
                first = False


def read_records(path: str) -> List[bytes]:
    recs: List[bytes] = []
    with open(path, "rb") as f:
        while True:
            h = f.read(4)
            if not h:
                break
            parts = []
            while True:
                (m,) = struct.unpack("<i", h)
                size = abs(m)
                parts.append(f.read(size))
                t = f.read(4)
                if len(t) != 4 or abs(struct.unpack("<i", t)[0]) != size:
                    raise ValueError(f"{path}: record framing mismatch")
                if m >= 0:
                    break
                h = f.read(4)
            recs.append(b"".join(parts))
    return recs


def _field(buf: bytes, shape) -> np.ndarray:
    nk, nj, ni = shape
    a = np.frombuffer(buf, dtype="<f8")
    if a.size != ni * nj * nk:
        raise ValueError(f"record holds {a.size} values, expected {ni * nj * nk}")
    return a.reshape(nk, nj, ni).astype(np.float64)


def _bytes(a) -> bytes:
    return np.ascontiguousarray(a, dtype="<f8").tobytes()


def write_grid(path: str, zones: Sequence[Tuple[np.ndarray, ...]]) -> None:
    """zones: per zone (x, y, z) or (x, y, z, d), each shaped (nk, nj, ni)."""
    write_records(path, [_bytes(a) for zone in zones for a in zone])


def read_grid(path: str, shapes: Sequence[Tuple[int, int, int]], turbulence: bool = False):
    """init.F90:1863-1880; shapes per zone as (nk, nj, ni)."""
    recs = read_records(path)
    per = 4 if turbulence else 3
    if len(recs) < per * len(shapes):
        raise ValueError(f"{path}: {len(recs)} records, expected {per * len(shapes)}")
    return [tuple(_field(recs[per * z + q], s) for q in range(per)) for z, s in enumerate(shapes)]


def write_phi_ini(path: str, zones: Sequence[Tuple[np.ndarray, np.ndarray]]) -> None:
    """zones: per zone (phi, phi_n), each shaped (nk, nj, ni)."""
    write_records(path, [_bytes(a) for zone in zones for a in zone])


def read_phi_ini(path: str, shapes: Sequence[Tuple[int, int, int]]):
    """init_LSM.F90:350-357: per zone the records phi and phi_n."""
    recs = read_records(path)
    if len(recs) < 2 * len(shapes):
        raise ValueError(f"{path}: {len(recs)} records, expected {2 * len(shapes)}")
    return [(_field(recs[2 * z], s), _field(recs[2 * z + 1], s)) for z, s in enumerate(shapes)]


def write_phi_restart(path: str, phi: np.ndarray, phi_n: np.ndarray) -> None:
    """main.F90:404-455 (one rank's owned nodes)."""
    write_records(path, [_bytes(phi), _bytes(phi_n)])


def read_phi_restart(path: str, shape: Tuple[int, int, int]):
    recs = read_records(path)
    if len(recs) != 2:
        raise ValueError(f"{path}: {len(recs)} records, expected 2")
    return _field(recs[0], shape), _field(recs[1], shape)


def grid_from_file(path: str, shape: Tuple[int, int, int], dc: float = 1.0, de: float = 1.0,
                   dz: float = 1.0, zone: int = 0, turbulence: bool = False):
    """A one-zone grid file as lsm_amd.inputs.Grid with the metrics of
    mg_metrics.F90:53-262 (dc = de = dz = 1 in the reference)."""
    from . import inputs
    shapes = [shape] * (zone + 1)
    x, y, z = read_grid(path, shapes, turbulence)[zone][:3]
    g = inputs.Grid(x, y, z)
    g.dc, g.de, g.dz = dc, de, dz
    return inputs.compute_metrics(g)


def debug_path(outdir: str, name: str, rank: int = 0) -> str:
    """global_debug.F90:44-47: trim(OUTPUTdir) // 'debugfiles/' // name // '.' //
    the rank as i2.2 (two digits, '**' from 100 on, as the Fortran edit
    descriptor writes it); OUTPUTdir is concatenated as given."""
    nn = f"{rank:02d}" if 0 <= rank <= 99 else "**"
    return f"{outdir.rstrip()}debugfiles/{name}.{nn}"


def write_debug(path: str, field: np.ndarray, ghost: Tuple[int, int, int] = (0, 0, 0)) -> None:
    """outputD3_real (global_debug.F90:62-99, through outputD1_real): the
    field of one rank, shaped (nk, nj, ni) with `ghost` = (igp, jgp, kgp)
    ghost layers on each side, written without them as one record."""
    igp, jgp, kgp = ghost
    a = np.asarray(field)
    nk, nj, ni = a.shape
    inner = a[kgp:nk - kgp, jgp:nj - jgp, igp:ni - igp]
    write_records(path, [_bytes(inner)])


def read_debug(path: str, shape: Tuple[int, int, int]) -> np.ndarray:
    """A debug dump of `shape` = (nk, nj, ni) nodes (without ghost layers)."""
    recs = read_records(path)
    if len(recs) != 1:
        raise ValueError(f"{path}: {len(recs)} records, expected 1")
    return _field(recs[0], shape)
