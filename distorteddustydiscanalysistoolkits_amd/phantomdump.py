"""Writes Phantom native binary dumps: synthetic benchmark inputs and the reader's test files.

The layout is the one ParticleIO.jl's reader consumes (read_phantom.jl:20-207): every record is
framed by a 4-byte tag before and after its payload (the payload length, as Fortran writes it);
the capture-pattern record holds i1 = 60769, r1 = 60878.0, i2 = 60878, iversion, i3 = 690706;
then a 100-character file identifier; eight typed header blocks (default int, int*1, int*2,
int*4, int*8, default real, real*4, real*8), each an int32 count record followed, when the
count is positive, by a record of 16-character tags and a record of values; then an int32
block count, one (int64 n, 8 x int32 counts-per-type) record per block, and per block the
arrays grouped by type slot, each a 16-character tag record and a data record.
"""
import struct

import numpy as np

# type slots of the header and array blocks
DEF_INT, I8, I16, I32, I64, DEF_REAL, R4, R8 = range(8)


def _np_type(slot, int_bytes, real_bytes):
    return {DEF_INT: np.int32 if int_bytes == 4 else np.int64, I8: np.int8, I16: np.int16, I32: np.int32,
            I64: np.int64, DEF_REAL: np.float32 if real_bytes == 4 else np.float64, R4: np.float32,
            R8: np.float64}[slot]


def _record(payload: bytes, tag_override=None) -> bytes:
    t = struct.pack("<i", len(payload))
    t2 = t if tag_override is None else struct.pack("<i", tag_override)
    return t + payload + t2


def _tag16(s: str) -> bytes:
    b = s.encode()
    assert len(b) <= 16
    return b + b" " * (16 - len(b))


def dump_bytes(header=None, blocks=(), ident="Phantom test dump", int_bytes=4, real_bytes=4, iversion=1,
               bad_end_tag_of_block=None) -> bytes:
    """header: {slot: [(tag, value), ...]}; blocks: [(n, [(slot, tag, array), ...]), ...] (the
    arrays of a block are written grouped by slot, in slot order)."""
    header = header or {}
    ti = np.int32 if int_bytes == 4 else np.int64
    tr = np.float32 if real_bytes == 4 else np.float64
    cap = (np.array([60769], ti).tobytes() + np.array([60878.0], tr).tobytes() + np.array([60878], ti).tobytes()
           + np.array([iversion], ti).tobytes() + np.array([690706], ti).tobytes())
    out = [_record(cap), _record(ident.encode().ljust(100))]
    for slot in range(8):
        items = header.get(slot, [])
        out.append(_record(struct.pack("<i", len(items))))
        if items:
            t = _np_type(slot, int_bytes, real_bytes)
            out.append(_record(b"".join(_tag16(k) for k, _ in items)))
            out.append(_record(np.array([v for _, v in items], t).tobytes()))
    out.append(_record(struct.pack("<i", len(blocks))))
    for n, arrays in blocks:
        nums = [sum(1 for a in arrays if a[0] == s) for s in range(8)]
        out.append(_record(struct.pack("<q", n) + struct.pack("<8i", *nums)))
    for b, (n, arrays) in enumerate(blocks):
        for slot in range(8):
            for s, tag, arr in arrays:
                if s != slot:
                    continue
                t = _np_type(slot, int_bytes, real_bytes)
                data = np.ascontiguousarray(arr, t)
                assert data.size == n
                out.append(_record(_tag16(tag)))
                bad = (len(data.tobytes()) + 1) if bad_end_tag_of_block == b else None
                out.append(_record(data.tobytes(), tag_override=bad))
    return b"".join(out)


def write_dump(path, **kw):
    with open(path, "wb") as f:
        f.write(dump_bytes(**kw))
    return path


def disc_dump(path, d, hfact=1.2, massoftype=None, with_rho=False, real_bytes=4):
    """A gas-only dump of a synthetic disc (dict of columns from disc.make_disc): x y z h as default
    reals, vx vy vz as real*4, itype as int*1; header time, hfact, massoftype."""
    n = d["x"].size
    m = float(d["m"][0]) if massoftype is None else massoftype
    arrays = [(DEF_REAL, k, d[k]) for k in ("x", "y", "z", "h")]
    if with_rho:
        arrays.append((R4, "rho", d["rho"]))
    arrays += [(R4, k, d[k]) for k in ("vx", "vy", "vz")]
    arrays.append((I8, "itype", np.ones(n, np.int8)))
    header = {DEF_INT: [("nparttot", n), ("ntypes", 8)],
              DEF_REAL: [("time", 0.5), ("hfact", hfact), ("massoftype", m)]}
    return write_dump(path, header=header, blocks=[(n, arrays), (0, [])], real_bytes=real_bytes)
