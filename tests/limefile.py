"""Minimal LIME / ILDG reader-writer in Python, written from the public LIME
specification, used by the tests as an independent check of libqcbflow's own
framing code.  Header: 144 bytes big-endian -- magic 0x456789ab (u32),
version (u16), flags (u16: 0x8000 MB, 0x4000 ME), payload length (u64),
type (128 bytes NUL-padded); payload padded with zeros to 8 bytes.
"""
import struct

import numpy as np

MAGIC = 0x456789AB


def header(rtype, length, mb, me, version=1):
    flags = (0x8000 if mb else 0) | (0x4000 if me else 0)
    t = rtype.encode()
    assert len(t) < 128
    return struct.pack(">IHHQ", MAGIC, version, flags, length) + t + b"\0" * (128 - len(t))


def records(data):
    """[(type, mb, me, payload_offset, length)] of every record in bytes `data`."""
    out = []
    pos = 0
    while pos + 144 <= len(data):
        magic, version, flags, length = struct.unpack(">IHHQ", data[pos:pos + 16])
        if magic != MAGIC:
            raise ValueError(f"bad magic at {pos}")
        rtype = data[pos + 16:pos + 144].split(b"\0", 1)[0].decode()
        out.append((rtype, bool(flags & 0x8000), bool(flags & 0x4000), pos + 144, length))
        pos += 144 + ((length + 7) // 8) * 8
    return out


def ildg_xml(dims, precision=64):
    lx, ly, lz, lt = dims
    return ('<?xml version="1.0" encoding="UTF-8"?>\n'
            "<ildgFormat><version>1.0</version><field>su3gauge</field>"
            f"<precision>{precision}</precision><lx>{lx}</lx><ly>{ly}</ly><lz>{lz}</lz>"
            f"<lt>{lt}</lt></ildgFormat>")


def write_ildg(path, links, dims, precision=64, framing="reference"):
    """links: V*72 doubles in the reference order.  framing "reference" writes
    ildg-format (MB=1, ME=1) then ildg-binary-data (MB=0, ME=1) as
    save_ildg_clime does (src/io/ildg.cpp:59-67); "standard" writes one message
    ildg-format (MB) + ildg-binary-data + ildg-data-lfn (ME)."""
    xml = ildg_xml(dims, precision).encode()
    dt = ">f8" if precision == 64 else ">f4"
    payload = np.ascontiguousarray(links, dtype=np.float64).reshape(-1).astype(dt).tobytes()
    pad = lambda b: b + b"\0" * ((8 - len(b) % 8) % 8)
    with open(path, "wb") as f:
        if framing == "reference":
            f.write(header("ildg-format", len(xml), True, True) + pad(xml))
            f.write(header("ildg-binary-data", len(payload), False, True) + payload)
        else:
            lfn = b"lfn://qcbflow/test"
            f.write(header("ildg-format", len(xml), True, False) + pad(xml))
            f.write(header("ildg-binary-data", len(payload), False, False) + pad(payload))
            f.write(header("ildg-data-lfn", len(lfn), False, True) + pad(lfn))


def read_ildg(path):
    data = open(path, "rb").read()
    recs = {r[0]: r for r in records(data)}
    _, _, _, off, length = recs["ildg-binary-data"]
    prec = 64
    if "ildg-format" in recs:
        _, _, _, xo, xl = recs["ildg-format"]
        xml = data[xo:xo + xl].decode()
        if "<precision>32</precision>" in xml:
            prec = 32
    dt = ">f8" if prec == 64 else ">f4"
    return np.frombuffer(data[off:off + length], dtype=dt).astype(np.float64)
