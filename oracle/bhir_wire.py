"""ORACLE — test infrastructure only.  Nothing in the product imports this.

Reader of recorded flush files (the BhIR wire format written by the bridge's
bhc_record_flushes, bohrium_amd/csrc/core/bh_ir.cpp) that turns them into the
trace records `omp_ve.replay` executes, so that the engine and the oracle run
the very same bytes.  Restated from the format description, independently of
the C++ reader:

  file:    "BHIRFILE", u32 version 1, then per flush: u64 archive size, the
           archive, and per base the archive introduces with data: u64 bytes,
           the bytes
  archive: "BHIR", u32 version 1, u64 nrepeats, u64 repeat condition,
           u64 instructions x (i32 opcode, u32 operands x view, i32 constant
           type, 16 constant bytes), u64 syncs x u64, u64 new bases x
           (u64 address, i32 type, i64 nelem, u8 has data)
  view:    u64 base (0: constant), i64 start, i64 ndim, ndim x i64 shape,
           ndim x i64 stride, u64 slide dims x 6 i64, i64 iteration counter,
           u64 resets x 3 i64
Bases are named by the recording process's addresses; an address freed in one
flush may name a new base later, so each introduction gets a fresh trace id.
This mirrors the reference's de-serializer (core/bh_ir.cpp:37-117): new bases
are introduced in the order of their first appearance in the instruction list.
"""
import struct

import numpy as np

# bh_type ids (reference core/codegen/types.json) -> numpy dtypes
TYPES = {0: np.bool_, 1: np.int8, 2: np.int16, 3: np.int32, 4: np.int64, 5: np.uint8, 6: np.uint16,
         7: np.uint32, 8: np.uint64, 9: np.float32, 10: np.float64, 11: np.complex64, 12: np.complex128}
R123 = 13
BH_FREE, BH_NONE, BH_TALLY, BH_RANDOM = 55, 57, 58, 71


class _Buf:
    def __init__(self, data):
        self.data, self.pos = data, 0

    def take(self, n):
        if self.pos + n > len(self.data):
            raise ValueError("recorded flush: truncated")
        b = self.data[self.pos:self.pos + n]
        self.pos += n
        return b

    def u(self, fmt):
        size = struct.calcsize("<" + fmt)
        return struct.unpack("<" + fmt, self.take(size))[0]


def parse_archive(data):
    """One archive -> dict(nrepeats, cond, instrs=[(opcode, [view|None], (type, 16 bytes))], syncs, news)."""
    b = _Buf(data)
    if b.take(4) != b"BHIR" or b.u("I") != 1:
        raise ValueError("recorded flush: not a version 1 BhIR archive")
    out = {"nrepeats": b.u("Q"), "cond": b.u("Q"), "instrs": [], "syncs": [], "news": []}
    for _ in range(b.u("Q")):
        opcode, nops = b.u("i"), b.u("I")
        ops = []
        for _ in range(nops):
            base, start, ndim = b.u("Q"), b.u("q"), b.u("q")
            shape = tuple(b.u("q") for _ in range(ndim))
            stride = tuple(b.u("q") for _ in range(ndim))
            slides = tuple(tuple(b.u("q") for _ in range(6)) for _ in range(b.u("Q")))
            b.u("q")  # iteration counter
            for _ in range(b.u("Q")):
                b.take(24)
            ops.append(None if base == 0 else (base, start, shape, stride, slides))
        ctype = b.u("i")
        cbytes = b.take(16)
        out["instrs"].append((opcode, ops, (ctype, cbytes)))
    out["syncs"] = [b.u("Q") for _ in range(b.u("Q"))]
    for _ in range(b.u("Q")):
        out["news"].append((b.u("Q"), b.u("i"), b.u("q"), b.u("B")))
    if b.pos != len(data):
        raise ValueError("recorded flush: trailing bytes in an archive")
    return out


def read_file(path):
    """[(archive dict, {address: numpy data of the introduced bases that carry data})] per flush."""
    with open(path, "rb") as f:
        raw = f.read()
    b = _Buf(raw)
    if b.take(8) != b"BHIRFILE" or b.u("I") != 1:
        raise ValueError("%s: not a recorded-flush file" % path)
    flushes = []
    while b.pos < len(raw):
        arc = parse_archive(b.take(b.u("Q")))
        data = {}
        for addr, t, nelem, has in arc["news"]:
            if has:
                nbytes = b.u("Q")
                dt = np.dtype(TYPES[t])
                if nbytes != nelem * dt.itemsize:
                    raise ValueError("recorded flush: data size does not match its base")
                data[addr] = np.frombuffer(b.take(nbytes), dtype=dt).copy()
        flushes.append((arc, data))
    return flushes


def to_trace(flushes):
    """Trace records for omp_ve.replay, and {address: trace id} of the bases alive after the last
    flush (the ones a replay's results are read from)."""
    trace, live = [], {}
    nxt = [0]
    for arc, data in flushes:
        for addr, t, nelem, has in arc["news"]:
            live[addr] = nxt[0]
            nxt[0] += 1
            trace.append(("new", live[addr], np.dtype(TYPES[t]).str, int(nelem)))
            if has:
                trace.append(("data", live[addr], data[addr]))
        freed = []
        for opcode, ops, (ctype, cbytes) in arc["instrs"]:
            if opcode == BH_FREE:
                freed.append(ops[0][0])
                trace.append(("free", live[ops[0][0]]))
                continue
            if opcode in (BH_NONE, BH_TALLY):
                continue
            rec_ops = [None if o is None else (live[o[0]], o[1], o[2], o[3]) for o in ops]
            has_const = any(o is None for o in ops) or opcode == BH_RANDOM
            const = None
            if has_const:
                if ctype == R123:
                    const = ("r123", struct.unpack("<QQ", cbytes))
                else:
                    dt = np.dtype(TYPES[ctype])
                    const = (dt.str, np.frombuffer(cbytes[:dt.itemsize], dtype=dt)[0].item())
            slides = tuple((o[4] if o is not None else ()) for o in ops)
            trace.append(("op", opcode, rec_ops, const, slides if any(slides) else None))
        trace.append(("flush", int(arc["nrepeats"]), None if arc["cond"] == 0 else live[arc["cond"]]))
        for addr in freed:
            live.pop(addr, None)
    return trace, dict(live)
