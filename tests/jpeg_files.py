"""A small JPEG encoder for the decoder's tests, written from ITU-T T.81:
sequential (baseline) and progressive Huffman coding with spectral selection
and successive approximation, one or three components with chosen sampling
factors, restart intervals.  Huffman tables are built from the symbol counts
of each image (canonical codes, T.81 Annex C and K.2), so the decoder's table
construction is exercised with code lengths other than the example tables'.

`coefficients` gives the quantised DCT blocks the files carry, so a test can
compare the decoded pixels with an FP64 inverse DCT of exactly those."""
from __future__ import annotations

import heapq
import math
import struct

import numpy as np

ZIGZAG = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14,
          21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60,
          61, 54, 47, 55, 62, 63]

_C = np.array([[(math.sqrt(0.5) if u == 0 else 1.0) * math.cos((2 * x + 1) * u * math.pi / 16) / 2
                for u in range(8)] for x in range(8)])  # _C[x, u]


def fdct(block: np.ndarray) -> np.ndarray:
    """Forward DCT of a level-shifted 8x8 block (T.81 A.3.3), FP64."""
    return _C.T @ block @ _C


def idct(coef: np.ndarray) -> np.ndarray:
    return _C @ coef @ _C.T


def _category(v: int) -> int:
    return 0 if v == 0 else int(abs(v)).bit_length()


def _bits_of(v: int, s: int) -> int:
    return v if v >= 0 else v + (1 << s) - 1


class _Bits:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, code: int, length: int):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((code >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)  # byte stuffing (F.1.2.3)
                self.acc = self.n = 0

    def flush(self):
        if self.n:
            self.put((1 << (8 - self.n)) - 1, 8 - self.n)  # pad with 1-bits


def _huffman(counts: dict) -> tuple[list, list, dict]:
    """Code lengths by Huffman's algorithm over the used symbols plus a dummy
    of count 0 that takes the all-ones code; canonical codes (C.2)."""
    syms = sorted(counts)
    if not syms:
        syms = [0]
        counts = {0: 1}
    heap = [(counts[s], i, [s]) for i, s in enumerate(syms)] + [(0, len(syms), [None])]
    heapq.heapify(heap)
    depth = {s: 0 for s in syms + [None]}
    k = len(heap)
    while len(heap) > 1:
        a = heapq.heappop(heap)
        b = heapq.heappop(heap)
        for s in a[2] + b[2]:
            depth[s] += 1
        heapq.heappush(heap, (a[0] + b[0], k, a[2] + b[2]))
        k += 1
    if max(depth.values()) > 16:  # flat code of one length: always within 16 bits here
        L = max(1, math.ceil(math.log2(len(syms) + 1)))
        depth = {s: L for s in syms + [None]}
    order = sorted(syms, key=lambda s: (depth[s], s))
    bits = [0] * 17
    for s in order:
        bits[depth[s]] += 1
    codes, code, prev = {}, 0, 0
    for s in order:
        code <<= depth[s] - prev
        prev = depth[s]
        codes[s] = (code, depth[s])
        code += 1
    return bits[1:], order, codes


def _segment(marker: int, payload: bytes) -> bytes:
    return struct.pack(">BBH", 0xFF, marker, len(payload) + 2) + payload


def coefficients(image: np.ndarray, sampling=(1, 1), quality_table: np.ndarray | None = None):
    """Quantised DCT blocks of an (H, W) grey or (H, W, 3) RGB uint8 image:
    JFIF YCbCr, chroma subsampled by `sampling` = (hmax, vmax) (box average),
    level shift, forward DCT, rounding division by the table (all ones if
    None).  Returns (components, tables): per component a dict with h, v and
    blocks (by, bx, 8, 8 natural order, int), and the quantisation tables."""
    img = np.asarray(image, np.float64)
    H, W = img.shape[:2]
    hmax, vmax = sampling
    if img.ndim == 2:
        planes = [(img, 1, 1, 0)]
    else:
        r, g, b = img[..., 0], img[..., 1], img[..., 2]
        y = 0.299 * r + 0.587 * g + 0.114 * b
        cb = -0.168736 * r - 0.331264 * g + 0.5 * b + 128
        cr = 0.5 * r - 0.418688 * g - 0.081312 * b + 128
        planes = [(y, hmax, vmax, 0), (cb, 1, 1, 1), (cr, 1, 1, 1)]
    q = [np.ones((8, 8)) if quality_table is None else np.asarray(quality_table, np.float64)] * 2
    mcux, mcuy = -(-W // (8 * hmax)), -(-H // (8 * vmax))
    comps = []
    for plane, h, v, tq in planes:
        fx, fy = hmax // h, vmax // v
        cw, ch = -(-W * h // hmax), -(-H * v // vmax)
        # box-average to the component's size (edge-replicated to whole factors)
        pad = np.pad(plane, ((0, ch * fy - H), (0, cw * fx - W)), mode="edge")
        sub = pad.reshape(ch, fy, cw, fx).mean(axis=(1, 3))
        bw, bh = mcux * h, mcuy * v
        full = np.pad(sub, ((0, bh * 8 - ch), (0, bw * 8 - cw)), mode="edge")
        blocks = np.zeros((bh, bw, 8, 8), np.int64)
        for by in range(bh):
            for bx in range(bw):
                f = fdct(full[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] - 128.0) / q[tq]
                blocks[by, bx] = np.sign(f) * np.floor(np.abs(f) + 0.5)
        comps.append({"h": h, "v": v, "tq": tq, "blocks": blocks, "bw": -(-cw // 8), "bh": -(-ch // 8)})
    return comps, q


def encode(width: int, height: int, comps, tables, progressive: bool = False, restart: int = 0,
           jfif: bool = True) -> bytes:
    """A JPEG file carrying exactly `comps` (from coefficients())."""
    hmax = max(c["h"] for c in comps)
    vmax = max(c["v"] for c in comps)
    mcux, mcuy = -(-width // (8 * hmax)), -(-height // (8 * vmax))
    out = bytearray(b"\xff\xd8")
    if jfif:
        out += _segment(0xE0, b"JFIF\x00\x01\x01\x00\x00\x01\x00\x01\x00\x00")
    for t, tab in enumerate(tables[:1 + (len(comps) > 1)]):
        out += _segment(0xDB, bytes([t]) + bytes(int(tab.reshape(-1)[ZIGZAG[k]]) for k in range(64)))
    ids = [1, 2, 3][:len(comps)]
    sof = struct.pack(">BHHB", 8, height, width, len(comps))
    for cid, c in zip(ids, comps):
        sof += bytes([cid, (c["h"] << 4) | c["v"], c["tq"]])
    out += _segment(0xC2 if progressive else 0xC0, sof)
    if restart:
        out += _segment(0xDD, struct.pack(">H", restart))

    def units(sc, interleaved):
        """(component index, by, bx) per data unit, with restart-interval breaks."""
        seq = []
        if interleaved:
            for my in range(mcuy):
                for mx in range(mcux):
                    mcu = []
                    for ci in sc:
                        c = comps[ci]
                        for v in range(c["v"]):
                            for h in range(c["h"]):
                                mcu.append((ci, my * c["v"] + v, mx * c["h"] + h))
                    seq.append(mcu)
        else:
            c = comps[sc[0]]
            for by in range(c["bh"]):
                for bx in range(c["bw"]):
                    seq.append([(sc[0], by, bx)])
        return seq

    def scan(sc, ss, se, ah, al):
        seq = units(sc, len(sc) > 1)
        dcodes = acodes = dbits = dorder = abits = aorder = None  # built from the first pass's counts
        # pass 1 collects the symbols, pass 2 writes them with the built tables
        for pass_ in (0, 1):
            dc_counts, ac_counts = {}, {}
            bw = _Bits()
            pred = {ci: 0 for ci in sc}

            def sym(table, counts, s, bits_len=0, bits_val=0):
                if pass_ == 0:
                    counts[s] = counts.get(s, 0) + 1
                else:
                    code, length = table[s]
                    bw.put(code, length)
                    if bits_len:
                        bw.put(bits_val, bits_len)

            for n, mcu in enumerate(seq):
                if restart and n and n % restart == 0:
                    if pass_:
                        bw.flush()
                        bw.out += bytes([0xFF, 0xD0 + ((n // restart - 1) % 8)])
                    pred = {ci: 0 for ci in sc}
                for ci, by, bx in mcu:
                    z = comps[ci]["blocks"][by, bx].reshape(-1)[ZIGZAG]
                    if ss == 0:
                        if ah == 0:
                            val = int(z[0]) >> al
                            diff = val - pred[ci]
                            pred[ci] = val
                            s = _category(diff)
                            sym(dcodes, dc_counts, s, s, _bits_of(diff, s))
                        elif pass_:
                            bw.put((int(z[0]) >> al) & 1, 1)
                        if se == 0:
                            continue
                    lo = max(ss, 1)  # a sequential scan codes DC, then AC 1-63
                    if ah == 0:  # AC first pass: magnitudes shifted by al, EOB runs of 1
                        vals = [int(np.sign(x)) * (abs(int(x)) >> al) for x in z[lo:se + 1]]
                        run = 0
                        last = max([i for i, x in enumerate(vals) if x != 0], default=-1)
                        for i, x in enumerate(vals[:last + 1]):
                            if x == 0:
                                run += 1
                                continue
                            while run > 15:
                                sym(acodes, ac_counts, 0xF0)
                                run -= 16
                            s = _category(x)
                            sym(acodes, ac_counts, (run << 4) | s, s, _bits_of(x, s))
                            run = 0
                        if last < se - lo:
                            sym(acodes, ac_counts, 0x00)
                        continue
                    # AC refinement (G.1.2.3): each new coefficient (magnitude 1 at
                    # this bit) is a symbol (zero-history run, 1) with its sign, then
                    # the correction bits of the nonzero coefficients passed on the
                    # way; 16 zero-history coefficients without a new one are a ZRL;
                    # after the last new one, EOB and the remaining correction bits
                    full = [abs(int(x)) for x in z]
                    hist = [(f >> ah) != 0 for f in full]  # nonzero after the previous pass
                    new = [not hist[i] and ((full[i] >> al) & 1) == 1 for i in range(64)]
                    last_new = max([i for i in range(ss, se + 1) if new[i]], default=ss - 1)
                    pending, zeros = [], 0

                    def emit(symbol, sign_bit=None):
                        if pass_ == 0:
                            ac_counts[symbol] = ac_counts.get(symbol, 0) + 1
                            return
                        code, length = acodes[symbol]
                        bw.put(code, length)
                        if sign_bit is not None:
                            bw.put(sign_bit, 1)
                        for b in pending:
                            bw.put(b, 1)

                    for k in range(ss, se + 1):
                        if hist[k]:
                            pending.append((full[k] >> al) & 1)
                        elif new[k]:
                            emit((zeros << 4) | 1, 0 if int(z[k]) < 0 else 1)
                            pending, zeros = [], 0
                        else:
                            zeros += 1
                            if zeros == 16 and k < last_new:
                                emit(0xF0)
                                pending, zeros = [], 0
                    if last_new < se:
                        emit(0x00)
            if pass_ == 0:
                dbits, dorder, dcodes = _huffman(dc_counts)
                abits, aorder, acodes = _huffman(ac_counts)
            else:
                bw.flush()
        hdr = bytearray()
        if ss == 0 and ah == 0:
            hdr += _segment(0xC4, bytes([0x00 | 0]) + bytes(dbits) + bytes(dorder))
        if se > 0:
            hdr += _segment(0xC4, bytes([0x10 | 0]) + bytes(abits) + bytes(aorder))
        sos = bytes([len(sc)]) + b"".join(bytes([ids[ci], 0x00]) for ci in sc) + bytes([ss, se, (ah << 4) | al])
        return bytes(hdr) + _segment(0xDA, sos) + bytes(bw.out)

    allc = list(range(len(comps)))
    if not progressive:
        out += scan(allc, 0, 63, 0, 0) if len(comps) > 1 else scan([0], 0, 63, 0, 0)
    else:
        out += scan(allc, 0, 0, 0, 1)             # DC, first pass, point transform 1
        for ci in allc:
            out += scan([ci], 1, 5, 0, 2)         # low AC band, first pass, point transform 2
            out += scan([ci], 6, 63, 0, 1)        # high AC band
        out += scan(allc, 0, 0, 1, 0)             # DC refinement
        for ci in allc:
            out += scan([ci], 1, 5, 2, 1)         # low AC band: two refinements
            out += scan([ci], 1, 5, 1, 0)
            out += scan([ci], 6, 63, 1, 0)        # high AC band: one
    out += b"\xff\xd9"
    return bytes(out)
