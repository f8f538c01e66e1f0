"""Small PNG and baseline-JPEG encoders that write the texture fixtures in
tests/golden/tex/ (test data for the image decoders; not used at test time).

PNG: ISO/IEC 15948 chunks, zlib-compressed scanlines with a chosen filter
per row, optional Adam7 interlacing.  JPEG: ITU-T T.81 baseline sequential
Huffman: JFIF YCbCr, chroma averaged over the sampling block, float forward
DCT, quantization by the Annex K example tables scaled by quality, Huffman
tables built from the symbol counts (code lengths limited to 16 as in
Annex K.2), optional restart interval and one scan per component.
"""
import heapq
import struct
import zlib

import numpy as np

# ------------------------------------------------------------------ PNG

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def _chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def _filter_rows(px, bpp, filters):
    """px: (rows, bytes) uint8 -> filtered bytes with a filter byte per row."""
    out = bytearray()
    prev = np.zeros(px.shape[1], np.int32)
    for y in range(px.shape[0]):
        x = px[y].astype(np.int32)
        a = np.concatenate([np.zeros(bpp, np.int32), x[:-bpp]])
        c = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        b = prev
        f = filters[y % len(filters)]
        if f == 0:
            r = x
        elif f == 1:
            r = x - a
        elif f == 2:
            r = x - b
        elif f == 3:
            r = x - (a + b) // 2
        else:
            p = a + b - c
            pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))
            r = x - pred
        out.append(f)
        out += (r & 0xFF).astype(np.uint8).tobytes()
        prev = x
    return bytes(out)


def png_encode(samples, ctype, palette=None, trns=None, interlace=False, filters=(0, 1, 2, 3, 4)):
    """samples: (h, w, spp) uint8 (palette indices for ctype 3)."""
    h, w, spp = samples.shape
    raw = b""
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        raw += _filter_rows(sub.reshape(sub.shape[0], -1), spp, filters)
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, int(interlace)))
    if palette is not None:
        out += _chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes())
    if trns is not None:
        out += _chunk(b"tRNS", np.asarray(trns, np.uint8).tobytes())
    z = zlib.compress(raw, 9)
    # split the stream over two IDAT chunks
    out += _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])
    return out + _chunk(b"IEND", b"")


# ------------------------------------------------------------------ JPEG

ZIGZAG = np.array([0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20,
                   13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52,
                   45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63])

# T.81 Annex K, Tables K.1 / K.2 (example luminance / chrominance tables)
Q_LUMA = np.array([16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56,
                   14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
                   49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99]).reshape(8, 8)
Q_CHROMA = np.array([17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99, 99, 99,
                     47, 66, 99, 99, 99, 99, 99, 99] + [99] * 32).reshape(8, 8)


def _qtable(base, quality):
    s = 5000 // quality if quality < 50 else 200 - 2 * quality
    return np.clip((base * s + 50) // 100, 1, 255)


_C = np.array([[(np.sqrt(0.5) if u == 0 else 1.0) * 0.5 * np.cos((2 * x + 1) * u * np.pi / 16) for x in range(8)]
               for u in range(8)])


def _fdct(block):
    return _C @ block @ _C.T


def _huff_lengths(freq):
    """Code length per symbol, limited to 16 bits (Annex K.2 adjustment);
    a reserved pseudo-symbol keeps the all-ones code unused."""
    syms = sorted(s for s, f in freq.items() if f > 0)
    if not syms:
        syms = [0]
    items = [(freq.get(s, 1), i, [s]) for i, s in enumerate(syms)] + [(0, len(syms), [-1])]
    heapq.heapify(items)
    length = {s: 0 for s in syms + [-1]}
    k = len(items)
    while len(items) > 1:
        f1, _, a = heapq.heappop(items)
        f2, _, b = heapq.heappop(items)
        for s in a + b:
            length[s] += 1
        heapq.heappush(items, (f1 + f2, k, a + b))
        k += 1
    bits = [0] * 33
    for s, L in length.items():
        bits[L] += 1
    for i in range(32, 16, -1):  # K.2 Figure K.3: move codes longer than 16 up
        while bits[i] > 0:
            j = i - 2
            while bits[j] == 0:
                j -= 1
            bits[i] -= 2
            bits[i - 1] += 1
            bits[j + 1] += 2
            bits[j] -= 1
    i = 16
    while bits[i] == 0:
        i -= 1
    bits[i] -= 1  # drop the reserved pseudo-symbol
    order = sorted(syms, key=lambda s: (length[s], s))
    return bits[1:17], order


def _huff_codes(bits, order):
    codes, code, k = {}, 0, 0
    for L in range(1, 17):
        for _ in range(bits[L - 1]):
            codes[order[k]] = (code, L)
            code += 1
            k += 1
        code <<= 1
    return codes


class _Bits:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, v, n):
        for i in range(n - 1, -1, -1):
            self.acc = (self.acc << 1) | ((v >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)
                self.acc, self.n = 0, 0

    def flush(self):
        while self.n:
            self.put(1, 1)


def _magnitude(v):
    s = int(abs(v)).bit_length()
    return s, (v if v >= 0 else v + (1 << s) - 1)


def jpeg_encode(rgb, sampling=(2, 2), quality=85, restart=0, separate_scans=False, grey=False):
    rgb = np.asarray(rgb, np.float64)
    H, W = rgb.shape[:2]
    if grey:
        planes = [rgb[..., 0]]
        fac = [(1, 1)]
    else:
        R, G, B = rgb[..., 0], rgb[..., 1], rgb[..., 2]
        Y = 0.299 * R + 0.587 * G + 0.114 * B
        Cb = -0.168736 * R - 0.331264 * G + 0.5 * B + 128
        Cr = 0.5 * R - 0.418688 * G - 0.081312 * B + 128
        planes = [Y, Cb, Cr]
        fac = [sampling, (1, 1), (1, 1)]
    hmax = max(f[0] for f in fac)
    vmax = max(f[1] for f in fac)
    mcux, mcuy = -(-W // (8 * hmax)), -(-H // (8 * vmax))
    qts = [_qtable(Q_LUMA, quality), _qtable(Q_CHROMA, quality)]
    comps = []
    for ci, (p, (h, v)) in enumerate(zip(planes, fac)):
        PW, PH = mcux * 8 * hmax, mcuy * 8 * vmax
        p = np.pad(p, ((0, PH - H), (0, PW - W)), mode="edge")
        fx, fy = hmax // h, vmax // v
        p = p.reshape(PH // fy, fy, PW // fx, fx).mean(axis=(1, 3))
        q = qts[0 if ci == 0 else 1]
        nby, nbx = p.shape[0] // 8, p.shape[1] // 8
        blocks = np.zeros((nby, nbx, 64), np.int32)
        for by in range(nby):
            for bx in range(nbx):
                F = _fdct(p[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] - 128.0)
                blocks[by, bx] = np.round(F / q).astype(np.int32).reshape(64)[ZIGZAG]
        comps.append(dict(id=ci + 1, h=h, v=v, tq=0 if ci == 0 else 1, blocks=blocks,
                          td=0 if ci == 0 else 1, ta=0 if ci == 0 else 1))

    def block_order(cs):
        """(component, block) in scan order; None marks a restart point."""
        seq = []
        if len(cs) == 1:
            c = cs[0]
            cols = -(-(-(-W * c["h"] // hmax)) // 8)
            rows = -(-(-(-H * c["v"] // vmax)) // 8)
            units = [[(c, by, bx)] for by in range(rows) for bx in range(cols)]
        else:
            units = [[(c, my * c["v"] + v, mx * c["h"] + h) for c in cs for v in range(c["v"]) for h in range(c["h"])]
                     for my in range(mcuy) for mx in range(mcux)]
        for i, u in enumerate(units):
            if restart and i and i % restart == 0:
                seq.append(None)
            seq.extend(u)
        return seq

    scans = [[c] for c in comps] if separate_scans else [comps]
    # symbol statistics per table over every scan
    freq = {("dc", 0): {}, ("dc", 1): {}, ("ac", 0): {}, ("ac", 1): {}}

    def symbols(seq):
        pred, out = {}, []
        for e in seq:
            if e is None:
                pred = {}
                out.append(None)
                continue
            c, by, bx = e
            blk = c["blocks"][by, bx]
            diff = int(blk[0]) - pred.get(c["id"], 0)
            pred[c["id"]] = int(blk[0])
            s, m = _magnitude(diff)
            syms = [("dc", c["td"], s, m, s)]
            run = 0
            last = max([k for k in range(1, 64) if blk[k] != 0], default=0)
            for k in range(1, last + 1):
                if blk[k] == 0:
                    run += 1
                    continue
                while run > 15:
                    syms.append(("ac", c["ta"], 0xF0, 0, 0))
                    run -= 16
                s, m = _magnitude(int(blk[k]))
                syms.append(("ac", c["ta"], (run << 4) | s, m, s))
                run = 0
            if last < 63:
                syms.append(("ac", c["ta"], 0x00, 0, 0))
            out.append(syms)
        return out

    scan_syms = [symbols(block_order(cs)) for cs in scans]
    for ss in scan_syms:
        for syms in ss:
            for t in syms or []:
                d = freq[(t[0], t[1])]
                d[t[2]] = d.get(t[2], 0) + 1
    tables = {k: _huff_lengths(v) for k, v in freq.items() if v}
    codes = {k: _huff_codes(*v) for k, v in tables.items()}

    out = bytearray(b"\xff\xd8")
    out += b"\xff\xe0" + struct.pack(">H", 16) + b"JFIF\x00\x01\x02\x00\x00\x01\x00\x01\x00\x00"
    for i, q in enumerate(qts[: 1 if grey else 2]):
        out += b"\xff\xdb" + struct.pack(">HB", 67, i) + bytes(int(v) for v in q.reshape(64)[ZIGZAG])
    out += b"\xff\xc0" + struct.pack(">HBHHB", 8 + 3 * len(comps), 8, H, W, len(comps))
    for c in comps:
        out += bytes([c["id"], (c["h"] << 4) | c["v"], c["tq"]])
    for (kind, th), (bits, order) in tables.items():
        body = bytes([(0 if kind == "dc" else 1) << 4 | th]) + bytes(bits) + bytes(order)
        out += b"\xff\xc4" + struct.pack(">H", 2 + len(body)) + body
    if restart:
        out += b"\xff\xdd" + struct.pack(">HH", 4, restart)
    for cs, ss in zip(scans, scan_syms):
        out += b"\xff\xda" + struct.pack(">HB", 6 + 2 * len(cs), len(cs))
        for c in cs:
            out += bytes([c["id"], (c["td"] << 4) | c["ta"]])
        out += bytes([0, 63, 0])
        bw = _Bits()
        rst = 0
        for syms in ss:
            if syms is None:
                bw.flush()
                out += bw.out
                out += bytes([0xFF, 0xD0 + rst])
                rst = (rst + 1) & 7
                bw = _Bits()
                continue
            for kind, th, sym, m, s in syms:
                code, L = codes[(kind, th)][sym]
                bw.put(code, L)
                if s:
                    bw.put(m, s)
        bw.flush()
        out += bw.out
    out += b"\xff\xd9"
    return bytes(out)
