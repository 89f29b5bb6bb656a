"""A small baseline JPEG encoder for tests (ITU-T T.81): it makes the inputs
whose quantized coefficients the decoder must return exactly.

Colour to YCbCr by the JFIF formulas, chroma subsampled by averaging, edges
padded by replication, an orthonormal 8x8 DCT, quantization by the T.81 Annex
K tables scaled to a quality, and per-image optimal Huffman tables (Annex K.2,
with the 16-bit length limit of K.3).  Options cover interleaved and
one-component-per-scan layouts, restart intervals, component IDs and the JFIF
and Adobe markers.  The encoder returns the bytes and, per component, the
quantized coefficients it wrote (blocks in natural order over the component's
MCU-padded grid, zero where a scan codes no block), in the decoder's layout.
"""
import numpy as np

ZIGZAG = np.array([
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14,
    21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53,
    60, 61, 54, 47, 55, 62, 63])

LUMA_Q = np.array([
    16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56,
    14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
    49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99])
CHROMA_Q = np.full(64, 99)
CHROMA_Q[[0, 1, 2, 3, 8, 9, 10, 11, 16, 17, 18, 24, 25]] = [17, 18, 24, 47, 18, 21, 26, 66, 24, 26, 56, 47, 66]


def scaled_table(base, quality):
    q = max(1, min(100, quality))
    scale = 5000 // q if q < 50 else 200 - 2 * q
    return np.clip((base * scale + 50) // 100, 1, 255).astype(np.int64)


def dct_matrix():
    d = np.zeros((8, 8))
    for u in range(8):
        c = np.sqrt(1 / 8) if u == 0 else np.sqrt(2 / 8)
        for x in range(8):
            d[u, x] = c * np.cos((2 * x + 1) * u * np.pi / 16)
    return d


def huffman_table(freq):
    """(bits[1..16], values) of an optimal length-limited code (T.81 K.2/K.3)."""
    f = np.zeros(257, np.int64)
    for s, c in freq.items():
        f[s] = c
    f[256] = 1  # reserved so that no code is all ones
    codesize = np.zeros(257, np.int64)
    others = np.full(257, -1, np.int64)
    while True:
        nz = [i for i in range(257) if f[i] > 0]
        if len(nz) < 2:
            break
        # least frequencies, ties to the larger index (as libjpeg does)
        v1 = min(nz, key=lambda i: (f[i], -i))
        v2 = min((i for i in nz if i != v1), key=lambda i: (f[i], -i))
        f[v1] += f[v2]
        f[v2] = 0
        codesize[v1] += 1
        while others[v1] >= 0:
            v1 = others[v1]
            codesize[v1] += 1
        others[v1] = v2
        codesize[v2] += 1
        while others[v2] >= 0:
            v2 = others[v2]
            codesize[v2] += 1
    bits = np.zeros(40, np.int64)
    for i in range(257):
        if codesize[i]:
            bits[codesize[i]] += 1
    for i in range(39, 16, -1):
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
    bits[i] -= 1  # drop the reserved symbol's code
    values = []
    for size in range(1, 33):
        for s in range(256):
            if codesize[s] == size:
                values.append(s)
    return [int(b) for b in bits[1:17]], values


def canonical_codes(bits, values):
    codes, code, k = {}, 0, 0
    for length in range(1, 17):
        for _ in range(bits[length - 1]):
            codes[values[k]] = (code, length)
            code += 1
            k += 1
        code <<= 1
    return codes


def category(v):
    return 0 if v == 0 else int(abs(v)).bit_length()


def magnitude_bits(v, s):
    return v if v >= 0 else v + (1 << s) - 1


class BitWriter:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, value, length):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((value >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)
                self.acc = 0
                self.n = 0

    def flush(self):
        while self.n != 0:
            self.put(1, 1)


def _segment(marker, payload):
    return bytes([0xFF, marker]) + (len(payload) + 2).to_bytes(2, "big") + bytes(payload)


# libjpeg's jpeg_simple_progression script for three components (jcparam.c):
# (components, Ss, Se, Ah, Al).
SIMPLE_PROGRESSION = [([0, 1, 2], 0, 0, 0, 1), ([0], 1, 5, 0, 2), ([2], 1, 63, 0, 1), ([1], 1, 63, 0, 1),
                      ([0], 6, 63, 0, 2), ([0], 1, 63, 2, 1), ([0, 1, 2], 0, 0, 1, 0), ([2], 1, 63, 1, 0),
                      ([1], 1, 63, 1, 0), ([0], 1, 63, 1, 0)]
GRAY_PROGRESSION = [([0], 0, 0, 0, 1), ([0], 1, 5, 0, 2), ([0], 6, 63, 0, 2), ([0], 1, 63, 2, 1),
                    ([0], 0, 0, 1, 0), ([0], 1, 63, 1, 0)]


def _scan_units(comps, scan, mcus_x, mcus_y):
    """A scan's MCUs as lists of (component, block row, block column)."""
    units = []
    if len(scan) == 1:
        c = comps[scan[0]]
        for by in range(-(-c["ch"] // 8)):
            for bx in range(-(-c["cw"] // 8)):
                units.append([(scan[0], by, bx)])
    else:
        for my in range(mcus_y):
            for mx in range(mcus_x):
                u = []
                for k in scan:
                    c = comps[k]
                    for v in range(c["v"]):
                        for hh in range(c["h"]):
                            u.append((k, my * c["v"] + v, mx * c["h"] + hh))
                units.append(u)
    return units


def _progressive_events(comps, units, ss, se, ah, al, restart):
    """The scan's output as events: ("s", table key, symbol), ("b", value,
    bits) and ("rst",), following libjpeg's jcphuff.c encoders."""
    ev = []
    state = {"eobrun": 0, "be": []}
    key_ac = None

    def emit_eobrun():
        if state["eobrun"] > 0:
            n = state["eobrun"].bit_length() - 1
            ev.append(("s", key_ac, n << 4))
            if n:
                ev.append(("b", state["eobrun"] & ((1 << n) - 1), n))
            state["eobrun"] = 0
            ev.extend(("b", bit, 1) for bit in state["be"])
            state["be"] = []

    pred = {}
    for i, u in enumerate(units):
        if restart and i > 0 and i % restart == 0:
            emit_eobrun()
            ev.append(("rst",))
        if i == 0 or (restart and i % restart == 0):
            pred = {}
        for (k, by, bx) in u:
            c = comps[k]
            z = [int(v) for v in c["coef"][by, bx][ZIGZAG]]
            if ss == 0 and ah == 0:
                v = z[0] >> al
                diff = v - pred.get(k, 0)
                pred[k] = v
                s = category(diff)
                ev.append(("s", ("dc", c["tq"]), s))
                if s:
                    ev.append(("b", magnitude_bits(diff, s), s))
            elif ss == 0:
                ev.append(("b", (z[0] >> al) & 1, 1))
            elif ah == 0:
                key_ac = ("ac", c["tq"])
                r = 0
                for kk in range(ss, se + 1):
                    t = abs(z[kk]) >> al
                    if t == 0:
                        r += 1
                        continue
                    emit_eobrun()
                    while r > 15:
                        ev.append(("s", key_ac, 0xF0))
                        r -= 16
                    n = t.bit_length()
                    ev.append(("s", key_ac, (r << 4) | n))
                    ev.append(("b", t if z[kk] > 0 else (~t) & ((1 << n) - 1), n))
                    r = 0
                if r > 0:
                    state["eobrun"] += 1
                    if state["eobrun"] == 0x7FFF:
                        emit_eobrun()
            else:
                key_ac = ("ac", c["tq"])
                absv = {kk: abs(z[kk]) >> al for kk in range(ss, se + 1)}
                eob = max([kk for kk in absv if absv[kk] == 1], default=0)
                r, br = 0, []
                for kk in range(ss, se + 1):
                    t = absv[kk]
                    if t == 0:
                        r += 1
                        continue
                    while r > 15 and kk <= eob:
                        emit_eobrun()
                        ev.append(("s", key_ac, 0xF0))
                        r -= 16
                        ev.extend(("b", bit, 1) for bit in br)
                        br = []
                    if t > 1:
                        br.append(t & 1)
                        continue
                    emit_eobrun()
                    ev.append(("s", key_ac, (r << 4) | 1))
                    ev.append(("b", 0 if z[kk] < 0 else 1, 1))
                    ev.extend(("b", bit, 1) for bit in br)
                    br = []
                    r = 0
                if r > 0 or br:
                    state["eobrun"] += 1
                    state["be"].extend(br)
                    if state["eobrun"] == 0x7FFF or len(state["be"]) > 1000 - 64 + 1:
                        emit_eobrun()
    emit_eobrun()
    return ev


def _write_events(ev, codes):
    out = BitWriter()
    rst = 0
    for e in ev:
        if e[0] == "s":
            code, length = codes[e[1]][e[2]]
            out.put(code, length)
        elif e[0] == "b":
            out.put(e[1], e[2])
        else:
            out.flush()
            out.out += bytes([0xFF, 0xD0 + rst])
            rst = (rst + 1) & 7
    out.flush()
    return out.out


def encode(image, quality=90, sampling=(2, 2), restart=0, ids=None, jfif=True, adobe=None,
           separate_scans=False, progressive=None):
    """Encodes an (h, w, 3) RGB or (h, w) gray uint8 image; returns (bytes,
    [per-component int16 coefficient arrays of shape (blocks_h, blocks_w, 64)]).
    progressive: True for libjpeg's simple progression, or a list of
    (components, Ss, Se, Ah, Al) scans."""
    img = np.asarray(image)
    gray = img.ndim == 2
    h, w = img.shape[:2]
    if gray:
        planes = [img.astype(np.float64)]
        factors = [(1, 1)]
        ids = ids or [1]
        tables = [0]
    else:
        rgb = img.astype(np.float64)
        if adobe == 0 or ids == [ord("R"), ord("G"), ord("B")]:
            planes = [rgb[..., 0], rgb[..., 1], rgb[..., 2]]
        else:
            r, g, b = rgb[..., 0], rgb[..., 1], rgb[..., 2]
            planes = [0.299 * r + 0.587 * g + 0.114 * b,
                      -0.168736 * r - 0.331264 * g + 0.5 * b + 128,
                      0.5 * r - 0.418688 * g - 0.081312 * b + 128]
        factors = [tuple(sampling), (1, 1), (1, 1)]
        ids = ids or [1, 2, 3]
        tables = [0, 1, 1]
    hmax = max(f[0] for f in factors)
    vmax = max(f[1] for f in factors)
    mcus_x = -(-w // (8 * hmax))
    mcus_y = -(-h // (8 * vmax))
    qts = [scaled_table(LUMA_Q, quality), scaled_table(CHROMA_Q, quality)]
    d = dct_matrix()
    comps = []
    for p, (fh, fv), tq in zip(planes, factors, tables):
        # subsample by averaging, then pad by replication to the MCU grid
        sh, sv = hmax // fh, vmax // fv
        ph, pw = -(-h // sv) * sv, -(-w // sh) * sh
        pp = np.pad(p, ((0, ph - h), (0, pw - w)), mode="edge")
        sub = pp.reshape(ph // sv, sv, pw // sh, sh).mean(axis=(1, 3))
        sub = np.clip(np.round(sub), 0, 255)
        bw, bh = mcus_x * fh, mcus_y * fv
        full = np.pad(sub, ((0, bh * 8 - sub.shape[0]), (0, bw * 8 - sub.shape[1])), mode="edge")
        blocks = full.reshape(bh, 8, bw, 8).transpose(0, 2, 1, 3) - 128.0
        coef = np.einsum("ux,abxy,vy->abuv", d, blocks, d)
        q = np.round(coef / qts[tq].reshape(8, 8)).astype(np.int64)
        comps.append({"id": ids[len(comps)], "h": fh, "v": fv, "tq": tq, "coef": q.reshape(bh, bw, 64),
                      "cw": -(-w * fh // hmax), "ch": -(-h * fv // vmax)})

    # scans: all components interleaved, or one component per scan
    scans = [list(range(len(comps)))] if not separate_scans else [[k] for k in range(len(comps))]
    header = bytearray(b"\xFF\xD8")
    if jfif:
        header += _segment(0xE0, b"JFIF\x00\x01\x01\x00\x00\x01\x00\x01\x00\x00")
    if adobe is not None:
        header += _segment(0xEE, b"Adobe\x00\x64\x00\x00\x00\x00" + bytes([adobe]))
    for t in range(1 if gray else 2):
        header += _segment(0xDB, bytes([t]) + bytes(int(v) for v in qts[t][ZIGZAG]))
    sof = bytearray([8]) + h.to_bytes(2, "big") + w.to_bytes(2, "big") + bytes([len(comps)])
    for c in comps:
        sof += bytes([c["id"], (c["h"] << 4) | c["v"], c["tq"]])
    header += _segment(0xC2 if progressive else 0xC0, sof)
    if restart:
        header += _segment(0xDD, restart.to_bytes(2, "big"))
    if progressive:
        script = progressive if progressive is not True else (GRAY_PROGRESSION if gray else SIMPLE_PROGRESSION)
        written = [np.zeros_like(c["coef"]) for c in comps]
        body = bytearray()
        for scan, ss, se, ah, al in script:
            units = _scan_units(comps, scan, mcus_x, mcus_y)
            for u in units:
                for (k, by, bx) in u:
                    lo, hi = (0, 1) if ss == 0 else (ss, se + 1)
                    zz = ZIGZAG[lo:hi]
                    written[k][by, bx, zz] = comps[k]["coef"][by, bx, zz]
            ev = _progressive_events(comps, units, ss, se, ah, al, restart)
            stats = {}
            for e in ev:
                if e[0] == "s":
                    stats.setdefault(e[1], {}).setdefault(e[2], 0)
                    stats[e[1]][e[2]] += 1
            codes = {}
            dht = bytearray()
            for (kind, t), freq in sorted(stats.items()):
                bits, values = huffman_table(freq)
                codes[(kind, t)] = canonical_codes(bits, values)
                dht += bytes([(0 if kind == "dc" else 0x10) | t]) + bytes(bits) + bytes(values)
            if dht:
                body += _segment(0xC4, dht)
            sos = bytearray([len(scan)])
            for k in scan:
                sos += bytes([comps[k]["id"], (comps[k]["tq"] << 4) | comps[k]["tq"]])
            sos += bytes([ss, se, (ah << 4) | al])
            body += _segment(0xDA, sos)
            body += _write_events(ev, codes)
        return bytes(header + body + b"\xFF\xD9"), written

    written = [np.zeros_like(c["coef"]) for c in comps]
    body = bytearray()
    for scan in scans:
        # the scan's MCUs as lists of (component, block row, block column)
        units = []
        if len(scan) == 1:
            c = comps[scan[0]]
            for by in range(-(-c["ch"] // 8)):
                for bx in range(-(-c["cw"] // 8)):
                    units.append([(scan[0], by, bx)])
        else:
            for my in range(mcus_y):
                for mx in range(mcus_x):
                    u = []
                    for k in scan:
                        c = comps[k]
                        for v in range(c["v"]):
                            for hh in range(c["h"]):
                                u.append((k, my * c["v"] + v, mx * c["h"] + hh))
                    units.append(u)
        # every block's symbols (DC predictors reset at each restart)
        coded = []
        pred = {}
        for i, u in enumerate(units):
            if i == 0 or (restart and i % restart == 0):
                pred = {k: 0 for k in scan}
            for (k, by, bx) in u:
                c = comps[k]
                z = c["coef"][by, bx][ZIGZAG]
                diff = int(z[0]) - pred[k]
                pred[k] = int(z[0])
                syms = [(("dc", c["tq"]), category(diff), diff)]
                nz = [j for j in range(1, 64) if z[j] != 0]
                run = 0
                for j in range(1, (nz[-1] if nz else 0) + 1):
                    if z[j] == 0:
                        run += 1
                        continue
                    while run > 15:
                        syms.append((("ac", c["tq"]), 0xF0, None))
                        run -= 16
                    syms.append((("ac", c["tq"]), (run << 4) | category(int(z[j])), int(z[j])))
                    run = 0
                if not nz or nz[-1] < 63:
                    syms.append((("ac", c["tq"]), 0x00, None))
                coded.append((i, k, by, bx, syms))
                written[k][by, bx] = c["coef"][by, bx]
        stats = {}
        for *_, syms in coded:
            for key, sym, _ in syms:
                stats.setdefault(key, {}).setdefault(sym, 0)
                stats[key][sym] += 1
        codes = {}
        dht = bytearray()
        for (kind, t), freq in sorted(stats.items()):
            bits, values = huffman_table(freq)
            codes[(kind, t)] = canonical_codes(bits, values)
            dht += bytes([(0 if kind == "dc" else 0x10) | t]) + bytes(bits) + bytes(values)
        body += _segment(0xC4, dht)
        sos = bytearray([len(scan)])
        for k in scan:
            sos += bytes([comps[k]["id"], (comps[k]["tq"] << 4) | comps[k]["tq"]])
        sos += bytes([0, 63, 0])
        body += _segment(0xDA, sos)
        out = BitWriter()
        rst = 0
        last_unit = 0
        for i, k, by, bx, syms in coded:
            if restart and i != last_unit and i % restart == 0:
                out.flush()
                out.out += bytes([0xFF, 0xD0 + rst])
                rst = (rst + 1) & 7
            last_unit = i
            for key, sym, val in syms:
                code, length = codes[key][sym]
                out.put(code, length)
                size = sym & 15
                if size:
                    out.put(magnitude_bits(val, size), size)
        out.flush()
        body += out.out
    return bytes(header + body + b"\xFF\xD9"), written
