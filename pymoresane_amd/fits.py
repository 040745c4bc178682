"""Minimal FITS primary-HDU reader/writer (astropy is not required).

Covers what FitsImage needs (pymoresane/main.py:618-646, 678-733): the primary
header as an ordered card list with typed values, and the data array for
BITPIX 8/16/32/64/-32/-64 in FITS big-endian order, numpy shape
(NAXISn, ..., NAXIS1), with BSCALE/BZERO applied.
"""

import numpy as np

BLOCK = 2880
CARD = 80
_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


class Header(object):
    """Ordered FITS header cards with dict-style access to values."""

    def __init__(self, cards=None):
        self.cards = list(cards or [])  # (key, value, comment)

    def __contains__(self, key):
        return any(k == key for k, _, _ in self.cards)

    def __getitem__(self, key):
        for k, v, _ in self.cards:
            if k == key:
                return v
        raise KeyError(key)

    def get(self, key, default=None):
        try:
            return self[key]
        except KeyError:
            return default

    def set(self, key, value, comment=""):
        for i, (k, _, c) in enumerate(self.cards):
            if k == key:
                self.cards[i] = (k, value, comment or c)
                return
        self.cards.append((key, value, comment))

    def __setitem__(self, key, value):
        self.set(key, value)

    def copy(self):
        return Header(self.cards)


def _parse_value(raw):
    s = raw.strip()
    if not s:
        return None
    if s.startswith("'"):
        end = 1
        out = []
        while end < len(s):
            if s[end] == "'":
                if end + 1 < len(s) and s[end + 1] == "'":
                    out.append("'")
                    end += 2
                    continue
                break
            out.append(s[end])
            end += 1
        return "".join(out).rstrip()
    v = s.split("/")[0].strip()
    if v == "T":
        return True
    if v == "F":
        return False
    try:
        return int(v)
    except ValueError:
        pass
    try:
        return float(v.replace("D", "E"))
    except ValueError:
        return v


def read_header(f):
    cards = []
    while True:
        block = f.read(BLOCK)
        if len(block) < BLOCK:
            raise ValueError("truncated FITS header")
        text = block.decode("ascii", "replace")
        for i in range(0, BLOCK, CARD):
            card = text[i:i + CARD]
            key = card[:8].strip()
            if key == "END":
                return Header(cards)
            if card[8:10] == "= ":
                body = card[10:]
                comment = ""
                if not body.strip().startswith("'") and "/" in body:
                    body, comment = body.split("/", 1)
                cards.append((key, _parse_value(body), comment.strip()))
            elif key:
                cards.append((key, None, card[8:].rstrip()))


def read(path):
    """(header, data) of the primary HDU."""
    with open(path, "rb") as f:
        hdr = read_header(f)
        naxis = int(hdr.get("NAXIS", 0))
        shape = tuple(int(hdr["NAXIS%d" % (i + 1)]) for i in range(naxis))[::-1]
        bitpix = int(hdr["BITPIX"])
        if bitpix not in _DTYPES:
            raise ValueError("unsupported BITPIX %d" % bitpix)
        count = int(np.prod(shape)) if shape else 0
        data = np.frombuffer(f.read(count * abs(bitpix) // 8), dtype=_DTYPES[bitpix]).reshape(shape)
    scale, zero = hdr.get("BSCALE", 1.0), hdr.get("BZERO", 0.0)
    data = data.astype(np.float64 if bitpix == -64 else np.float32 if bitpix == -32 else np.float64)
    if scale != 1.0 or zero != 0.0:
        data = data * scale + zero
    return hdr, data


def _format_value(v):
    if isinstance(v, bool):
        return "%20s" % ("T" if v else "F")
    if isinstance(v, (int, np.integer)):
        return "%20d" % v
    if isinstance(v, (float, np.floating)):
        return "%20s" % repr(float(v)).upper()
    if v is None:
        return " " * 20
    s = str(v).replace("'", "''")
    return "'%-8s'" % s


def write(path, data, header=None):
    """Write a float32 primary HDU (BITPIX -32), copying WCS-like cards from ``header``."""
    data = np.ascontiguousarray(data, dtype=">f4")
    cards = [("SIMPLE", True, ""), ("BITPIX", -32, ""), ("NAXIS", data.ndim, "")]
    for i, n in enumerate(data.shape[::-1]):
        cards.append(("NAXIS%d" % (i + 1), int(n), ""))
    skip = {"SIMPLE", "BITPIX", "NAXIS", "EXTEND", "BSCALE", "BZERO", "END"}
    if header is not None:
        for k, v, c in header.cards:
            if k in skip or (k.startswith("NAXIS") and k[5:].isdigit()):
                continue
            cards.append((k, v, c))
    lines = []
    for k, v, c in cards:
        if v is None and k in ("COMMENT", "HISTORY", ""):
            lines.append(("%-8s%s" % (k, c))[:CARD].ljust(CARD))
            continue
        body = "%-8s= %s" % (k[:8], _format_value(v))
        if c:
            body += " / " + c
        lines.append(body[:CARD].ljust(CARD))
    lines.append("END".ljust(CARD))
    text = "".join(lines)
    text += " " * ((-len(text)) % BLOCK)
    raw = data.tobytes()
    raw += b"\0" * ((-len(raw)) % BLOCK)
    with open(path, "wb") as f:
        f.write(text.encode("ascii"))
        f.write(raw)
