"""FITS product assembly after the core, without astropy.

The reference engine turns each saved iteration into a five-HDU FITS file
(updc/_engine.py:824-861): a primary HDU with the run metadata, SCI and WHT
image extensions (target WCS header plus metadata), and the UPDC_BKG and
UPDC_PSFGEOM binary tables, written with CHECKSUM/DATASUM
(_engine.py:1421-1428) and then re-opened and validated
(_engine.py:1432-1446).  This module writes and validates the same products
with numpy alone, following the FITS standard (version 4.0): 80-byte
header cards in 2880-byte blocks, big-endian data, and the 32-bit ones'
complement checksum with its 16-character ASCII encoding (the FITS checksum
convention of Seaman, Pence and Rots).

Header cards are (keyword, value, comment) triples; ("HISTORY", text, "")
and ("COMMENT", text, "") are commentary cards.  Values are bool, int, float
or str; strings longer than one card continue on CONTINUE cards (the long
string convention).
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Iterable, Sequence

import numpy as np

BLOCK = 2880
Card = tuple  # (keyword, value, comment)


# ---------------------------------------------------------------- header cards
def _fmt_float(v: float) -> str:
    if not np.isfinite(v):
        raise ValueError("FITS header values must be finite")
    s = repr(float(v)).upper()
    if "E" in s:
        m, e = s.split("E")
        if "." not in m:
            m += ".0"
        s = f"{m}E{e}"
    elif "." not in s:
        s += ".0"
    if len(s) > 20:  # the fixed-format field is 20 characters
        s = f"{v:.12E}"
    return s


def _fmt_value(v) -> str:
    """Fixed-format value field (columns 11-30) of a numeric or logical card."""
    if isinstance(v, (bool, np.bool_)):
        return f"{'T' if v else 'F':>20}"
    if isinstance(v, (int, np.integer)):
        return f"{int(v):>20}"
    if isinstance(v, (float, np.floating)):
        return f"{_fmt_float(float(v)):>20}"
    raise TypeError(f"unsupported FITS value {v!r}")


def _quote(s: str) -> str:
    s = s.replace("'", "''")
    return "'" + s.ljust(8) + "'"


def _card(text: str) -> str:
    if len(text) > 80 or any(not (32 <= ord(ch) <= 126) for ch in text):
        raise ValueError(f"invalid FITS card: {text!r}")
    return text.ljust(80)


def format_card(key: str, value, comment: str = "") -> list[str]:
    """The 80-character card(s) of one keyword."""
    key = key.upper()
    if len(key) > 8:
        raise ValueError(f"keyword longer than 8 characters: {key}")
    if key in ("HISTORY", "COMMENT", ""):
        text = str(value)
        return [_card(key.ljust(8) + text[i:i + 72]) for i in range(0, max(len(text), 1), 72)]
    head = key.ljust(8) + "= "
    if isinstance(value, str):
        # one card while it fits, else '&'-terminated pieces and CONTINUE cards
        if len(_quote(value)) <= 70 and len(head + _quote(value)) <= 80:
            text = head + _quote(value)
            if comment and len(text) + 3 < 80:
                text = (text + " / " + comment)[:80]
            return [_card(text)]
        pieces, rest = [], value
        while rest:
            take = 67
            while take > 1 and len(rest[:take].replace("'", "''")) > 67:
                take -= 1
            pieces.append(rest[:take])
            rest = rest[take:]
        cards = []
        for i, p in enumerate(pieces):
            last = i == len(pieces) - 1
            body = "'" + p.replace("'", "''") + ("" if last else "&") + "'"
            cards.append(_card((head if i == 0 else "CONTINUE  ") + body))
        return cards
    text = head + _fmt_value(value)
    if comment:
        text = (text + " / " + comment)[:80]
    return [_card(text)]


def header_bytes(cards: Iterable[Card]) -> bytes:
    lines = []
    for c in cards:
        lines += format_card(*c) if len(c) == 3 else format_card(c[0], c[1])
    lines.append("END".ljust(80))
    raw = "".join(lines).encode("ascii")
    return raw + b" " * (-len(raw) % BLOCK)


def _pad(data: bytes, fill: bytes = b"\0") -> bytes:
    return data + fill * (-len(data) % BLOCK)


# ---------------------------------------------------------------- checksums
def ones_complement_sum(data: bytes, start: int = 0) -> int:
    """32-bit ones' complement sum of big-endian words (the length is a
    multiple of 4, as every FITS block is), with end-around carry."""
    if len(data) % 4:
        raise ValueError("checksummed length must be a multiple of 4")
    w = np.frombuffer(data, dtype=">u4").astype(np.uint64)
    s = int(start)
    for i in range(0, w.size, 1 << 24):  # partial sums stay below 2**57
        s += int(w[i:i + (1 << 24)].sum())
    while s >> 32:
        s = (s & 0xFFFFFFFF) + (s >> 32)
    return s


_EXCLUDE = set(range(0x3A, 0x41)) | set(range(0x5B, 0x61))


def encode_checksum(value: int, complement: bool = True) -> str:
    """The 16-character ASCII form of a 32-bit checksum: each byte spread over
    four characters offset from '0', punctuation avoided by moving one unit
    between the two characters of a pair, then rotated right by one."""
    if complement:
        value = ~value & 0xFFFFFFFF
    asc = [0] * 16
    for i in range(4):
        byte = (value >> (24 - 8 * i)) & 0xFF
        q, r = byte // 4 + 0x30, byte % 4
        ch = [q + r, q, q, q]
        changed = True
        while changed:
            changed = False
            for j in (0, 2):
                if ch[j] in _EXCLUDE or ch[j + 1] in _EXCLUDE:
                    ch[j] += 1
                    ch[j + 1] -= 1
                    changed = True
        for j in range(4):
            asc[4 * j + i] = ch[j]
    asc = asc[15:] + asc[:15]
    return bytes(asc).decode("ascii")


def hdu_bytes(cards: Sequence[Card], data: bytes, checksum: bool = True) -> bytes:
    """Header + data of one HDU, with DATASUM and CHECKSUM so that the ones'
    complement sum of the whole HDU is all ones."""
    data = _pad(data)
    if not checksum:
        return header_bytes(cards) + data
    dsum = ones_complement_sum(data)
    base = list(cards) + [("CHECKSUM", "0" * 16, "HDU checksum"), ("DATASUM", str(dsum), "data unit checksum")]
    hsum = ones_complement_sum(header_bytes(base))
    total = hsum + dsum
    while total >> 32:
        total = (total & 0xFFFFFFFF) + (total >> 32)
    base[-2] = ("CHECKSUM", encode_checksum(total), "HDU checksum")
    return header_bytes(base) + data


# ---------------------------------------------------------------- HDUs
def primary_hdu(cards: Sequence[Card]) -> bytes:
    head = [("SIMPLE", True, "conforms to FITS standard"), ("BITPIX", 8, "array data type"),
            ("NAXIS", 0, "number of array dimensions"), ("EXTEND", True, "")]
    return hdu_bytes(head + list(cards), b"")


def image_hdu(data: np.ndarray, cards: Sequence[Card]) -> bytes:
    """IMAGE extension of a 2-D float32 plane (BITPIX -32, big-endian)."""
    a = np.ascontiguousarray(data, dtype=">f4")
    if a.ndim != 2:
        raise ValueError("image extensions here are 2-D")
    head = [("XTENSION", "IMAGE", "Image extension"), ("BITPIX", -32, "array data type"),
            ("NAXIS", 2, "number of array dimensions"), ("NAXIS1", a.shape[1], ""), ("NAXIS2", a.shape[0], ""),
            ("PCOUNT", 0, "number of parameters"), ("GCOUNT", 1, "number of groups")]
    return hdu_bytes(head + list(cards), a.tobytes())


_TFORM = {"J": ">i4", "K": ">i8", "D": ">f8", "E": ">f4", "L": "S1"}


@dataclass
class Column:
    name: str
    format: str       # 'J', 'K', 'D', 'E' or '<n>A'
    array: object


def _col_dtype(fmt: str):
    if fmt.endswith("A"):
        return f"S{int(fmt[:-1] or 1)}"
    return _TFORM[fmt]


def bintable_hdu(columns: Sequence[Column], extname: str, cards: Sequence[Card] = ()) -> bytes:
    """BINTABLE extension: one row per element, big-endian fields in column order."""
    n = len(np.asarray(columns[0].array))
    dt = np.dtype([(c.name, _col_dtype(c.format)) for c in columns])
    rows = np.zeros(n, dtype=dt)
    for c in columns:
        a = np.asarray(c.array)
        if len(a) != n:
            raise ValueError("table columns differ in length")
        if c.format.endswith("A"):
            rows[c.name] = np.array([str(x).encode("ascii") for x in a], dtype=_col_dtype(c.format))
        else:
            rows[c.name] = a
    head = [("XTENSION", "BINTABLE", "binary table extension"), ("BITPIX", 8, "array data type"),
            ("NAXIS", 2, "number of array dimensions"), ("NAXIS1", dt.itemsize, "length of dimension 1"),
            ("NAXIS2", n, "length of dimension 2"), ("PCOUNT", 0, "number of group parameters"),
            ("GCOUNT", 1, "number of groups"), ("TFIELDS", len(columns), "number of table fields")]
    for i, c in enumerate(columns, 1):
        head += [(f"TTYPE{i}", c.name, ""), (f"TFORM{i}", c.format, "")]
    head.append(("EXTNAME", extname, "extension name"))
    return hdu_bytes(head + list(cards), rows.tobytes())


# ---------------------------------------------------------------- the product
def _set(cards: list, key: str, value, comment: str = "") -> None:
    """astropy's header[key] = value: replace the first card of that keyword,
    else append."""
    for i, c in enumerate(cards):
        if c[0].upper() == key and key not in ("HISTORY", "COMMENT"):
            cards[i] = (key, value, comment or c[2])
            return
    cards.append((key, value, comment))


def _extend(cards: list, more: Iterable[Card]) -> None:
    """astropy's header.extend(other, update=True): existing keywords take the
    new value in place, commentary cards and new keywords are appended."""
    for c in more:
        if c[0].upper() in ("HISTORY", "COMMENT"):
            cards.append(c)
        else:
            _set(cards, c[0].upper(), c[1], c[2] if len(c) > 2 else "")


def psf_geometry_columns(psf_records: Sequence[dict], psf_names: Sequence[str]) -> list[Column]:
    """UPDC_PSFGEOM columns (_engine.py:786-821)."""
    def arr(key, dtype=np.float64):
        return np.asarray([r[key] for r in psf_records], dtype=dtype)
    names = list(psf_names)
    methods = [str(r.get("method", "")) for r in psf_records]
    wcs = [str(r.get("wcs_used", "")) for r in psf_records]
    w = lambda xs: f"{max(1, max(map(len, xs)))}A"
    return [
        Column("EXPOSURE", "J", np.arange(1, len(psf_records) + 1, dtype=np.int32)),
        Column("PSFNAME", w(names), names),
        Column("ANCHOR_X", "D", arr("anchor_x")), Column("ANCHOR_Y", "D", arr("anchor_y")),
        Column("OVERSAMP", "J", arr("oversample", np.int32)),
        Column("CENTER_X", "D", arr("center_target_x")), Column("CENTER_Y", "D", arr("center_target_y")),
        Column("EXTENT_X", "D", arr("extent_x")), Column("EXTENT_Y", "D", arr("extent_y")),
        Column("IN_W", "J", arr("input_width", np.int32)), Column("IN_H", "J", arr("input_height", np.int32)),
        Column("MAP_W", "J", arr("mapped_width", np.int32)), Column("MAP_H", "J", arr("mapped_height", np.int32)),
        Column("OUT_W", "J", arr("output_width", np.int32)), Column("OUT_H", "J", arr("output_height", np.int32)),
        Column("INPUT_SUM", "D", arr("input_sum")), Column("OUTPUT_SUM", "D", arr("output_sum")),
        Column("CROP_RETN", "D", arr("crop_retained_flux")), Column("CROP_LOSS", "D", arr("crop_discarded_flux")),
        Column("FLUX_ERR", "D", arr("flux_rel_error")), Column("OVERLAPS", "K", arr("overlaps", np.int64)),
        Column("METHOD", w(methods), methods), Column("WCS_USED", w(wcs), wcs),
    ]


def background_columns(input_metrics: Sequence[dict]) -> list[Column]:
    """UPDC_BKG columns (_engine.py:849-859)."""
    names = [Path(m["source"]).name for m in input_metrics]
    return [
        Column("EXPOSURE", "J", np.arange(1, len(input_metrics) + 1, dtype=np.int32)),
        Column("INPUTFILE", f"{max(map(len, names))}A", names),
        Column("BACKGROUND", "D", [m["background"] for m in input_metrics]),
        Column("BKG_SIGMA", "D", [m["background_sigma"] for m in input_metrics]),
        Column("POIS_PEDESTAL", "D", [m["poisson_pedestal"] for m in input_metrics]),
        Column("AUTO_PEDESTAL", "D", [m["auto_positivity_pedestal"] for m in input_metrics]),
        Column("N_NONPOS", "K", np.array([m["nonpositive_valid_before_pedestal"] for m in input_metrics],
                                         dtype=np.int64)),
        Column("SKY_VAR_POISSON", "D", [m["sky_var_poisson"] for m in input_metrics]),
    ]


def iteration_product(data: np.ndarray, wht: np.ndarray, target_cards: Sequence[Card],
                      metadata: Sequence[Card], input_metrics: Sequence[dict], iteration: int, filename: str,
                      weight_mode: str, freeze_rows: Sequence[dict], psf_records: Sequence[dict],
                      psf_names: Sequence[str]) -> bytes:
    """The bytes of one iteration's FITS product, HDU by HDU as
    _engine.py:824-861 assembles it."""
    tot = sum(int(r["n_total_frozen"]) for r in freeze_rows)
    new = sum(int(r["n_new_frozen"]) for r in freeze_rows)
    prim = []
    _extend(prim, metadata)
    _set(prim, "FILENAME", filename)
    _set(prim, "UPDCITER", int(iteration))
    _set(prim, "MUFRTOT", tot, "Frozen detector pixels at this iteration")
    _set(prim, "MUFRNEW", new, "Newly frozen detector pixels")

    def ext_cards(extname):
        c = list(target_cards)
        _set(c, "EXTNAME", extname, "extension name")
        _extend(c, metadata)
        _set(c, "FILENAME", filename)
        _set(c, "UPDCITER", int(iteration))
        return c

    sci = ext_cards("SCI")
    _set(sci, "MUFRTOT", tot, "Frozen detector pixels at this iteration")
    _set(sci, "MUFRNEW", new, "Newly frozen detector pixels")
    _set(sci, "SCIMODE", "SIGNED-X", "Signed multiplicative reconstruction")
    wcards = ext_cards("WHT")
    unit = next((str(c[1]) for c in target_cards if c[0].upper() == "BUNIT"), "MJy/sr")
    _set(wcards, "BUNIT", "1" if weight_mode == "mask" else f"({unit})^-2")
    _set(wcards, "WHTTYPE", "SUM(W*A/J)")
    return b"".join([
        primary_hdu(prim),
        image_hdu(data, sci),
        image_hdu(wht, wcards),
        bintable_hdu(background_columns(input_metrics), "UPDC_BKG"),
        bintable_hdu(psf_geometry_columns(psf_records, psf_names), "UPDC_PSFGEOM"),
    ])


def write_iteration_product(path: Path, *args, **kw) -> Path:
    """Write iteration_product(...) to `path` (through a temporary file and a rename)."""
    path = Path(path)
    tmp = path.with_name(path.name + ".tmp")
    tmp.write_bytes(iteration_product(*args, **kw))
    tmp.replace(path)
    return path


# ---------------------------------------------------------------- reading and validation
@dataclass
class Hdu:
    cards: list        # (keyword, raw value text) in order
    header: dict       # keyword -> parsed value (first occurrence)
    start: int         # byte offset of the header
    data_start: int
    data_bytes: int    # unpadded
    end: int           # offset after the padded data


def _parse_value(text: str):
    t = text.strip()
    if t.startswith("'"):
        out, i = [], 1
        while i < len(t):
            if t[i] == "'":
                if i + 1 < len(t) and t[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(t[i])
            i += 1
        return "".join(out).rstrip()
    v = t.split("/")[0].strip()
    if v in ("T", "F"):
        return v == "T"
    try:
        return int(v)
    except ValueError:
        return float(v.replace("D", "E"))


def read_hdus(raw: bytes) -> list[Hdu]:
    """Split a FITS file into HDUs (headers parsed, data located)."""
    out, pos = [], 0
    while pos < len(raw):
        cards, header, p = [], {}, pos
        while True:
            block = raw[p:p + BLOCK]
            if len(block) < BLOCK:
                raise ValueError("truncated FITS header")
            p += BLOCK
            done = False
            for k in range(0, BLOCK, 80):
                card = block[k:k + 80].decode("ascii")
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= ":
                    cards.append((key, card[10:]))
                    header.setdefault(key, _parse_value(card[10:]))
                elif key:
                    cards.append((key, card[8:].rstrip()))
            if done:
                break
        # data size: |BITPIX|/8 x GCOUNT x (PCOUNT + NAXIS1 x ... x NAXISn)
        naxis = header.get("NAXIS", 0)
        nbytes = 0
        if naxis:
            prod = 1
            for i in range(1, naxis + 1):
                prod *= header[f"NAXIS{i}"]
            nbytes = abs(header["BITPIX"]) // 8 * header.get("GCOUNT", 1) * (header.get("PCOUNT", 0) + prod)
        end = p + nbytes + (-nbytes % BLOCK)
        out.append(Hdu(cards, header, pos, p, nbytes, end))
        pos = end
    return out


def image_data(raw: bytes, h: Hdu) -> np.ndarray:
    return np.frombuffer(raw, dtype=">f4", count=h.header["NAXIS1"] * h.header["NAXIS2"],
                         offset=h.data_start).reshape(h.header["NAXIS2"], h.header["NAXIS1"])


def table_rows(raw: bytes, h: Hdu) -> np.ndarray:
    n = h.header["TFIELDS"]
    dt = np.dtype([(h.header[f"TTYPE{i}"], _col_dtype(h.header[f"TFORM{i}"])) for i in range(1, n + 1)])
    if dt.itemsize != h.header["NAXIS1"]:
        raise ValueError("TFORM widths disagree with NAXIS1")
    return np.frombuffer(raw, dtype=dt, count=h.header["NAXIS2"], offset=h.data_start)


def verify_checksums(raw: bytes, h: Hdu) -> tuple[bool, bool]:
    """(CHECKSUM holds, DATASUM holds) for one HDU."""
    data = raw[h.data_start:h.end]
    ds = ones_complement_sum(data)
    ok_data = str(h.header.get("DATASUM", "")) == str(ds)
    ok_hdu = ones_complement_sum(raw[h.start:h.end]) == 0xFFFFFFFF
    return ok_hdu, ok_data


def validate_product(path: Path, exposures: int) -> None:
    """The engine's checks of a written product (_engine.py:1432-1446): finite
    SCI, finite non-negative WHT, one table row per exposure, and valid
    CHECKSUM/DATASUM on every extension.  Raises ValueError."""
    raw = Path(path).read_bytes()
    hdus = {h.header.get("EXTNAME", "PRIMARY"): h for h in read_hdus(raw)}
    for name in ("SCI", "WHT", "UPDC_BKG", "UPDC_PSFGEOM"):
        if name not in hdus:
            raise ValueError(f"missing {name} HDU: {path}")
    if not np.isfinite(image_data(raw, hdus["SCI"])).all():
        raise ValueError(f"SCI contains nonfinite values: {path}")
    w = image_data(raw, hdus["WHT"])
    if not np.isfinite(w).all() or np.any(w < 0):
        raise ValueError(f"WHT contains invalid values: {path}")
    if len(table_rows(raw, hdus["UPDC_BKG"])) != exposures:
        raise ValueError(f"UPDC_BKG row mismatch: {path}")
    if len(table_rows(raw, hdus["UPDC_PSFGEOM"])) != exposures:
        raise ValueError(f"UPDC_PSFGEOM row mismatch: {path}")
    for name in ("SCI", "WHT", "UPDC_BKG", "UPDC_PSFGEOM"):
        if verify_checksums(raw, hdus[name]) != (True, True):
            raise ValueError(f"{name} FITS checksum failure: {path}")
