"""Seeded cases for the host-format utilities: GUPPI RAW headers (80-byte
cards) in the forms real and hand-edited headers take, and filterbank header
fields; plus helpers that ask a library (this one, or the reference's own C
built into oracle/_ref) for its answers."""
from __future__ import annotations

import ctypes as C
import os
import random
import string

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

INT_KEYS = ["BLOCSIZE", "NPOL", "OBSNCHAN", "NBITS", "DIRECTIO", "PKTIDX", "BEAM_ID", "NBEAM",
            "REFBEAM", "NANTS", "STT_IMJD", "STT_SMJD", "PKTSIZE", "NPKT"]
DBL_KEYS = ["OBSFREQ", "OBSBW", "TBIN", "CHAN_BW", "AZ", "ZA"]
STR_KEYS = ["RA_STR", "DEC_STR", "SRC_NAME", "TELESCOP", "BACKEND", "DATE-OBS"]
# keys asked for but usually absent, and near-miss decoys
QUERY_KEYS = INT_KEYS + DBL_KEYS + STR_KEYS + ["MISSING", "NPOLS", "OBSFRE", "TBI"]
DECOYS = ["NPOLX", "NPOLS", "OBSFREQQ", "XBLOCSIZ", "TBINS", "NBITSX"]


def _int_text(rng):
    v = rng.choice([rng.randint(0, 64), rng.randint(-40, 40), rng.randint(0, 1 << 28),
                    rng.randint(1 << 31, 1 << 33), rng.randint(0, 1 << 40)])
    return rng.choice([
        str(v), f"+{abs(v)}", f"{-abs(v)}", hex(abs(v)), f"0{rng.randint(1, 77)}",
        f"{v}.0", f"{v + rng.random():.3f}", f"{-v - rng.random():.3f}", f"{v}e0",
        f"{rng.uniform(1, 9):.2f}E{rng.randint(0, 3)}", f"{rng.uniform(1, 9):.2f}D{rng.randint(0, 3)}",
        f"{rng.uniform(1, 9):.1f}d{rng.randint(0, 2)}", f"{rng.randint(1, 9)}D{rng.randint(0, 3)}",
        f"#{abs(v)}", f"'{v}'", f"{v}abc", "-0", "0", "-1", " 17 ",
    ])


def _dbl_text(rng):
    v = rng.choice([rng.uniform(-2000, 2000), rng.uniform(0, 1e-5), rng.uniform(1e3, 1e4)])
    return rng.choice([
        repr(v), f"{v:.6f}", f"{v:.6e}", f"{v:.6E}", f"{v:.6e}".replace("e", "D"),
        f"{v:.4e}".replace("e", "d"), f"'{v:.5f}'", f"{v:.3f}xyz", str(int(v)), "-0",
    ])


def _sexa(rng, hours):
    a = rng.uniform(0, 24) if hours else rng.uniform(-90, 90)
    sign = "-" if a < 0 else rng.choice(["", "+"])
    a = abs(a)
    d, m = int(a), int(a * 60) % 60
    s = (a * 3600) % 60
    return rng.choice([f"{sign}{d:02d}:{m:02d}:{s:06.3f}", f"{sign}{d}:{m}", f"{sign}{a:.5f}",
                       f"{sign}{d}::{m}", f"{sign}{d}:{m}:{s:.1f}:9"])


def _str_text(rng, key):
    if key == "RA_STR":
        v = _sexa(rng, True)
    elif key == "DEC_STR":
        v = _sexa(rng, False)
    elif key == "DATE-OBS":
        v = "2021-05-%02dT12:34:56" % rng.randint(1, 28)
    else:
        n = rng.choice([3, 8, 12, 30, 60])
        v = "".join(rng.choice(string.ascii_letters + string.digits + "_+-. /") for _ in range(n))
    form = rng.random()
    if form < 0.6:
        return "'" + v.ljust(8) + "'"
    if form < 0.7:
        return '"' + v + '"'
    if form < 0.8:
        return "'" + v                       # unterminated quote
    return v.replace("/", "_")              # unquoted


def _card(rng, key, value):
    comment = rng.choice(["", "", " / a comment", " / it's here", " /"])
    k = key.lower() if rng.random() < 0.05 else key
    if rng.random() < 0.05 and len(k) <= 6:
        k = " " * rng.randint(1, 2) + k
    if rng.random() < 0.7 and not value.startswith(("'", '"')):
        body = f"{k:<8}= {value:>20}{comment}"
    else:
        body = f"{k:<8}= {value}{comment}"
    if rng.random() < 0.03:
        body = f"{k}={value}"               # no blanks at all
    return body[:80].ljust(80)


def raw_header(rng, directio=None) -> bytes:
    """One header: random subset of keywords in random order, some decoys and
    duplicates, END card; NUL-terminated."""
    keys = INT_KEYS + DBL_KEYS + STR_KEYS
    chosen = [k for k in keys if rng.random() < 0.8]
    cards = []
    for k in chosen:
        if rng.random() < 0.1:
            cards.append(_card(rng, rng.choice(DECOYS), _int_text(rng)))
        if k in INT_KEYS:
            v = _int_text(rng)
        elif k in DBL_KEYS:
            v = _dbl_text(rng)
        else:
            v = _str_text(rng, k)
        if k == "DIRECTIO" and directio is not None:
            v = str(directio)
        cards.append(_card(rng, k, v))
        if rng.random() < 0.05:              # a duplicate: the first one counts
            cards.append(_card(rng, k, _int_text(rng)))
    rng.shuffle(cards)
    cards.append("END".ljust(80))
    return "".join(cards).encode("latin-1")


def raw_headers(n, seed=1):
    rng = random.Random(seed)
    return [raw_header(rng) for _ in range(n)]


def ref_libs():
    """(rawutils, fbutils) from oracle/_ref, or None when not built."""
    import rawspec_amd as R
    pr, pf = os.path.join(REF_DIR, "libref_rawutils.so"), os.path.join(REF_DIR, "libref_fbutils.so")
    if not (os.path.exists(pr) and os.path.exists(pf)):
        return None
    return (R.host_format_prototypes(C.CDLL(pr), "rawspec_raw_"),
            R.host_format_prototypes(C.CDLL(pf), "fb_"))


def raw_answers(L, hdr: bytes) -> dict:
    """Everything a library answers for one NUL-terminated header."""
    import rawspec_amd as R
    buf = C.create_string_buffer(hdr + b"\0" * 16)
    out = {}
    for k in QUERY_KEYS:
        kb = k.encode()
        out[f"s32.{k}"] = L.rawspec_raw_get_s32(buf, kb, -7)
        out[f"u32.{k}"] = L.rawspec_raw_get_u32(buf, kb, 7)
        out[f"s64.{k}"] = L.rawspec_raw_get_s64(buf, kb, -9)
        out[f"u64.{k}"] = L.rawspec_raw_get_u64(buf, kb, 9)
        out[f"dbl.{k}"] = repr(L.rawspec_raw_get_dbl(buf, kb, 0.5))
        s = C.create_string_buffer(b"\x01" * 81)
        L.rawspec_raw_get_str(buf, kb, b"dflt", s, 80)
        out[f"str.{k}"] = s.raw[:80].split(b"\x01")[0].decode("latin-1")
    h = R.RawHdr()
    L.rawspec_raw_parse_header(buf, C.byref(h))
    for name, _t in R.RawHdr._fields_:
        if name in ("hdr_pos", "hdr_size"):
            continue
        v = getattr(h, name)
        out[f"hdr.{name}"] = v.decode("latin-1") if isinstance(v, bytes) else (
            repr(v) if isinstance(v, float) else v)
    raw = C.create_string_buffer(hdr, len(hdr))
    out["size0"] = L.rawspec_raw_header_size(raw, len(hdr), 0)
    out["size1"] = L.rawspec_raw_header_size(raw, len(hdr), 1)
    return out


def fb_random_header(rng):
    import rawspec_amd as R
    h = R.FbHdr()
    h.machine_id, h.telescope_id, h.data_type = rng.randint(0, 30), rng.randint(-1, 64), rng.randint(1, 2)
    h.barycentric, h.pulsarcentric = rng.choice([0, 0, 1]), rng.choice([0, 0, 1])
    h.src_raj, h.src_dej = rng.uniform(0, 24), rng.uniform(-90, 90)
    h.az_start, h.za_start = rng.uniform(0, 360), rng.uniform(0, 90)
    h.fch1, h.foff = rng.uniform(100, 20000), rng.choice([-1, 1]) * rng.uniform(1e-7, 4.0)
    h.nchans, h.nbeams, h.ibeam = rng.randint(1, 1 << 26), rng.randint(1, 64), rng.randint(0, 64)
    h.nbits, h.nifs = rng.choice([8, 16, 32]), rng.choice([1, 4])
    h.tstart, h.tsamp = rng.uniform(50000, 60000), rng.uniform(1e-7, 20)
    pick = string.ascii_letters + string.digits + "_-.+ "
    h.source_name = "".join(rng.choice(pick) for _ in range(rng.choice([0, 1, 8, 40, 79, 80]))).encode()
    h.rawdatafile = "".join(rng.choice(pick) for _ in range(rng.choice([1, 20, 60, 79, 80]))).encode()
    return h


def fb_headers(n, seed=2):
    rng = random.Random(seed)
    return [(fb_random_header(rng), rng.choice([0, 0, rng.randint(0, 300), rng.randint(300, 2500)]))
            for _ in range(n)]


def fb_fields(h) -> dict:
    out = {}
    for name, _t in type(h)._fields_:
        v = getattr(h, name)
        out[name] = v.decode("latin-1") if isinstance(v, bytes) else (
            v.hex() if isinstance(v, float) else v)
    return out


def fb_from_fields(d):
    import rawspec_amd as R
    h = R.FbHdr()
    for name, _t in R.FbHdr._fields_:
        v = d[name]
        if isinstance(getattr(h, name), bytes):
            setattr(h, name, v.encode("latin-1"))
        elif isinstance(getattr(h, name), float):
            setattr(h, name, float.fromhex(v))
        else:
            setattr(h, name, v)
    return h


def fb_bytes(L, h, minlen) -> bytes:
    """The header a library writes into a buffer for h padded to minlen."""
    buf = C.create_string_buffer(16384)
    end = L.fb_buf_write_padded_header(buf, C.byref(h), minlen)
    return buf.raw[: end - C.addressof(buf)]
