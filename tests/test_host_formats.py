"""GUPPI RAW header (rawspec_raw_*) and SIGPROC filterbank header (fb_*)
utilities exported by librawspec, checked against

  * the reference's own C (rawspec_rawutils.c, hget.h, rawspec_fbutils.c)
    compiled from its sources into oracle/_ref (skipped where absent),
  * fixtures that code answered (tests/golden/host_formats.json, written by
    tests/golden/make_golden.py --host), so the checks run without it too,
  * closed forms: sexagesimal conversions, card arithmetic of header sizes,
    the byte layout of filterbank values.

All of it is host code; no GPU is used."""
from __future__ import annotations

import ctypes as C
import json
import os
import random
import struct
import sys
import tempfile

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import rawspec_amd as R  # noqa: E402
from tests import host_cases as H  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def L():
    return R.lib()


@pytest.fixture(scope="module")
def ref():
    libs = H.ref_libs()
    if libs is None:
        pytest.skip("oracle/_ref not built (needs the reference sources: make -C oracle ref)")
    return libs


def _diff(a, b):
    return {k: (a[k], b[k]) for k in a if a[k] != b[k]}


# ---- struct layouts ---------------------------------------------------------

def test_struct_layouts_match_reference():
    """rawspec_raw_hdr_t and fb_hdr_t have the reference's sizes and offsets."""
    from tests.golden import make_golden as G
    want = json.load(open(os.path.join(GOLD, "host_formats_layout.json")))
    got = G.host_layouts(os.path.join(H.ROOT, "include"))
    assert got == want
    assert C.sizeof(R.RawHdr) == want["rawspec_raw_hdr_t"]["sizeof"]
    assert C.sizeof(R.FbHdr) == want["fb_hdr_t"]["sizeof"]
    for cls, t in ((R.RawHdr, "rawspec_raw_hdr_t"), (R.FbHdr, "fb_hdr_t")):
        for name, _ty in cls._fields_:
            assert getattr(cls, name).offset == want[t][name], (t, name)


# ---- GUPPI RAW header cards -------------------------------------------------

def test_raw_answers_match_reference_fuzz(L, ref):
    """Every getter, for every keyword, on 300 random headers (ints in decimal,
    hex, octal, with fractions, E/D exponents, quoted, junk-suffixed; strings
    quoted, unquoted, unterminated; comments; lowercase and indented keywords;
    decoy and duplicate cards), plus parse_header and header_size."""
    ref_raw, _ = ref
    for i, h in enumerate(H.raw_headers(300, seed=7)):
        d = _diff(H.raw_answers(L, h), H.raw_answers(ref_raw, h))
        assert not d, (i, d)


def test_raw_answers_match_golden(L):
    gold = json.load(open(os.path.join(GOLD, "host_formats.json")))["raw"]
    assert len(gold) >= 16
    for case in gold:
        got = H.raw_answers(L, case["header"].encode("latin-1"))
        assert got == case["expect"], _diff(got, case["expect"])


def _hdr(cards):
    return ("".join(c[:80].ljust(80) for c in cards) + "END".ljust(80)).encode() + b"\0"


def test_raw_getters_closed_forms(L):
    h = C.create_string_buffer(_hdr([
        "NPOL    =                    4", "OBSNCHAN=                   64",
        "BLOCSIZE=            134217728", "PKTIDX  =          12345678901",
        "OBSFREQ =             1500.125", "OBSBW   =                -187.5",
        "TBIN    =             3.413D-7", "NBITS   = 0x8",
        "NANTS   =                 010", "SRC_NAME= 'VOYAGER1'",
        "TELESCOP= 'GBT     ' / telescope", "RA_STR  = '17:10:03.984'",
        "DEC_STR = '-12:24:42.00'", "STT_IMJD=                59000",
        "STT_SMJD=                43200", "DIRECTIO=                    1",
        "ROUNDUP =                 2.5", "ROUNDDN =                -2.5",
        "NEARUP  =               2.9995", "NEARDN  =              -2.9995",
    ]))
    assert L.rawspec_raw_get_s32(h, b"NPOL", 0) == 4
    assert L.rawspec_raw_get_s32(h, b"npol", 0) == 4          # keywords are case-insensitive
    assert L.rawspec_raw_get_s32(h, b"NPO", 0) == 0           # a prefix is not the keyword
    assert L.rawspec_raw_get_s32(h, b"ABSENT", -3) == -3
    assert L.rawspec_raw_get_u64(h, b"PKTIDX", 0) == 12345678901
    assert L.rawspec_raw_get_s64(h, b"PKTIDX", 0) == 12345678901
    assert L.rawspec_raw_get_u32(h, b"NBITS", 0) == 8          # base-0 integer parsing
    assert L.rawspec_raw_get_u32(h, b"NANTS", 0) == 8          # leading 0 is octal
    assert L.rawspec_raw_get_dbl(h, b"TBIN", 0) == 3.413e-7    # Fortran D exponent
    assert L.rawspec_raw_get_dbl(h, b"OBSBW", 0) == -187.5
    # s32 truncates toward zero after a 0.001 nudge away from zero
    assert L.rawspec_raw_get_s32(h, b"ROUNDUP", 0) == 2
    assert L.rawspec_raw_get_s32(h, b"ROUNDDN", 0) == -2
    assert L.rawspec_raw_get_s32(h, b"NEARUP", 0) == 3
    assert L.rawspec_raw_get_s32(h, b"NEARDN", 0) == -3
    p = R.RawHdr()
    L.rawspec_raw_parse_header(h, C.byref(p))
    assert (p.npol, p.obsnchan, p.blocsize, p.pktidx, p.nbits, p.nants) == (4, 64, 134217728,
                                                                          12345678901, 8, 8)
    assert p.src_name == b"VOYAGER1" and p.telescop == b"GBT"
    assert p.ra == 17 + 10 / 60 + 3.984 / 3600
    assert p.dec == -(12 + 24 / 60 + 42.0 / 3600)
    assert p.mjd == 59000.5
    assert (p.beam_id, p.nbeam, p.refbeam, p.directio) == (-1, -1, -1, 1)


DMS_CASES = ["12:30:00", "-00:30:00", "+05:06:07.5", "17", "-17.25", "1:2", "1::2", ":30",
             "", "-", "+", "12:30:15:99", "abc", "0x10:1", " 3: 4", "-0:0:36", "1e1:1"]


def test_dmsstr_closed_forms(L):
    def f(s):
        return L.rawspec_raw_dmsstr_to_d(C.create_string_buffer(s.encode()))
    assert f("12:30:00") == 12.5
    assert f("-00:30:00") == -0.5          # the sign applies to the whole angle
    assert f("+05:06:07.5") == 5 + 6 / 60 + 7.5 / 3600
    assert f("-17.25") == -17.25
    assert f("1:2") == 1 + 2 / 60
    assert f("12:30:15:99") == 12 + 30 / 60 + 15 / 3600   # a fourth field is ignored


def test_dmsstr_matches_reference(L, ref):
    ref_raw, _ = ref
    rng = random.Random(3)
    cases = DMS_CASES + [H._sexa(rng, rng.random() < 0.5) for _ in range(300)]
    for s in cases:
        a = L.rawspec_raw_dmsstr_to_d(C.create_string_buffer(s.encode()))
        b = ref_raw.rawspec_raw_dmsstr_to_d(C.create_string_buffer(s.encode()))
        assert a == b or (a != a and b != b), (s, a, b)


@pytest.mark.parametrize("ncards", [0, 1, 5, 6, 7, 31, 32, 319])
def test_header_size_closed_form(L, ncards):
    h = ("".join(f"K{i:06d} = {i}".ljust(80) for i in range(ncards)) + "END".ljust(80)).encode()
    data = h + b"\x07" * 1024
    buf = C.create_string_buffer(data, len(data))
    n = 80 * (ncards + 1)
    assert L.rawspec_raw_header_size(buf, len(data), 0) == n
    assert L.rawspec_raw_header_size(buf, len(data), 1) == (n + 511) // 512 * 512
    # no END within len bytes
    assert L.rawspec_raw_header_size(buf, n - 80, 0) == 0


def test_header_size_matches_reference(L, ref):
    ref_raw, _ = ref
    rng = random.Random(5)
    for _ in range(200):
        n = rng.randint(0, 100)
        body = b"".join(("X%06d = 1" % i).ljust(80).encode() for i in range(n))
        tail = rng.choice([b"END".ljust(80), b"END", b"ENDX".ljust(80), b" END".ljust(80), b""])
        data = body + tail + bytes(rng.randrange(256) for _ in range(rng.randint(0, 700)))
        ln = rng.randint(max(0, len(data) - 200), len(data))
        for dio in (0, 1, 7):
            a = L.rawspec_raw_header_size(C.create_string_buffer(data, len(data)), ln, dio)
            b = ref_raw.rawspec_raw_header_size(C.create_string_buffer(data, len(data)), ln, dio)
            assert a == b, (n, tail, ln, dio, a, b)


def _valid_header(rng, blocsize, directio):
    cards = [
        f"BLOCSIZE= {blocsize:20d}", f"NPOL    = {rng.choice([1, 2, 4]):20d}",
        f"OBSNCHAN= {rng.choice([1, 8, 64]):20d}", f"NBITS   = {rng.choice([4, 8, 16]):20d}",
        f"OBSFREQ = {rng.uniform(1000, 9000):20.6f}", f"OBSBW   = {rng.choice([-1, 1]) * 187.5:20.3f}",
        f"TBIN    = {rng.uniform(1e-7, 1e-5):20.6E}", f"PKTIDX  = {rng.randint(0, 1 << 40):20d}",
        f"DIRECTIO= {directio:20d}", f"STT_IMJD= {rng.randint(50000, 60000):20d}",
        f"STT_SMJD= {rng.randint(0, 86399):20d}", f"RA_STR  = '{H._sexa(rng, True)}'",
        f"DEC_STR = '{H._sexa(rng, False)}'", f"SRC_NAME= 'SRC{rng.randint(0, 999)}'",
        f"TELESCOP= '{rng.choice(['GBT', 'Parkes', 'ATA', 'MeerKAT'])}'",
    ]
    for _ in range(rng.randint(0, 40)):
        cards.append(f"PAD{rng.randint(0, 9999):04d} = {rng.randint(0, 99)}")
    if rng.random() < 0.3:
        cards.append(f"NANTS   = {rng.randint(1, 64):20d}")
    if rng.random() < 0.3:
        cards.append(f"BEAM_ID = {rng.randint(0, 64):20d}")
    rng.shuffle(cards)
    return ("".join(c.ljust(80) for c in cards) + "END".ljust(80)).encode()


def _raw_file(path, rng, nblocks, drop=None):
    with open(path, "wb") as f:
        for b in range(nblocks):
            dio = rng.choice([0, 1])
            bs = rng.choice([512, 4096, 33280, 131072])
            h = _valid_header(rng, bs, dio)
            if drop is not None and b == nblocks - 1:
                h = b"".join(h[i:i + 80] for i in range(0, len(h), 80)
                             if not h[i:i + 80].startswith(drop.encode()))
            pad = (-len(h)) % 512 if dio else 0
            data = bytes([0]) + bytes(rng.randrange(256) for _ in range(bs - 1))
            f.write(h + b" " * pad + data)


def _read_all(lib, path):
    fd = os.open(path, os.O_RDONLY)
    out = []
    try:
        while True:
            h = R.RawHdr()
            pos = lib.rawspec_raw_read_header(fd, C.byref(h))
            fields = {n: getattr(h, n) for n, _t in R.RawHdr._fields_}
            out.append((pos, fields))
            if pos <= 0:
                break
            os.lseek(fd, h.blocsize, os.SEEK_CUR)
    finally:
        os.close(fd)
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_read_header_walks_file_like_reference(L, ref, seed):
    """Block by block through a RAW file (mixed DIRECTIO padding, block sizes,
    cards): positions, parsed fields, header sizes, and the end-of-file 0."""
    ref_raw, _ = ref
    rng = random.Random(seed)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.raw")
        _raw_file(p, rng, 6)
        ours, theirs = _read_all(L, p), _read_all(ref_raw, p)
    assert ours == theirs
    assert ours[-1][0] == 0 and len(ours) == 7


@pytest.mark.parametrize("drop", ["BLOCSIZE", "NPOL", "OBSNCHAN", "OBSFREQ", "OBSBW", "TBIN",
                                  "PKTIDX"])
def test_read_header_missing_required_keyword(L, ref, drop, capfd):
    ref_raw, _ = ref
    rng = random.Random(11)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.raw")
        _raw_file(p, rng, 2, drop=drop)
        ours = _read_all(L, p)
        out_ours = capfd.readouterr().err
        theirs = _read_all(ref_raw, p)
        out_theirs = capfd.readouterr().err
    assert ours[-1][0] == -1 and [o[0] for o in ours] == [t[0] for t in theirs]
    assert out_ours == out_theirs and drop in out_ours


def test_read_header_short_file(L):
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "short.raw")
        open(p, "wb").write(b"END".ljust(79))
        fd = os.open(p, os.O_RDONLY)
        try:
            assert L.rawspec_raw_read_header(fd, C.byref(R.RawHdr())) == 0
        finally:
            os.close(fd)


def test_cards_after_end_are_not_searched(L):
    """Only the cards up to END belong to the header: a keyword-like card in
    the data that follows is not found.  (The reference's search runs on to
    the first NUL byte, reading whatever follows END.)"""
    h = _hdr(["NPOL    =                    2"])[:-1] + "NBITS   =                   16".ljust(80).encode()
    assert L.rawspec_raw_get_s32(C.create_string_buffer(h), b"NBITS", 8) == 8
    assert L.rawspec_raw_get_s32(C.create_string_buffer(h), b"NPOL", 0) == 2


# ---- SIGPROC filterbank headers ---------------------------------------------

def test_fb_value_encoding(L):
    b = C.create_string_buffer(256)
    end = L.fb_buf_write_int(b, -2)
    assert end - C.addressof(b) == 4 and b.raw[:4] == struct.pack("<i", -2)
    end = L.fb_buf_write_double(b, 1420.40575)
    assert end - C.addressof(b) == 8 and b.raw[:8] == struct.pack("<d", 1420.40575)
    end = L.fb_buf_write_string(b, b"nchans")
    assert b.raw[: end - C.addressof(b)] == struct.pack("<i", 6) + b"nchans"
    end = L.fb_buf_write_string(b, b"x" * 100)          # strings are limited to 80
    assert b.raw[:4] == struct.pack("<i", 80) and end - C.addressof(b) == 84
    end = L.fb_buf_write_angle(b, -12.5)                # degrees -> ddmmss.s
    assert b.raw[:8] == struct.pack("<d", -123000.0)


def test_fb_angles_closed_forms(L):
    assert L.fb_ddd_to_dms(12.5) == 123000.0
    assert L.fb_ddd_to_dms(-0.5) == -3000.0
    assert L.fb_ddd_to_dms(1 + 2 / 60 + 3 / 3600) == pytest.approx(10203.0, abs=1e-9)
    assert L.fb_dms_to_ddd(123000.0) == 12.5
    assert L.fb_dms_to_ddd(-3000.0) == -0.5
    rng = random.Random(4)
    for _ in range(200):
        x = rng.uniform(-90, 90)
        assert L.fb_dms_to_ddd(L.fb_ddd_to_dms(x)) == pytest.approx(x, abs=1e-9)


def test_fb_angles_match_reference(L, ref):
    _, ref_fb = ref
    rng = random.Random(8)
    for x in [0.0, -0.0, 1e-12, -1e-12, 59.99999999, 89.9999999999] + [
            rng.uniform(-400, 400) for _ in range(500)]:
        assert L.fb_ddd_to_dms(x) == ref_fb.fb_ddd_to_dms(x), x
        y = rng.uniform(-900000, 900000)
        assert L.fb_dms_to_ddd(y) == ref_fb.fb_dms_to_ddd(y), y


def test_fb_header_bytes_match_reference(L, ref):
    """Headers (with padding to random minimum lengths) are byte-identical."""
    _, ref_fb = ref
    for i, (h, m) in enumerate(H.fb_headers(300, seed=12)):
        assert H.fb_bytes(L, h, m) == H.fb_bytes(ref_fb, h, m), (i, m)


def test_fb_header_bytes_match_golden(L):
    gold = json.load(open(os.path.join(GOLD, "host_formats.json")))["fb"]
    assert len(gold) >= 10
    for case in gold:
        h = H.fb_from_fields(case["hdr"])
        assert H.fb_bytes(L, h, case["minlen"]).hex() == case["bytes"]


def test_fb_header_structure(L):
    """Keyword order, value types, conditional keywords and the padding rule,
    read straight off the bytes."""
    h = H.fb_random_header(random.Random(1))
    h.barycentric, h.pulsarcentric = 1, 0
    raw = H.fb_bytes(L, h, 0)
    items, p = [], 0
    types = {"machine_id": "i", "telescope_id": "i", "data_type": "i", "barycentric": "i",
             "pulsarcentric": "i", "src_raj": "d", "src_dej": "d", "az_start": "d",
             "za_start": "d", "fch1": "d", "foff": "d", "nchans": "i", "nbeams": "i",
             "ibeam": "i", "nbits": "i", "tstart": "d", "tsamp": "d", "nifs": "i",
             "source_name": "s", "rawdatafile": "s", "HEADER_START": "", "HEADER_END": ""}
    while p < len(raw):
        n, = struct.unpack_from("<i", raw, p)
        k = raw[p + 4: p + 4 + n].decode()
        p += 4 + n
        t = types[k]
        if t == "i":
            v, = struct.unpack_from("<i", raw, p); p += 4
        elif t == "d":
            v, = struct.unpack_from("<d", raw, p); p += 8
        elif t == "s":
            m, = struct.unpack_from("<i", raw, p)
            v = raw[p + 4: p + 4 + m].decode(); p += 4 + m
        else:
            v = None
        items.append((k, v))
    keys = [k for k, _ in items]
    assert keys[0] == "HEADER_START" and keys[-1] == "HEADER_END"
    assert "barycentric" in keys and "pulsarcentric" not in keys
    d = dict(items)
    assert d["nchans"] == h.nchans and d["fch1"] == h.fch1 and d["tsamp"] == h.tsamp
    assert d["src_raj"] == L.fb_ddd_to_dms(h.src_raj)
    assert d["rawdatafile"] == h.rawdatafile.decode() and d["source_name"] == h.source_name.decode()
    # padding: minlen far above the plain size adds "rawdatafile" entries first
    padded = H.fb_bytes(L, h, len(raw) + 1000)
    assert len(padded) >= len(raw) + 1000 - 100
    assert padded.count(b"rawdatafile") > raw.count(b"rawdatafile")


def test_fb_fd_writer_equals_buffer_writer(L):
    for h, m in H.fb_headers(40, seed=21):
        want = H.fb_bytes(L, h, m)
        with tempfile.TemporaryFile() as f:
            n = L.fb_fd_write_padded_header(f.fileno(), C.byref(h), m)
            f.seek(0)
            got = f.read()
        assert n == len(want) and got == want


def _read_back(lib, raw):
    o = R.FbHdr()
    ln = C.c_size_t(0)
    buf = C.create_string_buffer(raw + b"\xff" * 32, len(raw) + 32)
    end = lib.fb_buf_read_header(buf, C.byref(o), C.byref(ln))
    return H.fb_fields(o), ln.value, (end or 0) - C.addressof(buf)


def test_fb_reader_round_trip_and_reference(L, ref):
    _, ref_fb = ref
    for i, (h, m) in enumerate(H.fb_headers(200, seed=31)):
        raw = H.fb_bytes(L, h, m)
        ours, theirs = _read_back(L, raw), _read_back(ref_fb, raw)
        assert ours == theirs, i
        f, ln, end = ours
        assert ln == len(raw) and end == len(raw)
        src = H.fb_fields(h)
        for k in ("nchans", "fch1", "foff", "tstart", "tsamp", "nbits", "nifs", "source_name",
                  "rawdatafile", "telescope_id", "machine_id", "ibeam", "nbeams"):
            assert f[k] == src[k], (i, k)
        assert float.fromhex(f["src_dej"]) == pytest.approx(h.src_dej, abs=1e-9)


def test_fb_reader_rejects_non_header(L):
    o = R.FbHdr()
    ln = C.c_size_t(77)
    buf = C.create_string_buffer(struct.pack("<i", 6) + b"NOTHDR" + b"\0" * 64)
    end = L.fb_buf_read_header(buf, C.byref(o), C.byref(ln))
    assert end == C.addressof(buf) and ln.value == 77


def test_fb_fd_reader_matches_reference(L, ref):
    """Reading from a file stores the header length and leaves the offset at
    the first data byte."""
    _, ref_fb = ref
    for h, m in H.fb_headers(30, seed=41):
        raw = H.fb_bytes(L, h, m)
        with tempfile.TemporaryFile() as f:
            f.write(raw + bytes(range(256)) * 40)
            res = []
            for lib in (L, ref_fb):
                f.seek(0)
                o = R.FbHdr()
                ln = C.c_size_t(0)
                r = lib.fb_fd_read_header(f.fileno(), C.byref(o), C.byref(ln))
                res.append((r, ln.value, os.lseek(f.fileno(), 0, os.SEEK_CUR), H.fb_fields(o)))
        assert res[0] == res[1]
        assert res[0][1] == len(raw) and res[0][2] == len(raw)


def test_fb_fd_reader_short_file(L):
    """A file holding only a header (shorter than the 4096-byte read) leaves
    the offset just past it.  (The reference seeks by the header length minus
    4096 from the end of the read, which lands before it.)"""
    h, _ = H.fb_headers(1, seed=5)[0]
    raw = H.fb_bytes(L, h, 0)
    with tempfile.TemporaryFile() as f:
        f.write(raw + b"\x01" * 100)
        f.seek(0)
        ln = C.c_size_t(0)
        r = L.fb_fd_read_header(f.fileno(), C.byref(R.FbHdr()), C.byref(ln))
        assert r == len(raw) == ln.value and os.lseek(f.fileno(), 0, os.SEEK_CUR) == len(raw)


def test_fb_reader_skips_other_sigproc_keywords(L):
    """Values of SIGPROC keywords outside fb_hdr_t (nsamples, refdm, ...) are
    skipped, so the fields after them still parse."""
    def s(x):
        return struct.pack("<i", len(x)) + x
    raw = (s(b"HEADER_START") + s(b"nsamples") + struct.pack("<i", 99) + s(b"refdm")
           + struct.pack("<d", 56.7) + s(b"nchans") + struct.pack("<i", 4096) + s(b"fch1")
           + struct.pack("<d", 1500.0) + s(b"HEADER_END"))
    f, ln, _ = _read_back(L, raw)
    assert f["nchans"] == 4096 and float.fromhex(f["fch1"]) == 1500.0 and ln == len(raw)


TELESCOPES = ["FAKE", "ARECIBO", "OOTY", "NANCAY", "PARKES", "JODRELL", "GBT", "GMRT",
              "EFFELSBERG", "140FT", "ATA", "SRT", "LEUSCHNER", "MEERKAT", "KAT-7", "VLA",
              "gbt", "Parkes", "MeerKAT", "GBTX", "", "GB", "  GBT", "Effelsberg_LOFAR",
              "LOFAR", "GMRT ", "unknown"]


def test_telescope_ids(L):
    assert L.fb_telescope_id(b"GBT") == 6 and L.fb_telescope_id(b"gbt") == 6
    assert L.fb_telescope_id(b"PARKES") == 4 and L.fb_telescope_id(b"ATA") == 9
    assert L.fb_telescope_id(b"140FT") == 9 and L.fb_telescope_id(b"SRT") == 10
    assert L.fb_telescope_id(b"nowhere") == -1


def test_telescope_ids_match_reference(L, ref):
    _, ref_fb = ref
    for t in TELESCOPES:
        assert L.fb_telescope_id(t.encode()) == ref_fb.fb_telescope_id(t.encode()), t


@pytest.mark.parametrize("nbits,directio", [(8, 1), (16, 0)])
def test_noise_file_reads_back(L, nbits, directio):
    """rawfile.write_noise: every block's header parses with the block
    geometry asked for, blocks follow each other at header + BLOCSIZE, the
    samples are the clamped rounded noise (mean near 0, sigma near the one
    asked for, integral, in range), and the file ends after the last block."""
    import numpy as np
    from rawspec_amd.rawfile import write_noise
    nchan, npol, ntpb, nblocks, sigma = 3, 2, 256, 3, 8.0
    with tempfile.TemporaryDirectory() as d:
        name = write_noise(os.path.join(d, "s"), nchan=nchan, npol=npol, nbits=nbits, ntpb=ntpb,
                           nblocks=nblocks, sigma=sigma, seed=5, directio=directio)
        blocs = nchan * ntpb * npol * 2 * nbits // 8
        walk = _read_all(L, name)
        assert [p > 0 for p, _f in walk] == [True] * nblocks + [False]
        ends = []
        for b, (pos, f) in enumerate(walk[:-1]):
            assert (f["obsnchan"], f["npol"], f["nbits"], f["blocsize"], f["pktidx"],
                    f["directio"]) == (nchan, npol, nbits, blocs, 32 * b, directio)
            if directio:
                assert pos % 512 == 0
            ends.append(pos + blocs)
        assert ends[-1] == os.path.getsize(name)
        raw = open(name, "rb").read()
        x = np.frombuffer(raw[walk[0][0]:walk[0][0] + blocs], dtype=np.int8 if nbits == 8 else np.int16)
        assert abs(x.mean()) < 0.5 and abs(x.std() - sigma) < 0.5
