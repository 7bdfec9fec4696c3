"""CPU checks for the rawspec program: the restatement of the reference
program's block sequencing, filterbank header derivation, output names and
per-antenna split (oracle/rawstem.py) against hand-worked cases and closed
forms; and the program's argument and input checks that end before any GPU
work."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle.rawstem as RS
from tests.cli_cases import CLI, run_cli
from tests.raw_stems import StemSpec, header, write_stem

Z = RS.ZERO


def _files(*files):
    return [[(1000 + 32 * s, True) for s in f] for f in files]


# ---- block sequencing (rawspec.c:905-1019) ---------------------------------

def test_sequence_plain_and_partial_buffer():
    bufs, why = RS.sequence(_files([0, 1, 2, 3, 4]), 2)
    assert why == "end"
    assert bufs == [[(0, 0), (0, 1)], [(0, 2), (0, 3)]]        # block 4 dropped


def test_sequence_gap_becomes_zero_blocks_across_files():
    bufs, _ = RS.sequence(_files([0, 1, 2, 3], [4, 5, 8, 9]), 1)
    flat = [b[0] for b in bufs]
    assert flat == [(0, 0), (0, 1), (0, 2), (0, 3), (1, 0), (1, 1), Z, Z, (1, 2), (1, 3)]


def test_sequence_step_learned_from_first_increase():
    # the step is 64, not 32: the first increase fixes it
    f = [[(1000, True), (1064, True), (1128, True), (1256, True)]]
    bufs, _ = RS.sequence(f, 1)
    assert [b[0] for b in bufs] == [(0, 0), (0, 1), (0, 2), Z, (0, 3)]


@pytest.mark.parametrize("steps,why,n", [
    ([0, 1, 2, 1, 4], "backwards", 3),
    ([0, 2, 3], "misaligned", 2),           # the first increase (two steps) is the step
    ([0, 1, 3, 4], "end", 5),               # 0 1 Z 3 4
])
def test_sequence_stops(steps, why, n):
    bufs, w = RS.sequence(_files(steps), 1)
    assert w == why and len(bufs) == n


def test_sequence_misaligned_and_null():
    f = [[(1000, True), (1032, True), (1080, True), (1096, True)]]
    bufs, why = RS.sequence(f, 1)
    assert why == "misaligned" and len(bufs) == 2
    # a repeated PKTIDX skips the rest of its file only
    f = [[(1000, True), (1032, True), (1032, True), (1064, True)], [(1096, True)]]
    bufs, why = RS.sequence(f, 1)
    assert [b[0] for b in bufs] == [(0, 0), (0, 1), Z, (1, 0)]


def test_sequence_incomplete_block_ends_stem():
    f = [[(1000, True), (1032, True), (1064, False), (1096, True)], [(1128, True)]]
    bufs, why = RS.sequence(f, 1)
    assert why == "incomplete" and len(bufs) == 2


def test_sequence_readme_counts():
    """1001 buffers of one block each: the number of complete buffers is what
    the dump counts of the reference's README are taken over."""
    bufs, _ = RS.sequence(_files(list(range(1001))), 1)
    assert len(bufs) == 1001
    assert (len(bufs) // 51, len(bufs) * 1024, len(bufs) // 3) == (19, 1025024, 333)


# ---- filterbank header derivation (rawspec.c:806-832) ----------------------

def _fields(**kw):
    base = dict(obsfreq=1420.0, obsbw=187.5, obsnchan=64, nants=1, tbin=3.2e-7, mjd=59000.5,
                ra=1.0, dec=2.0, telescope_id=6, src_name="X", rawdatafile="x.0000.raw",
                beam_id=-1, refbeam=-1, Nc=64, Nt=1024, Na=16, Npolout=1, schan=0)
    base.update(kw)
    return RS.fb_fields(**base), base


@pytest.mark.parametrize("obsbw", [187.5, -187.5, 3.0])
@pytest.mark.parametrize("Nt", [8, 1024, 7, 1 << 20])
@pytest.mark.parametrize("nants,obsnchan,schan,Nc", [(1, 64, 0, 64), (1, 64, 5, 20), (4, 64, 0, 64),
                                                     (4, 64, 19, 8)])
def test_fine_channel_frequencies_closed_form(obsbw, Nt, nants, obsnchan, schan, Nc):
    """The header's fch1/foff put fine channel j of coarse channel c (after
    the FFT shift) at that channel's sky frequency: the centre of coarse
    channel schan + c of one antenna's band, plus (j - floor(Nt/2)) fine
    channel widths."""
    f, b = _fields(obsbw=obsbw, Nt=Nt, nants=nants, obsnchan=obsnchan, schan=schan,
                   Nc=Nc if nants == 1 else Nc)
    ncpa = obsnchan // nants
    width = obsbw / ncpa
    assert f["foff"] == pytest.approx(width / Nt, rel=1e-15)
    nc_out = f["nchans"] // Nt
    for c in [0, nc_out - 1]:
        cc = (schan % ncpa) + c
        centre = b["obsfreq"] + (cc - (ncpa - 1) / 2) * width
        for j in [0, Nt // 2, Nt - 1]:
            got = f["fch1"] + (c * Nt + j) * f["foff"]
            assert got == pytest.approx(centre + (j - Nt // 2) * width / Nt, abs=1e-9 * abs(obsbw))
    assert f["tsamp"] == pytest.approx(b["tbin"] * Nt * b["Na"])
    assert f["nchans"] == Nc * Nt // nants


def test_fields_passthrough():
    f, _ = _fields(beam_id=3, refbeam=2, Npolout=-4, src_name="S" * 100)
    assert (f["ibeam"], f["refbeam"], f["nifs"], f["nbits"], f["machine_id"], f["data_type"]) == \
        (3, 2, 4, 32, 20, 1)
    assert f["source_name"] == "S" * 80 and f["nfpc"] == 1024


# ---- names and per-antenna split (rawspec_file.c) ---------------------------

def test_output_names():
    assert RS.output_name("/data/x/guppi_1", 0) == "/data/x/guppi_1.rawspec.0000.fil"
    assert RS.output_name("/data/x/guppi_1", 12, dest="/out") == "/out/guppi_1.rawspec.0012.fil"
    assert RS.output_name("g", 1, ant=3) == "g-ant003.rawspec.0001.fil"
    assert RS.output_name("/a/g", 2, dest="/o", ics=True) == "/o/g-ics.rawspec.0002.fil"


def test_split_payload_closed_form():
    """Value v = 1000*dump + 100*pol + chan*Nt + fine: antenna a's file holds,
    in order, for every dump and pol, channels a*Ncpa .. (a+1)*Ncpa-1."""
    nds, nifs, Nant, Ncpa, Nt = 2, 4, 3, 2, 5
    d = np.array([[[1000 * k + 100 * q + c * Nt + j for c in range(Nant * Ncpa) for j in range(Nt)]
                   for q in range(nifs)] for k in range(nds)], dtype=np.float32).reshape(-1)
    files = RS.file_payload([(d, nds)], nifs=nifs, Nant=Nant, per_ant=True)
    for a in range(Nant):
        got = np.frombuffer(files[a], dtype=np.float32).reshape(nds, nifs, Ncpa * Nt)
        for k in range(nds):
            for q in range(nifs):
                want = [1000 * k + 100 * q + c * Nt + j for c in range(a * Ncpa, (a + 1) * Ncpa)
                        for j in range(Nt)]
                assert list(got[k, q]) == want
    one = RS.file_payload([(d, nds), (d + 1, nds)], nifs=nifs, Nant=Nant, per_ant=False)
    assert one[0] == d.tobytes() + (d + 1).tobytes()


# ---- the program's own checks, before any GPU work --------------------------

needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="rawspec program not built")


@needs_cli
@pytest.mark.parametrize("args,msg", [
    ([], "a file stem must be specified"),
    (["-p", "3", "x"], "number of output pols must be 1 or +/- 4"),
    (["-f", "1024", "-t", "1,2", "x"], "must specify same number of FFT and integration lengths"),
    (["-s", "4", "x"], "nchan must be non-zero if schan is non-zero"),
    (["-f", "1,2,3,4,5", "x"], "up to 4 fine channel counts supported"),
    (["-d", "127.0.0.1:9", "-p", "4", "x"], "full-pol mode is not supported for network output"),
])
def test_argument_errors(args, msg):
    r = run_cli(args, check=False)
    assert r.returncode == 1 and msg in r.stderr


@needs_cli
def test_missing_stem_counts_nothing(tmp_path):
    r = run_cli([str(tmp_path / "nothing")], check=False)
    assert r.returncode == 0
    assert "using default FFT and integration lengths" in r.stdout
    for i in range(3):
        assert f"output product {i}: 0 spectra" in r.stdout


@needs_cli
def test_bad_block_geometry_is_reported(tmp_path):
    spec = StemSpec(obsnchan=8, ntpb=64)
    stem = str(tmp_path / "bad")
    write_stem(spec, stem)
    raw = open(stem + ".0000.raw", "rb").read()
    raw = raw.replace(b"BLOCSIZE=                 2048", b"BLOCSIZE=                 2050")
    open(stem + ".0000.raw", "wb").write(raw)
    r = run_cli(["-f", "64", "-t", "1", stem], check=False)
    assert r.returncode == 0 and "bad block geometry" in r.stdout


@needs_cli
def test_bad_obsnchan_nants_is_reported(tmp_path):
    spec = StemSpec(obsnchan=6, nants=4, ntpb=64)
    stem = str(tmp_path / "bad")
    write_stem(spec, stem)
    r = run_cli(["-f", "64", "-t", "1", stem], check=False)
    assert r.returncode == 0 and "bad obsnchan/nants: 6 % 4 != 0" in r.stderr


def test_header_cards_parse_back():
    """The stems' headers carry what the parser reads (GUPPI card layout)."""
    import ctypes as C

    import rawspec_amd as R
    spec = StemSpec(obsnchan=16, nants=2, nbits=16, directio=1, beam_id=3)
    h = header(spec, 4242)
    assert len(h) % 512 == 0
    L = R.lib()
    p = R.RawHdr()
    L.rawspec_raw_parse_header(C.create_string_buffer(h + b"\0"), C.byref(p))
    assert (p.obsnchan, p.nants, p.nbits, p.npol, p.pktidx, p.blocsize, p.directio, p.beam_id) == \
        (16, 2, 16, 2, 4242, spec.blocsize, 1, 3)
    assert p.telescop == b"GBT" and p.src_name == b"VOYAGER1"
