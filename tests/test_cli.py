"""The rawspec program end to end on the GPU: synthetic GUPPI RAW stems in,
filterbank files (or UDP packets) out, checked against the oracle.

What must hold, per case: the output file names; the filterbank header
bytes (fields derived as the reference program derives them, written by the
header writer that matches the reference's byte for byte); the number of
dumps (block sequencing with PKTIDX gaps, stops and dropped partial
buffers); and every spectrum within the plain path's 1e-5 error bound of the
FP64 oracle fed the same blocks."""
from __future__ import annotations

import os
import socket
import threading

import numpy as np
import pytest

import oracle.rawstem as RS
from tests import fbh5_read as F
from tests.cli_cases import (CLI, StemSpec, check_fil, check_split, expected, header_fields,
                             run_cli, write_stem)

pytestmark = pytest.mark.gpu


def _run_products(tmp_path, spec, *, Nts, Nas, Npolout, extra=(), schan=0, nchan=0,
                  name="stem", dest=True, sync=False):
    stem = str(tmp_path / name)
    paths = write_stem(spec, stem)
    out = tmp_path / "out"
    out.mkdir(exist_ok=True)
    args = ["-f", ",".join(map(str, Nts)), "-t", ",".join(map(str, Nas)),
            "-p", ",".join(map(str, Npolout))]
    if dest:
        args += ["-d", str(out)]
    if schan or nchan:
        args += ["-s", str(schan), "-n", str(nchan)]
    if sync:
        args += ["--sync-ingest"]
    r = run_cli([*args, *extra, stem])
    pw, ic, cfg, reason = expected(spec, Nts=Nts, Nas=Nas, Npolout=Npolout, schan=schan,
                                   nchan=nchan)
    return stem, str(out) if dest else None, os.path.basename(paths[0]), r, pw, cfg, reason


def _check_products(stem, dest, bfname, spec, pw, cfg, Npolout, schan=0):
    worst = 0.0
    for i in range(cfg.No):
        f = header_fields(spec, cfg, i, schan=schan, rawdatafile=bfname)
        f["nifs"] = abs(Npolout[i])
        path = RS.output_name(stem, i, dest=dest)
        worst = max(worst, check_fil(path, f, pw[i], nifs=abs(cfg.Npolout[i]), Nc=cfg.Nc,
                                     Nt=cfg.Nts[i], Npolout=cfg.Npolout[i]))
    return worst


def test_8ch_int8_stem_nfft1024(tmp_path):
    """8 coarse channels x 2 pols int8, 4 blocks of 131072 samples, tone at
    bin 300 of channel 3; Nfft 1024, Nint 16, Stokes I."""
    spec = StemSpec(obsnchan=8, ntpb=131072, files=[[0, 1, 2, 3]])
    stem, dest, bf, r, pw, cfg, reason = _run_products(tmp_path, spec, Nts=[1024], Nas=[16],
                                                       Npolout=[1])
    assert reason == "end" and len(pw[0]) == 4 and cfg.Nds[0] == 8
    worst = _check_products(stem, dest, bf, spec, pw, cfg, [1])
    assert "output product 0: 32 spectra" in r.stdout
    print(f"8-channel stem: worst rel err {worst:.2e}")


def test_gaps_files_and_products(tmp_path):
    """Two files, one PKTIDX gap (a zero block), three products with uneven
    Nfft / Nint and all three pol modes; sync and pipelined ingest agree."""
    spec = StemSpec(obsnchan=8, ntpb=4096, files=[[0, 1, 2, 3], [4, 5, 7, 8, 9]])
    kw = dict(Nts=[1024, 64, 4096], Nas=[2, 128, 1], Npolout=[1, 4, -4])
    stem, dest, bf, r, pw, cfg, reason = _run_products(tmp_path, spec, **kw)
    assert reason == "end"
    assert [len(pw[i]) for i in range(3)] == [10, 5, 10]
    _check_products(stem, dest, bf, spec, pw, cfg, kw["Npolout"])
    first = {i: open(RS.output_name(stem, i, dest=dest), "rb").read() for i in range(3)}
    _run_products(tmp_path, spec, **kw, sync=True)
    for i in range(3):
        assert open(RS.output_name(stem, i, dest=dest), "rb").read() == first[i]


@pytest.mark.parametrize("case,files,emitted", [
    ("backwards", [[0, 1, 2, 3, 1, 5, 6]], 4),
    ("misaligned", [[0, 1, 2, 3.5, 4]], 3),
    ("null", [[0, 1, 2, 2, 3, 4], [5, 6, 7]], 8),        # 0 1 2 Z Z 5 6 7
    ("incomplete", [[0, 1, 2, 3, 4]], 4),
])
def test_stem_stops(tmp_path, case, files, emitted):
    """A backward or misaligned PKTIDX step, or a block cut short by the end
    of the file, ends the stem; a repeated PKTIDX ends the file."""
    spec = StemSpec(obsnchan=4, ntpb=2048, files=files, directio=1,
                    truncate_last=1000 if case == "incomplete" else 0)
    kw = dict(Nts=[512], Nas=[4], Npolout=[1])
    stem, dest, bf, r, pw, cfg, reason = _run_products(tmp_path, spec, **kw)
    assert reason == case and cfg.Nb == 1 and len(pw[0]) == emitted
    assert {"backwards": "backwards jump", "misaligned": "misaligned jump",
            "null": "null jump", "incomplete": "incomplete block"}[case] in r.stdout + r.stderr
    _check_products(stem, dest, bf, spec, pw, cfg, [1])


def test_channel_subset_and_directio(tmp_path):
    spec = StemSpec(obsnchan=16, ntpb=2048, directio=1, obsbw=-187.5,
                    files=[[0, 1, 2, 3, 4, 5]], tone=(5, 20.0, 100.0, 512.0))
    kw = dict(Nts=[512, 8], Nas=[4, 64], Npolout=[4, -4])
    stem, dest, bf, r, pw, cfg, _ = _run_products(tmp_path, spec, **kw, schan=3, nchan=6)
    assert cfg.Nc == 6 and cfg.input_conjugated == 1
    _check_products(stem, dest, bf, spec, pw, cfg, kw["Npolout"], schan=3)


def test_int16_stem(tmp_path):
    spec = StemSpec(obsnchan=4, nbits=16, ntpb=8192, sigma=1000.0, files=[[0, 1, 2, 3]],
                    tone=(1, 3000.0, 77.0, 2048.0))
    kw = dict(Nts=[2048, 16], Nas=[2, 256], Npolout=[-4, 1])
    stem, dest, bf, r, pw, cfg, _ = _run_products(tmp_path, spec, **kw)
    _check_products(stem, dest, bf, spec, pw, cfg, kw["Npolout"])


def test_int4_stem(tmp_path):
    """4-bit samples go through the library's expanding copy."""
    spec = StemSpec(obsnchan=4, nbits=4, ntpb=4096, files=[[0, 1, 2, 3]])
    kw = dict(Nts=[256], Nas=[8], Npolout=[1])
    stem, dest, bf, r, pw, cfg, _ = _run_products(tmp_path, spec, **kw, extra=["--stats"])
    assert '"mode": "sync"' in r.stderr
    _check_products(stem, dest, bf, spec, pw, cfg, kw["Npolout"])


def test_split_antennas(tmp_path):
    """NANTS > 1 without -S or -i splits the output per antenna."""
    spec = StemSpec(obsnchan=12, nants=3, ntpb=2048, files=[[0, 1, 2, 3]])
    kw = dict(Nts=[256, 8], Nas=[8, 256], Npolout=[1, -4])
    stem = str(tmp_path / "multi")
    paths = write_stem(spec, stem)
    r = run_cli(["-f", "256,8", "-t", "8,256", "-p", "1,-4", "-d", str(tmp_path), stem])
    assert "Splitting output per 3 antennas" in r.stdout
    pw, _, cfg, _ = expected(spec, **kw)
    for i in range(2):
        f = header_fields(spec, cfg, i, rawdatafile=os.path.basename(paths[0]))
        files = [RS.output_name(stem, i, dest=str(tmp_path), ant=a) for a in range(3)]
        check_split(files, f, pw[i], nifs=abs(cfg.Npolout[i]), Nant=3, Ncpa=4, Nt=cfg.Nts[i],
                    Npolout=cfg.Npolout[i])


def test_incoherent_sum(tmp_path):
    """-i with per-antenna weights writes only the incoherent-sum file."""
    spec = StemSpec(obsnchan=8, nants=4, ntpb=2048, files=[[0, 1, 2, 3]])
    w = [1.0, 0.5, 1.5, 0.75]
    stem = str(tmp_path / "ics")
    paths = write_stem(spec, stem)
    run_cli(["-f", "512", "-t", "4", "-p", "4", "-i", ",".join(map(str, w)), "-d", str(tmp_path),
             stem])
    pw, ic, cfg, _ = expected(spec, Nts=[512], Nas=[4], Npolout=[4], ics=True, weights=w)
    f = header_fields(spec, cfg, 0, rawdatafile=os.path.basename(paths[0]))
    ics_path = RS.output_name(stem, 0, dest=str(tmp_path), ics=True)
    check_fil(ics_path, f, ic[0], nifs=4, Nc=2, Nt=512, Npolout=4)
    assert not os.path.exists(RS.output_name(stem, 0, dest=str(tmp_path)))


def test_headers_file_and_outidx(tmp_path):
    """-H saves each block's header (through its END card, without the
    DIRECTIO padding) to STEM.rawspec.headers; -o numbers the outputs."""
    from tests.raw_stems import header, pktidx_of
    spec = StemSpec(obsnchan=4, ntpb=1024, directio=1, files=[[0, 1], [2, 3]])
    stem = str(tmp_path / "h")
    write_stem(spec, stem)
    run_cli(["-f", "256", "-t", "4", "-H", "-o", "7", stem])
    one = header(spec, spec.pktidx0)
    n = 80 + next(k for k in range(0, len(one), 80) if one[k:k + 80] == b"END".ljust(80))
    assert len(one) % 512 == 0 and len(one) > n
    want = b"".join(header(spec, pktidx_of(spec, s))[:n] for s in [0, 1, 2, 3])
    assert open(stem + ".rawspec.headers", "rb").read() == want
    assert os.path.exists(stem + ".rawspec.0007.fil")


def test_two_stems_restart_integration(tmp_path):
    """A second stem with the same geometry restarts the integrations."""
    spec = StemSpec(obsnchan=4, ntpb=2048, files=[[0, 1, 2]], seed=3)
    spec2 = StemSpec(obsnchan=4, ntpb=2048, files=[[0, 1, 2, 3, 4]], seed=4)
    s1, s2 = str(tmp_path / "a"), str(tmp_path / "b")
    write_stem(spec, s1)
    write_stem(spec2, s2)
    run_cli(["-f", "1024", "-t", "4", "-d", str(tmp_path), s1, s2])  # Nis = 2
    for sp, s in ((spec, s1), (spec2, s2)):
        pw, _, cfg, _ = expected(sp, Nts=[1024], Nas=[4], Npolout=[1])
        assert len(pw[0]) == len(sp.files[0]) // 2
        _check_products(s, str(tmp_path), os.path.basename(s) + ".0000.raw", sp, pw, cfg, [1])


def test_network_output(tmp_path):
    """UDP packets carry the same spectra as the file output, with per-packet
    headers whose fch1 / nchans / tstart describe each slice."""
    import ctypes as C

    import rawspec_amd as R
    spec = StemSpec(obsnchan=8, ntpb=4096, files=[[0, 1, 2, 3]])
    stem = str(tmp_path / "n")
    paths = write_stem(spec, stem)
    rx = socket.socket(socket.AF_INET, socket.SOCK_DGRAM)
    rx.setsockopt(socket.SOL_SOCKET, socket.SO_RCVBUF, 8 << 20)
    rx.bind(("127.0.0.1", 0))
    rx.settimeout(0.5)
    port = rx.getsockname()[1]
    pkts = []
    done = threading.Event()

    def recv():
        while not done.is_set():
            try:
                pkts.append(rx.recv(65536))
            except socket.timeout:
                pass
    th = threading.Thread(target=recv)
    th.start()
    try:
        run_cli(["-f", "1024,8", "-t", "4,2048", "-r", "0.2", "-d", f"127.0.0.1:{port}", stem])
    finally:
        import time
        time.sleep(1.0)
        done.set()
        th.join()
        rx.close()
    pw, _, cfg, _ = expected(spec, Nts=[1024, 8], Nas=[4, 2048], Npolout=[1, 1])
    L = R.lib()
    by_nchan = {}
    for p in pkts:
        h = R.FbHdr()
        n = C.c_size_t(0)
        L.fb_buf_read_header(C.create_string_buffer(p, len(p)), C.byref(h), C.byref(n))
        assert (0x6c + n.value) % 512 == 0
        data = np.frombuffer(p[n.value:], dtype=np.float32)
        by_nchan.setdefault(h.foff, []).append((h.tstart, h.fch1, h.nchans, data))
    for i in range(2):
        f = header_fields(spec, cfg, i, rawdatafile=os.path.basename(paths[0]))
        got = sorted(by_nchan[f["foff"]], key=lambda t: (t[0], t[1]))
        nchans = f["nchans"]
        want = np.concatenate([d for d, _ in pw[i]]).reshape(-1, nchans)
        # reassemble spectra in time order, channel slices in frequency order
        spectra = []
        per_pkt = max(1, 2048 // nchans) if nchans < 2048 else 1
        k = 0
        while k < len(got):
            if nchans >= 2048:
                row = np.concatenate([g[3] for g in got[k:k + nchans // 2048]])
                assert got[k][1] == pytest.approx(f["fch1"])
                spectra.append(row)
                k += nchans // 2048
            else:
                spectra.extend(got[k][3].reshape(-1, nchans))
                k += 1
        spectra = np.array(spectra)
        assert spectra.shape == want.shape, (i, spectra.shape, want.shape, per_pkt)
        den = np.maximum(want, want.mean(axis=1, keepdims=True))
        assert float((np.abs(spectra - want) / den).max()) <= 1e-5


FBH5_OK = os.path.exists(os.path.join(os.path.dirname(CLI), "librawspec_fbh5.so")) and F.available()


def _fil_payload(path):
    raw = open(path, "rb").read()
    n = raw.index(b"HEADER_END") + len(b"HEADER_END")
    return np.frombuffer(raw[n:], dtype=np.float32)


@pytest.mark.skipif(not FBH5_OK, reason="FBH5 plugin or h5dump not available")
@pytest.mark.parametrize("case", ["one", "split", "ics"])
def test_fbh5_output(tmp_path, case):
    """-j writes FBH5 files holding the same spectra as the SIGPROC files of
    the same run, with the filterbank header as data set attributes."""
    nants = 1 if case == "one" else 2
    spec = StemSpec(obsnchan=8, nants=nants, ntpb=4096, files=[[0, 1, 2, 3]])
    stem = str(tmp_path / "h")
    paths = write_stem(spec, stem)
    fil, h5 = tmp_path / "fil", tmp_path / "h5"
    fil.mkdir()
    h5.mkdir()
    args = ["-f", "1024,8", "-t", "2,512", "-p", "-4,1"]
    if case == "ics":
        args += ["-i", "1.0,0.5"]
    run_cli([*args, "-d", str(fil), stem])
    r = run_cli([*args, "-j", "-d", str(h5), stem])
    assert "FBH5" in r.stdout
    _, _, cfg, _ = expected(spec, Nts=[1024, 8], Nas=[2, 512], Npolout=[-4, 1])
    for i in range(2):
        f = header_fields(spec, cfg, i, rawdatafile=os.path.basename(paths[0]))
        names = ({"one": [dict()], "split": [dict(ant=a) for a in range(nants)],
                  "ics": [dict(ics=True)]})[case]
        for kw in names:
            fa, da, info, data = F.read(RS.output_name(stem, i, dest=str(h5), **kw)[:-4] + ".h5")
            want = _fil_payload(RS.output_name(stem, i, dest=str(fil), **kw))
            assert fa["CLASS"] == "FILTERBANK" and fa["VERSION"] == "2.0"
            assert info["shape"][1:] == (f["nifs"], f["nchans"])
            assert info["chunk"] == (cfg.Nds[i], 1, cfg.Nts[i])
            assert np.array_equal(data.reshape(-1), want)
            for k in ("nchans", "nfpc", "nifs", "nbits", "telescope_id", "machine_id"):
                assert da[k] == f[k], k
            for k in ("fch1", "foff", "tstart", "tsamp", "src_raj", "src_dej"):
                assert da[k] == f[k], k
            assert da["source_name"] == f["source_name"]
