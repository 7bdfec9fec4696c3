"""Runs the rawspec program on synthetic stems and builds what its output
files must hold from the oracle: the block sequence and headers restated in
oracle/rawstem.py, the spectra from the FP64 oracle."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

import oracle as O
import oracle.rawstem as RS
from tests.raw_stems import StemSpec, block_data, pktidx_of, read_fil, seq_of, write_stem
from tests.spectro_cases import PLAIN_TOL, rel_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# RAWSPEC_AMD_CLI names another build of the program (e.g. a sanitizer build)
CLI = os.environ.get("RAWSPEC_AMD_CLI") or os.path.join(ROOT, "rawspec_amd", "bin", "rawspec")


def run_cli(args, timeout=300, check=True):
    if not os.path.exists(CLI):
        raise RuntimeError(f"{CLI} is missing: build it with `python -m rawspec_amd.build`")
    print("rawspec", " ".join(args), flush=True)
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)
    if check and r.returncode != 0:
        raise AssertionError(f"rawspec failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    return r


def _files_meta(spec: StemSpec):
    """PKTIDX and completeness of every block of every file."""
    meta = []
    for fi, steps in enumerate(spec.files):
        last = fi == len(spec.files) - 1
        m = []
        for k, s in enumerate(steps):
            complete = not (last and spec.truncate_last and k == len(steps) - 1)
            m.append((pktidx_of(spec, s), complete))
        meta.append(m)
    return meta


def cfg_for(spec: StemSpec, Nc, Nts, Nas, Npolout, ics=False, weights=None):
    No = len(Nts)
    cfg = O.make_cfg(No=No, Np=spec.npol, Nant=spec.nants, Nc=Nc, Ntpb=spec.ntpb,
                     Nbps=spec.nbits, Npolout=list(Npolout), Nts=list(Nts), Nas=list(Nas),
                     input_conjugated=1 if spec.obsbw < 0 else 0,
                     incoherently_sum=1 if ics else 0,
                     Naws=len(weights) if weights is not None else 0)
    assert O.derive(cfg) == 0
    return cfg


def expected(spec: StemSpec, *, Nts, Nas, Npolout, schan=0, nchan=0, ics=False, weights=None):
    """Per product: list of (power, Nds) dumps, list of (ics, Nds) dumps; the
    oracle's cfg; the end reason."""
    Nc = nchan or spec.obsnchan
    cfg = cfg_for(spec, Nc, Nts, Nas, Npolout, ics, weights)
    buffers, reason = RS.sequence(_files_meta(spec), cfg.Nb)
    st = O.State(cfg, weights=weights)
    pw = {i: [] for i in range(cfg.No)}
    ic = {i: [] for i in range(cfg.No)}
    cache = {}
    for buf in buffers:
        blocks = []
        for e in buf:
            if e is RS.ZERO:
                blocks.append(np.zeros(Nc * spec.ntpb * spec.npol * 2,
                                       dtype=np.int16 if spec.nbits == 16 else np.int8))
            else:
                fi, k = e
                seq = seq_of(spec.files[fi][k])
                if seq not in cache:
                    cache[seq] = block_data(spec, seq)
                d = cache[seq].reshape(spec.obsnchan, -1)[schan:schan + Nc]
                blocks.append(np.ascontiguousarray(d).reshape(-1))
        dev = O.device_layout(blocks, Nc, spec.ntpb, spec.npol)
        _, outs, icss = st.process(dev, 1)
        for i in outs:
            pw[i].append((outs[i], cfg.Nds[i]))
            if ics:
                ic[i].append((icss[i], cfg.Nds[i]))
    st.close()
    return pw, ic, cfg, reason


def header_fields(spec: StemSpec, cfg, i, *, schan=0, rawdatafile):
    import rawspec_amd as R
    L = R.lib()
    raw = R.RawHdr()
    h = (header_text(spec) + b"\0")
    L.rawspec_raw_parse_header(C.create_string_buffer(h), C.byref(raw))
    return RS.fb_fields(obsfreq=raw.obsfreq, obsbw=raw.obsbw, obsnchan=raw.obsnchan,
                        nants=raw.nants, tbin=raw.tbin, mjd=raw.mjd, ra=raw.ra, dec=raw.dec,
                        telescope_id=L.fb_telescope_id(raw.telescop),
                        src_name=raw.src_name.decode(), rawdatafile=rawdatafile,
                        beam_id=raw.beam_id, refbeam=raw.refbeam, Nc=cfg.Nc,
                        Nt=cfg.Nts[i], Na=cfg.Nas[i], Npolout=cfg.Npolout[i], schan=schan)


def header_text(spec: StemSpec) -> bytes:
    from tests.raw_stems import header
    return header(spec, spec.pktidx0)


def fb_header_bytes(fields) -> bytes:
    import rawspec_amd as R
    L = R.lib()
    h = R.FbHdr()
    for k, v in fields.items():
        setattr(h, k, v.encode() if isinstance(v, str) else v)
    buf = C.create_string_buffer(8192)
    end = L.fb_buf_write_header(buf, C.byref(h))
    return buf.raw[: end - C.addressof(buf)]


def check_fil(path, fields, dumps, *, nifs, Nc, Nt, Npolout, tol=PLAIN_TOL):
    """Header bytes identical to the expected header; spectra within tol of
    the oracle's (the shared error metric)."""
    raw = open(path, "rb").read()
    want_hdr = fb_header_bytes(fields)
    assert raw[: len(want_hdr)] == want_hdr, path
    got = np.frombuffer(raw[len(want_hdr):], dtype=np.float32)
    want = np.concatenate([np.asarray(d, dtype=np.float64) for d, _ in dumps]) if dumps else np.zeros(0)
    assert got.size == want.size, (path, got.size, want.size)
    if not dumps:
        return 0.0
    nds_total = sum(n for _, n in dumps)
    mode = 1 if Npolout == 1 else Npolout
    err = rel_err(got, want, nds_total, nifs, Nc, Nt, mode)
    worst = float(err.max())
    assert worst <= tol, (path, worst)
    return worst


def check_split(paths, fields, dumps, *, nifs, Nant, Ncpa, Nt, Npolout, tol=PLAIN_TOL):
    """Per-antenna files against the oracle's dumps split per antenna."""
    want = RS.file_payload(dumps, nifs=nifs, Nant=Nant, per_ant=True)
    hdr = fb_header_bytes(fields)
    nds_total = sum(n for _, n in dumps)
    mode = 1 if Npolout == 1 else Npolout
    worst = 0.0
    for a, p in enumerate(paths):
        raw = open(p, "rb").read()
        assert raw[: len(hdr)] == hdr, p
        got = np.frombuffer(raw[len(hdr):], dtype=np.float32)
        ref = np.frombuffer(want[a], dtype=np.float32).astype(np.float64)
        assert got.size == ref.size, (p, got.size, ref.size)
        if ref.size:
            worst = max(worst, float(rel_err(got, ref, nds_total, nifs, Ncpa, Nt, mode).max()))
    assert worst <= tol, worst
    return worst


__all__ = ["StemSpec", "write_stem", "read_fil", "run_cli", "expected", "header_fields",
           "fb_header_bytes", "check_fil", "check_split", "CLI"]
