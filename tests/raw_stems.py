"""Synthetic GUPPI RAW stems for the rawspec program's tests and bench.

A stem is a list of files STEM.0000.raw, STEM.0001.raw, ...; each file a
sequence of blocks, each block an ASCII header of 80-byte cards ended by an
END card (padded to a 512-byte boundary when DIRECTIO = 1) followed by
BLOCSIZE bytes of [chan][time][pol][re, im] samples.  Sample values come
from the oracle's seeded generator, keyed by the block's sequence number, so
a test can regenerate any block."""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

import oracle as O


@dataclass
class StemSpec:
    obsnchan: int = 8            # coarse channels in the file (all antennas)
    nants: int = 1
    npol: int = 2
    nbits: int = 8
    ntpb: int = 4096             # time samples per block
    obsfreq: float = 1420.0
    obsbw: float = 187.5
    tbin: float = 5.12e-6 * 8
    pktidx0: int = 1000
    dpktidx: int = 32
    src_name: str = "VOYAGER1"
    telescop: str = "GBT"
    ra_str: str = "17:10:03.984"
    dec_str: str = "12:24:42.00"
    stt_imjd: int = 59000
    stt_smjd: int = 43200
    beam_id: int | None = None
    refbeam: int | None = None
    directio: int = 0
    seed: int = 7
    sigma: float = 8.0
    tone: tuple | None = (3, 20.0, 300.0, 1024.0)   # (chan, amp, bin, nfft)
    # files: list of lists of PKTIDX steps (block k of a file has PKTIDX
    # pktidx0 + dpktidx * steps[k]; a fractional step makes a misaligned one)
    files: list = field(default_factory=lambda: [[0, 1, 2, 3]])
    truncate_last: int = 0       # bytes cut from the end of the last file

    @property
    def blocsize(self) -> int:
        return self.obsnchan * self.ntpb * self.npol * 2 * self.nbits // 8


def header(spec: StemSpec, pktidx: int) -> bytes:
    from rawspec_amd.rawfile import header as hdr, observation_cards
    extra = []
    if spec.beam_id is not None:
        extra.append(("BEAM_ID", spec.beam_id))
    if spec.refbeam is not None:
        extra.append(("REFBEAM", spec.refbeam))
    cards = observation_cards(obsnchan=spec.obsnchan, npol=spec.npol, nbits=spec.nbits,
                              ntpb=spec.ntpb, pktidx=pktidx, obsfreq=spec.obsfreq,
                              obsbw=spec.obsbw, tbin=spec.tbin, nants=spec.nants,
                              directio=spec.directio, telescop=spec.telescop,
                              src_name=spec.src_name, ra_str=spec.ra_str, dec_str=spec.dec_str,
                              stt_imjd=spec.stt_imjd, stt_smjd=spec.stt_smjd, extra=extra)
    return hdr(cards, spec.directio)


def pktidx_of(spec: StemSpec, step) -> int:
    return spec.pktidx0 + int(round(spec.dpktidx * step))


def seq_of(step) -> int:
    """Generator key of the block at a PKTIDX step."""
    return int(step) if float(step).is_integer() else 100000 + int(round(step * 1000))


def block_data(spec: StemSpec, seq: int) -> np.ndarray:
    """Samples of the block with generator key seq (see seq_of)."""
    if spec.nbits == 4:
        v = O.gen_block(spec.obsnchan, spec.ntpb, spec.npol, 8, spec.seed, 3.0, 127, seq,
                        *(spec.tone if spec.tone else (-1, 0.0, 0.0, 1.0)))
        return np.clip(v.astype(np.int16), -8, 7).astype(np.int8)
    clip = 32767 if spec.nbits == 16 else 127
    tc, ta, tb, tn = spec.tone if spec.tone else (-1, 0.0, 0.0, 1.0)
    return O.gen_block(spec.obsnchan, spec.ntpb, spec.npol, spec.nbits, spec.seed, spec.sigma,
                       clip, seq, tc, ta, tb, tn)


def block_bytes(spec: StemSpec, data: np.ndarray) -> bytes:
    if spec.nbits == 4:
        v = data.astype(np.int16)
        b = ((v[0::2] & 0xF) << 4) | (v[1::2] & 0xF)
        return b.astype(np.uint8).tobytes()
    return data.tobytes()


def write_stem(spec: StemSpec, stem: str) -> list[str]:
    """Writes the stem's files; returns their paths."""
    paths = []
    for fi, steps in enumerate(spec.files):
        p = f"{stem}.{fi:04d}.raw"
        with open(p, "wb") as f:
            for s in steps:
                f.write(header(spec, pktidx_of(spec, s)))
                f.write(block_bytes(spec, block_data(spec, seq_of(s))))
        paths.append(p)
    if spec.truncate_last and paths:
        sz = os.path.getsize(paths[-1])
        os.truncate(paths[-1], sz - spec.truncate_last)
    return paths


def read_fil(path: str):
    """(header fields, header length, float32 payload) of a filterbank file."""
    import ctypes as C

    import rawspec_amd as R
    raw = open(path, "rb").read()
    L = R.lib()
    h = R.FbHdr()
    n = C.c_size_t(0)
    buf = C.create_string_buffer(raw, len(raw))
    L.fb_buf_read_header(buf, C.byref(h), C.byref(n))
    return h, n.value, np.frombuffer(raw[n.value:], dtype=np.float32)
