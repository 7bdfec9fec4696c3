"""Writing GUPPI RAW files: one header of 80-byte "KEYWORD = value" cards
ended by an END card (padded with blanks to a multiple of 512 bytes when
DIRECTIO = 1), then BLOCSIZE bytes of samples [chan][time][pol][re, im];
repeated per block.  Files of a stem are STEM.0000.raw, STEM.0001.raw, ...
Reading them is the C library's job (rawspec_raw_read_header).

As a program it writes a synthetic observation, e.g. 32 blocks of 64
channels x 2 pols of complex int8 Gaussian noise, 524288 samples per block:
    python -m rawspec_amd.rawfile STEM --nchan 64 --ntpb 524288 --blocks 32"""
from __future__ import annotations

import argparse
import sys


def card(key: str, value) -> str:
    """One 80-character card.  Strings are quoted; numbers right-aligned in
    the value field (columns 11-30)."""
    v = f"'{value}'" if isinstance(value, str) else (repr(float(value)) if isinstance(value, float)
                                                     else str(value))
    return f"{key[:8]:<8}= {v:>20}"[:80].ljust(80)


def header(cards, directio: int = 0) -> bytes:
    """Header bytes for (key, value) pairs."""
    h = ("".join(card(k, v) for k, v in cards) + "END".ljust(80)).encode("ascii")
    if directio:
        h += b" " * ((-len(h)) % 512)
    return h


def observation_cards(*, obsnchan, npol, nbits, ntpb, pktidx, obsfreq=1420.0, obsbw=187.5,
                      tbin=4.096e-5, nants=1, directio=0, telescop="GBT", src_name="SYNTHETIC",
                      ra_str="17:10:03.984", dec_str="12:24:42.00", stt_imjd=59000,
                      stt_smjd=43200, extra=()):
    """The cards rawspec reads, for one block."""
    blocsize = obsnchan * ntpb * npol * 2 * nbits // 8
    return [("BACKEND", "GUPPI   "), ("TELESCOP", telescop), ("OBSFREQ", float(obsfreq)),
            ("OBSBW", float(obsbw)), ("OBSNCHAN", obsnchan), ("NPOL", npol), ("NBITS", nbits),
            ("TBIN", float(tbin)), ("BLOCSIZE", blocsize), ("PKTIDX", pktidx),
            ("DIRECTIO", directio), ("NANTS", nants), ("SRC_NAME", src_name),
            ("RA_STR", ra_str), ("DEC_STR", dec_str), ("STT_IMJD", stt_imjd),
            ("STT_SMJD", stt_smjd), *extra]


def write_noise(stem: str, *, nchan: int, npol: int = 2, nbits: int = 8, ntpb: int,
                nblocks: int, sigma: float = 8.0, seed: int = 1, directio: int = 1) -> str:
    """STEM.0000.raw with nblocks blocks of complex Gaussian noise of `sigma`
    counts per component, rounded and clamped to the SNORM range of nbits
    (8 or 16); PKTIDX advances by 32 per block.  Returns the file name."""
    import numpy as np
    if nbits not in (8, 16):
        raise ValueError("nbits must be 8 or 16")
    lim, dt = (127, np.int8) if nbits == 8 else (32767, np.int16)
    rng = np.random.default_rng(seed)
    name = stem + ".0000.raw"
    with open(name, "wb") as f:
        for b in range(nblocks):
            f.write(header(observation_cards(obsnchan=nchan, npol=npol, nbits=nbits, ntpb=ntpb,
                                             pktidx=32 * b, directio=directio), directio))
            x = rng.standard_normal(nchan * ntpb * npol * 2, dtype=np.float32)
            x *= np.float32(sigma)
            np.rint(x, out=x)
            np.clip(x, -lim, lim, out=x)
            f.write(x.astype(dt).tobytes())
    return name


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="write a GUPPI RAW file of complex Gaussian noise")
    ap.add_argument("stem", help="writes STEM.0000.raw")
    ap.add_argument("--nchan", type=int, required=True)
    ap.add_argument("--npol", type=int, default=2)
    ap.add_argument("--nbits", type=int, default=8)
    ap.add_argument("--ntpb", type=int, required=True, help="time samples per block")
    ap.add_argument("--blocks", type=int, default=1)
    ap.add_argument("--sigma", type=float, default=8.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--directio", type=int, default=1)
    a = ap.parse_args(argv)
    print(write_noise(a.stem, nchan=a.nchan, npol=a.npol, nbits=a.nbits, ntpb=a.ntpb,
                      nblocks=a.blocks, sigma=a.sigma, seed=a.seed, directio=a.directio))
    return 0


if __name__ == "__main__":
    sys.exit(main())
