"""Writing GUPPI RAW files: one header of 80-byte "KEYWORD = value" cards
ended by an END card (padded with blanks to a multiple of 512 bytes when
DIRECTIO = 1), then BLOCSIZE bytes of samples [chan][time][pol][re, im];
repeated per block.  Files of a stem are STEM.0000.raw, STEM.0001.raw, ...
Reading them is the C library's job (rawspec_raw_read_header)."""
from __future__ import annotations


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
