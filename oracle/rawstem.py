"""CPU restatement of how the reference rawspec program turns a GUPPI RAW
stem into spectrometer buffers and filterbank files.  Test infrastructure
only: the tests compare the program built on librawspec against it.

  sequence()      the block loop, rawspec.c:905-1019 (PKTIDX gaps, stops)
  fb_fields()     the filterbank header, rawspec.c:541-551 and 806-832
  file_payload()  what each output file holds, rawspec_file.c:102-202
  output_name()   output file names, rawspec_file.c:15-45, :73-98
"""
from __future__ import annotations

import numpy as np

ZERO = None   # a zero-filled block in a sequence


def sequence(files, Nb):
    """files: per file, the PKTIDX of each block in file order, with
    `complete` flags: [[(pktidx, complete), ...], ...].  Returns the
    emitted blocks as (file, index) pairs or ZERO, grouped into the complete
    buffers of Nb blocks that get processed (a trailing partial buffer is
    dropped), plus the reason the stem ended.

    rawspec.c:905-1019: the first PKTIDX increase fixes the step; a block
    `k` steps after the previous one is preceded by k-1 zero blocks; a
    backward or misaligned step ends the stem; a repeated PKTIDX (once the
    step is known) ends the file; a block cut short by the end of the file
    ends the stem."""
    emitted = []
    pktidx = None
    step = 0
    reason = "end"
    stop = False
    for fi, blocks in enumerate(files):
        for k, (pi, complete) in enumerate(blocks):
            if pktidx is None:
                pktidx = pi
            if step == 0 and pi > pktidx:
                step = pi - pktidx
            if pi - pktidx != step:
                if pi < pktidx:
                    reason, stop = "backwards", True
                    break
                if (pi - pktidx) % step != 0:
                    reason, stop = "misaligned", True
                    break
                if pi == pktidx:
                    reason = "null"
                    break            # rest of this file skipped
                while pi - pktidx != step:
                    pktidx += step
                    emitted.append(ZERO)
            if not complete:
                reason, stop = "incomplete", True
                break
            emitted.append((fi, k))
            pktidx = pi
        if stop:
            break
    nfull = len(emitted) // Nb
    return [emitted[b * Nb:(b + 1) * Nb] for b in range(nfull)], reason


def fb_fields(*, obsfreq, obsbw, obsnchan, nants, tbin, mjd, ra, dec, telescope_id,
              src_name, rawdatafile, beam_id, refbeam, Nc, Nt, Na, Npolout, schan):
    """Filterbank header fields of one output product (rawspec.c:541-551,
    806-832).  Nc: coarse channels processed; schan: the first of them
    (absolute, antenna offset included).  Integer divisions are C's."""
    ncpa = obsnchan // nants
    foff = obsbw / ncpa / Nt
    fch1 = (obsfreq - obsbw * (ncpa - 1) / ((2 * obsnchan) // nants)
            - (Nt // 2) * foff + (schan % ncpa) * obsbw / ncpa)
    return dict(machine_id=20, telescope_id=telescope_id, data_type=1, barycentric=0,
                pulsarcentric=0, src_raj=ra, src_dej=dec, az_start=0.0, za_start=0.0,
                fch1=fch1, foff=foff, nchans=(Nc * Nt) // nants, nbeams=1, ibeam=beam_id,
                nbits=32, tstart=mjd, tsamp=tbin * Nt * Na, nifs=abs(Npolout),
                source_name=src_name[:80], rawdatafile=rawdatafile[:80], nfpc=Nt,
                refbeam=refbeam)


def fine_channel_freqs(f):
    """Centre frequency of every fine channel a header describes."""
    return f["fch1"] + f["foff"] * np.arange(f["nchans"])


def file_payload(dumps, *, nifs, Nant, per_ant):
    """Bytes of the spectra in each output file, given the dumps as
    (power buffer, Nds) pairs, each buffer [Nds][nifs][Nc][Nt] float32.  One
    file: the buffers back to back (rawspec_file.c:111-129).  Split per
    antenna: for every dump spectrum and pol, antenna a's run of Nc/Nant
    channels goes to file a (rawspec_file.c:132-170)."""
    if not per_ant or Nant == 1:
        return [b"".join(np.asarray(d, dtype=np.float32).tobytes() for d, _ in dumps)]
    outs = [bytearray() for _ in range(Nant)]
    for d, nds in dumps:
        d = np.asarray(d, dtype=np.float32).reshape(nds, nifs, Nant, -1)
        for k in range(nds):
            for j in range(nifs):
                for a in range(Nant):
                    outs[a] += d[k, j, a].tobytes()
    return [bytes(o) for o in outs]


def output_name(stem, idx, dest=None, ant=None, ics=False):
    """Output file name of product idx (rawspec_file.c:15-45, 73-98)."""
    s = stem + ("-ics" if ics else "") + (f"-ant{ant:03d}" if ant is not None else "")
    if dest:
        s = dest + "/" + s.rsplit("/", 1)[-1]
    return f"{s}.rawspec.{idx:04d}.fil"
