"""Writes the committed golden fixtures under tests/golden/.

    python tests/golden/make_golden.py            # FFT vectors (numpy, FP64)
    python tests/golden/make_golden.py --layout   # + reference struct layout

fft_vectors.npz        seeded complex inputs and numpy.fft outputs (FP64):
                       fwd_<n> = np.fft.fft(x_<n>)       (kernel exp(-2 pi i km/n))
                       inv_<n> = np.fft.ifft(x_<n>) * n  (kernel exp(+2 pi i km/n),
                                                          unnormalised)
rawspec_ctx_layout.json  sizeof/offsetof of struct rawspec_context_s as laid
                       out by the reference header (rawspec.h:32-148), from a
                       probe compiled here against that header with gcc.  Only
                       the numbers are committed.
host_formats_layout.json  the same for rawspec_raw_hdr_t
                       (rawspec_rawutils.h:7-28) and fb_hdr_t
                       (rawspec_fbutils.h:70-122)          (--layout)
host_formats.json      GUPPI RAW headers and filterbank header fields with
                       what the reference's own rawspec_rawutils.c /
                       rawspec_fbutils.c, compiled into oracle/_ref by
                       `make -C oracle ref`, return for them   (--host)
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SIZES = [1, 2, 3, 5, 7, 8, 12, 15, 16, 64, 97, 128, 1000, 1024, 4096]
FIELDS = ["No", "Np", "Nant", "Nc", "Ntpb", "exit_soon", "Nbc", "Nbps", "Npolout", "Nts",
          "Nas", "dump_callback", "user_data", "Nb", "Nb_host", "h_blkbufs", "gpu_index",
          "input_conjugated", "incoherently_sum", "Naws", "Aws", "h_pwrbuf", "h_pwrbuf_size",
          "h_icsbuf", "Nds", "Ntmax", "gpu_ctx"]


def fft_vectors():
    out = {}
    for n in SIZES:
        rng = np.random.default_rng(1234 + n)
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x_{n}"] = x
        out[f"fwd_{n}"] = np.fft.fft(x)
        out[f"inv_{n}"] = np.fft.ifft(x) * n
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


def probe_source():
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "rawspec.h"',
             "int main(void) {",
             '  printf("{\\"sizeof\\": %zu", sizeof(rawspec_context));']
    for f in FIELDS:
        lines.append(f'  printf(", \\"{f}\\": %zu", offsetof(rawspec_context, {f}));')
    lines += ['  printf("}\\n");', "  return 0;", "}"]
    return "\n".join(lines) + "\n"


def layout(include_dir):
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(src, "w").write(probe_source())
        subprocess.run(["gcc", "-I", include_dir, src, "-o", exe], check=True)
        return json.loads(subprocess.run([exe], check=True, capture_output=True,
                                         text=True).stdout)


HOST_STRUCTS = {
    ("rawspec_rawutils.h", "rawspec_raw_hdr_t"): [
        "directio", "blocsize", "npol", "obsnchan", "nbits", "pktidx", "obsfreq", "obsbw",
        "tbin", "ra", "dec", "mjd", "beam_id", "nbeam", "refbeam", "nants", "src_name",
        "telescop", "hdr_pos", "hdr_size"],
    ("rawspec_fbutils.h", "fb_hdr_t"): [
        "machine_id", "telescope_id", "data_type", "barycentric", "pulsarcentric", "src_raj",
        "src_dej", "az_start", "za_start", "fch1", "foff", "nchans", "nbeams", "ibeam", "nbits",
        "tstart", "tsamp", "nifs", "source_name", "rawdatafile", "nfpc", "refbeam"],
}


def host_layouts(include_dir):
    """{type: {"sizeof": .., field: offset}} of the host-format structs."""
    lines = ["#include <stdio.h>", "#include <stddef.h>"]
    lines += [f'#include "{h}"' for (h, _t) in HOST_STRUCTS]
    lines += ["int main(void) {", '  printf("{");']
    first = True
    for (_h, t), fields in HOST_STRUCTS.items():
        sep = "" if first else ", "
        first = False
        lines.append(f'  printf("{sep}\\"{t}\\": {{\\"sizeof\\": %zu", sizeof({t}));')
        for f in fields:
            lines.append(f'  printf(", \\"{f}\\": %zu", offsetof({t}, {f}));')
        lines.append('  printf("}");')
    lines += ['  printf("}\\n");', "  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(src, "w").write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-I", include_dir, src, "-o", exe], check=True)
        return json.loads(subprocess.run([exe], check=True, capture_output=True,
                                         text=True).stdout)


def host_fixtures():
    """Cases from tests/host_cases.py answered by the reference's own code
    (oracle/_ref)."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from tests import host_cases as H
    ref_raw, ref_fb = H.ref_libs()
    raw = [{"header": h.decode("latin-1"), "expect": H.raw_answers(ref_raw, h)}
           for h in H.raw_headers(16, seed=99)]
    fb = [{"hdr": H.fb_fields(hd), "minlen": m, "bytes": H.fb_bytes(ref_fb, hd, m).hex()}
          for hd, m in H.fb_headers(12, seed=98)]
    json.dump({"raw": raw, "fb": fb}, open(os.path.join(HERE, "host_formats.json"), "w"),
              indent=0)


if __name__ == "__main__":
    fft_vectors()
    if "--layout" in sys.argv:
        ref = "/root/reference"
        res = layout(ref)
        json.dump(res, open(os.path.join(HERE, "rawspec_ctx_layout.json"), "w"), indent=1)
        print("reference layout:", res)
        res = host_layouts(ref)
        json.dump(res, open(os.path.join(HERE, "host_formats_layout.json"), "w"), indent=1)
        print("reference host-format layouts:", res)
    if "--host" in sys.argv:
        host_fixtures()
