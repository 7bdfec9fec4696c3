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


if __name__ == "__main__":
    fft_vectors()
    if "--layout" in sys.argv:
        ref = "/root/reference"
        res = layout(ref)
        json.dump(res, open(os.path.join(HERE, "rawspec_ctx_layout.json"), "w"), indent=1)
        print("reference layout:", res)
