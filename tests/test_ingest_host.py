"""The rawspec program's ingest (rawspec_amd/csrc/cli/cli_ingest.c) on the
CPU: its read pool, staging slots and issuing thread run against host-only
stand-ins for HIP and for the library's processing calls, which checksum
every buffer they receive (tests/ingest_harness/).  Checked: every complete
buffer arrives once, in order, with the right bytes in the device layout
[chan][Nb][row], in both ingest modes, for several pool sizes, slot counts
and copy speeds, with PKTIDX-gap zero blocks; a read that runs past the end
of the file stops the stem after the buffers before it."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS_DIR = os.path.join(ROOT, "tests", "ingest_harness")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ingest") / "harness")
    # RAWSPEC_SANITIZE (e.g. "address,undefined" or "thread") instruments it
    san = os.environ.get("RAWSPEC_SANITIZE")
    sflags = [f"-fsanitize={san}", "-fno-omit-frame-pointer", "-g",
              "-fno-sanitize-recover=all"] if san else []
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-pthread", "-Wall", *sflags, f"-I{HARNESS_DIR}",
                    f"-I{ROOT}/include", f"-I{ROOT}/rawspec_amd/csrc/cli",
                    f"{ROOT}/rawspec_amd/csrc/cli/cli_ingest.c", f"{HARNESS_DIR}/harness.c",
                    "-o", out], check=True)
    return out


def _checksum(b: np.ndarray) -> int:
    w = (np.arange(b.size, dtype=np.uint64) % 65521) + 1
    return int((b.astype(np.uint64) * w).sum(dtype=np.uint64))


def _expected(data, nc, nb, cb, nblocks, zeros, file_blocks):
    """Checksums of the complete buffers, and how many come before a block
    the file cannot fill."""
    blocks = []
    f = 0
    for k in range(nblocks):
        if k in zeros:
            blocks.append(np.zeros(nc * cb, dtype=np.uint8))
        else:
            if f >= file_blocks:
                blocks.append(None)
            else:
                blocks.append(data[f * nc * cb:(f + 1) * nc * cb])
            f += 1
    sums = []
    for b0 in range(0, nblocks - nb + 1, nb):
        buf = blocks[b0:b0 + nb]
        if any(x is None for x in buf):
            break
        dev = np.stack([x.reshape(nc, cb) for x in buf], axis=1)   # [chan][Nb][row]
        sums.append(_checksum(dev.reshape(-1)))
    return sums


def _run(harness, path, nc, nb, cb, nblocks, mode, readers, slots, delay, zeros):
    r = subprocess.run([harness, path, str(nc), str(nb), str(cb), str(nblocks), str(mode),
                        str(readers), str(slots), str(delay), *map(str, sorted(zeros))],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    bufs = [ln.split() for ln in lines if ln.startswith("buf ")]
    end = [ln.split() for ln in lines if ln.startswith("end ")][0]
    assert [int(b[1]) for b in bufs] == list(range(len(bufs)))
    return [int(b[2]) for b in bufs], int(end[1]), r.stderr


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("readers,slots,delay", [(1, 2, 0), (3, 2, 300), (8, 3, 0), (8, 5, 500),
                                                 (16, 2, 100)])
def test_buffers_in_order_with_zero_blocks(harness, tmp_path, mode, readers, slots, delay):
    rng = np.random.default_rng(readers * 10 + slots)
    nc, nb, cb, nblocks = 600, 2, 96, 23          # 600 rows of 96 B: several jobs per block
    zeros = {5, 6, 13}
    file_blocks = nblocks - len(zeros)
    data = rng.integers(0, 256, file_blocks * nc * cb, dtype=np.uint8)
    p = str(tmp_path / "f.raw")
    data.tofile(p)
    got, rc, err = _run(harness, p, nc, nb, cb, nblocks, mode, readers, slots, delay, zeros)
    want = _expected(data, nc, nb, cb, nblocks, zeros, file_blocks)
    assert rc == 0 and len(want) == nblocks // nb
    assert got == want


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("nb,short_block", [(1, 7), (2, 7), (3, 4), (2, 0)])
def test_read_past_end_stops_after_earlier_buffers(harness, tmp_path, mode, nb, short_block):
    """The file holds fewer blocks than the stem asks for: every buffer
    before the first one that cannot be filled is processed, none after."""
    rng = np.random.default_rng(nb + short_block)
    nc, cb, nblocks = 40, 128, 12
    data = rng.integers(0, 256, short_block * nc * cb + 100, dtype=np.uint8)
    p = str(tmp_path / "f.raw")
    data.tofile(p)
    got, rc, err = _run(harness, p, nc, nb, cb, nblocks, mode, 4, 2, 0, set())
    want = _expected(data, nc, nb, cb, nblocks, set(), short_block)
    assert rc == 1 and "incomplete block at EOF" in err
    assert got == want


def test_many_buffers_stress(harness, tmp_path):
    """Many small buffers through the pipelined ingest with a slow copy: the
    pool, the slots and the issuing thread hand buffers over without loss,
    duplication or overlap of library calls."""
    rng = np.random.default_rng(99)
    nc, nb, cb, nblocks = 64, 1, 256, 400
    data = rng.integers(0, 256, nblocks * nc * cb, dtype=np.uint8)
    p = str(tmp_path / "f.raw")
    data.tofile(p)
    for readers, slots, delay in [(12, 2, 50), (2, 4, 0), (5, 3, 20)]:
        got, rc, _ = _run(harness, p, nc, nb, cb, nblocks, 1, readers, slots, delay, set())
        assert rc == 0 and got == _expected(data, nc, nb, cb, nblocks, set(), nblocks)
