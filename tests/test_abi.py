"""The C-ABI boundary, checked without a GPU:
  * struct rawspec_context_s in include/rawspec.h has the reference's layout
    (golden offsets from the reference header, tests/golden),
  * the ctypes mirror agrees with the compiled header,
  * librawspec.so loads and exports every function include/*.h declares,
  * rawspec_initialize() applies the reference's validation and derivation
    rules exactly as the oracle restates them (these run before any device
    call, so they are observable on a CPU-only machine).
"""
import ctypes as C
import json
import os
import random
import re
import subprocess
import tempfile

import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
GOLD = os.path.join(ROOT, "tests", "golden", "rawspec_ctx_layout.json")


@pytest.fixture(scope="session")
def rlib():
    import rawspec_amd as R
    if not os.path.exists(R.lib_path()):
        from rawspec_amd import build
        build.build()
    return R


def test_header_layout_matches_reference():
    from tests.golden.make_golden import layout
    ours = layout(INC)
    ref = json.load(open(GOLD))
    assert ours == ref


def test_ctypes_mirror_matches_header(rlib):
    ref = json.load(open(GOLD))
    st = rlib.RawspecContext
    assert C.sizeof(st) == ref["sizeof"]
    for name, _ in st._fields_:
        assert getattr(st, name).offset == ref[name], name


def _declared_functions():
    names = set()
    for h in ("rawspec.h", "rawspec_amd.h", "rawspec_rawutils.h", "rawspec_fbutils.h"):
        txt = open(os.path.join(INC, h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b((?:rawspec|get|fb)_\w+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function(rlib):
    from rawspec_amd.build import exported_symbols
    declared = _declared_functions()
    assert {"rawspec_initialize", "rawspec_cleanup", "rawspec_start_processing",
            "rawspec_wait_for_completion", "get_cufft_version", "rawspec_raw_read_header",
            "fb_fd_write_padded_header", "fb_telescope_id"} <= declared
    exported = set(exported_symbols(rlib.lib_path()))
    missing = declared - exported
    assert not missing, missing
    lib = rlib.lib()
    for name in declared:
        assert hasattr(lib, name)


def test_version_strings(rlib):
    assert rlib.lib().get_librawspec_version().decode()
    assert "HIP" in rlib.lib().get_cufft_version().decode()


def _lib_init(rlib, kw):
    c = rlib.RawspecContext()
    for k in ("No", "Np", "Nc", "Ntpb", "Nbps", "Nbc", "Nb", "Nant"):
        if k in kw:
            setattr(c, k, kw[k])
    for i in range(len(kw.get("Nts", ()))):
        c.Nts[i] = kw["Nts"][i]
        c.Nas[i] = kw["Nas"][i]
        c.Npolout[i] = kw["Npolout"][i]
    rc = rlib.lib().rawspec_initialize(C.byref(c))
    if rc == 0:                      # a GPU is present: release it again
        rlib.lib().rawspec_cleanup(C.byref(c))
    return rc, c


def _random_cfg(rng):
    No = rng.randint(1, 4)
    Ntpb = rng.choice([1, 3, 5]) * (1 << rng.randint(4, 12))
    nts = [1 << rng.randint(0, 14) for _ in range(No)]
    if rng.random() < 0.15:
        nts[0] = rng.choice([96, 1000, 0])
    nas = [rng.choice([1, 2, 3, 4, 16, 51, 128, 0 if rng.random() < 0.05 else 8]) for _ in range(No)]
    return dict(No=No, Np=rng.choice([1, 2, 2, 3]), Nc=rng.choice([1, 4, 6, 10, 88]), Ntpb=Ntpb,
                Nbps=rng.choice([0, 4, 8, 16, 16, 12]), Nbc=rng.choice([0, 0, 1, 2, 4]),
                Nb=rng.choice([0, 0, 0, 1, 2, 3]), Nant=rng.choice([0, 1, 2]),
                Nts=nts, Nas=nas, Npolout=[rng.choice([1, 4, -4, 2]) for _ in range(No)])


def test_library_validation_matches_oracle(rlib, capfd):
    rng = random.Random(2024)
    nvalid = ninvalid = 0
    for _ in range(3000):
        kw = _random_cfg(rng)
        cfg = O.make_cfg(**kw)
        orc = O.derive(cfg)
        rc, c = _lib_init(rlib, kw)
        if orc != 0:
            ninvalid += 1
            assert rc == 1 and not c.gpu_ctx, (kw, O.ERRORS[orc])
            continue
        nvalid += 1
        # derived before any device work: identical to the oracle's
        assert (c.Nb, c.Nb_host, c.Nbc, c.Nbps, c.Ntmax, c.Nant) == \
               (cfg.Nb, cfg.Nb_host, cfg.Nbc, cfg.Nbps, cfg.Ntmax, cfg.Nant), kw
        for i in range(kw["No"]):
            assert c.Npolout[i] == cfg.Npolout[i]
            assert c.Nds[i] == cfg.Nds[i]
            assert c.h_pwrbuf_size[i] == 4 * cfg.pwr_len[i]
    capfd.readouterr()
    assert nvalid > 50 and ninvalid > 50


def test_plan_validates_like_initialize_and_sizes_buffers(rlib, capfd):
    """rawspec_amd_plan() applies rawspec_initialize()'s validation and
    derivation without allocating (it runs here, with no GPU) and reports at
    least the input buffer plus one integration buffer per product on the
    device, and exactly the pinned host buffers rawspec_initialize() would
    allocate."""
    rng = random.Random(77)
    nvalid = ninvalid = 0
    for _ in range(1500):
        kw = _random_cfg(rng)
        cfg = O.make_cfg(**kw)
        orc = O.derive(cfg)
        c = rlib.RawspecContext()
        for k in ("No", "Np", "Nc", "Ntpb", "Nbps", "Nbc", "Nb", "Nant"):
            setattr(c, k, kw[k])
        for i in range(kw["No"]):
            c.Nts[i], c.Nas[i], c.Npolout[i] = kw["Nts"][i], kw["Nas"][i], kw["Npolout"][i]
        dev, host = C.c_size_t(0), C.c_size_t(0)
        rc = rlib.lib().rawspec_amd_plan(C.byref(c), C.byref(dev), C.byref(host))
        # lengths: powers of two, or any other length up to 2048 (Bluestein)
        lengths_ok = all(n & (n - 1) == 0 or n <= 2048 for n in kw["Nts"][:kw["No"]])
        if orc != 0 or not lengths_ok:
            ninvalid += 1
            assert rc == 1, kw
            continue
        nvalid += 1
        assert rc == 0, kw
        assert not c.gpu_ctx
        assert c.Nds[0] == cfg.Nds[0] and c.Nb == cfg.Nb
        pwr = sum(c.h_pwrbuf_size[i] for i in range(kw["No"]))
        blocks = cfg.Nb_host * kw["Ntpb"] * kw["Np"] * kw["Nc"] * 2 * (cfg.Nbps // 8)
        assert host.value == pwr + blocks, kw
        in_bytes = cfg.Nb * kw["Ntpb"] * kw["Np"] * kw["Nc"] * 2 * (cfg.Nbps // 8)
        assert dev.value >= in_bytes + pwr, kw
    capfd.readouterr()
    assert nvalid > 50 and ninvalid > 50


def test_plan_bytes_for_the_64_channel_three_product_workload(rlib):
    """64 channels x 2 pols x 8 bits, 2 blocks of 2^19 samples; products
    2^20 x 51, 8 x 128, 1024 x 3072, full Stokes.  The device plan holds the
    256 MiB input, the integration buffers (1 GiB for 2^20 and its 1 GiB
    inter-buffer accumulator since it integrates over 51 buffers, 8 MiB for
    Nfft 8 with 1024 spectra per dump, 1 MiB for 1024), the 1536 MiB
    four-step intermediate (all 64 channels of one buffer, or batches of 8
    buffers x 12 channels), the 8 GiB store of first-pass output the 2^20
    product holds across one-buffer calls (8 buffers x 1 GiB), and twiddle
    tables; the rest is the small path's partial sums, bounded by one more
    integration buffer set."""
    c = rlib.RawspecContext()
    c.No, c.Np, c.Nc, c.Ntpb, c.Nbps, c.Nb = 3, 2, 64, 1 << 19, 8, 2
    for i, (nt, na) in enumerate([(1 << 20, 51), (8, 128), (1024, 3072)]):
        c.Nts[i], c.Nas[i], c.Npolout[i] = nt, na, -4
    dev, host = C.c_size_t(0), C.c_size_t(0)
    assert rlib.lib().rawspec_amd_plan(C.byref(c), C.byref(dev), C.byref(host)) == 0
    MiB = 1 << 20
    fixed = 256 * MiB + 2 * 1024 * MiB + 8 * MiB + 1 * MiB + 1536 * MiB + 8 * 1024 * MiB
    tw = (4096 + 1024) * 8          # master W_4096 table, pass A's fine table W_N^r, r < 1024
    assert fixed + tw <= dev.value <= fixed + tw + 64 * MiB, dev.value
    # pinned host: three dump buffers (blocks are allocated too: h_blkbufs NULL)
    assert host.value == 1024 * MiB + 8 * MiB + 1 * MiB + 2 * 256 * MiB // 2


def test_no_kernel_needs_scratch_memory():
    """Every kernel of the built library keeps its values in registers: the
    HIP runtime would otherwise allocate scratch memory for every wave the
    device can hold (hundreds of MiB) outside what rawspec_amd_plan()
    reports and admission control counts.  Read from the compiler's
    resource report that the build writes."""
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "rawspec_amd", "_build", "kernel_resources.tsv")
    if not os.path.exists(path):
        pytest.skip("the library was not built here")
    with open(path) as f:
        head = f.readline().rstrip("\n").split("\t")
        rows = [dict(zip(head, line.rstrip("\n").split("\t"))) for line in f]
    assert len(rows) > 100
    spill = [r["kernel"] for r in rows if int(r["ScratchSize [bytes/lane]"]) != 0]
    assert not spill, spill[:5]
