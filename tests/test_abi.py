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
    for h in ("rawspec.h", "rawspec_amd.h"):
        txt = open(os.path.join(INC, h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b((?:rawspec|get)_\w+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function(rlib):
    from rawspec_amd.build import exported_symbols
    declared = _declared_functions()
    assert {"rawspec_initialize", "rawspec_cleanup", "rawspec_start_processing",
            "rawspec_wait_for_completion", "get_cufft_version"} <= declared
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
