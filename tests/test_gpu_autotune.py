"""GPU checks of the launch-plan autotuner (rawspec_amd_autotune): after
tuning for one-buffer or batch calls, and when tuning in the middle of a run
(which restarts every integration, as rawspec_reset_integration does), the
spectra match the FP64 oracle at the FP32 bar of spectro_cases.py and no
dump fires during tuning; the report names every product and never picks a
plan slower than the default.
"""
import re

import pytest

from tests.spectro_cases import (MFMA_FAST, MFMA_SPLIT, MFMA_FLOOR_MULT, PLAIN_TOL, check_results,
                                 run_calls, unsplit_products)

pytestmark = pytest.mark.gpu

LINE = re.compile(r"p(\d+) N(\d+) (small|fourstep) .*: ([\d.]+) ms \(default ([\d.]+) ms, "
                  r"(\d+) candidates\)")


def _check_report(rep, nts):
    lines = rep.strip().splitlines()
    print("  " + "\n  ".join(lines), flush=True)
    assert len(lines) == len(nts)
    for i, (ln, nt) in enumerate(zip(lines, nts)):
        m = LINE.match(ln)
        assert m, ln
        assert int(m.group(1)) == i and int(m.group(2)) == nt
        assert m.group(3) == ("small" if nt <= 4096 else "fourstep")
        assert float(m.group(4)) <= float(m.group(5)) and int(m.group(6)) >= 1


def test_autotune_then_process_three_products():
    kw = dict(Nc=4, Ntpb=65536, Nts=[131072, 8, 1024], Nas=[3, 128, 64], Npolout=[-4, -4, -4],
              tone=(1, 20.0, 300.0, 1024.0))
    r, cfg, got, reps = run_calls(["autotune"] + [1] * 6, **kw)
    _check_report(reps[0], kw["Nts"])
    check_results(r, cfg, tol=PLAIN_TOL)
    assert [len(got[i]) for i in range(3)] == [2, 6, 6]


def test_autotune_mid_run_restarts_integrations():
    # two buffers into an integration of three, tune, then a batch call
    kw = dict(Nc=2, Ntpb=65536, Nts=[131072, 64], Nas=[3, 2048], Npolout=[4, 1],
              tone=(0, 20.0, 300.0, 1024.0))
    r, cfg, got, reps = run_calls([1, 1, "autotune:3", 3, 1, 1, 1], **kw)
    _check_report(reps[0], kw["Nts"])
    check_results(r, cfg, tol=PLAIN_TOL)
    assert len(got[0]) == 2          # the interrupted integration never dumps


def test_autotune_matrix_core_engine():
    kw = dict(Nc=2, Ntpb=1 << 20, Nts=[1 << 20, 1024], Nas=[2, 64], Npolout=[-4, -4], mfma=MFMA_SPLIT,
              tone=(1, 20.0, 300.0, 1024.0))
    r, cfg, got, reps = run_calls(["autotune", 2, 1, 1], **kw)
    _check_report(reps[0], kw["Nts"])
    check_results(r, cfg, tol=PLAIN_TOL, floor_mult=MFMA_FLOOR_MULT)


def test_autotune_fused_pair_at_level_1():
    # at engine level 1 the 8-point and the 1024-point products run as one
    # kernel: the 1024-point product's candidates retune it, the 8-point
    # product's leave it as it is; one-buffer and batch calls afterwards
    kw = dict(Nc=4, Ntpb=65536, Nts=[131072, 8, 1024], Nas=[3, 128, 64], Npolout=[-4, -4, -4],
              mfma=MFMA_FAST, tone=(1, 20.0, 300.0, 1024.0))
    r, cfg, got, reps = run_calls(["autotune", 1, 3, 1, 1], **kw)
    _check_report(reps[0], kw["Nts"])
    check_results(r, cfg, tol=PLAIN_TOL,
                  tol_by_product=unsplit_products(kw["Nts"], kw["Nas"], mfma=MFMA_FAST))
    assert [len(got[i]) for i in range(3)] == [2, 6, 6]
