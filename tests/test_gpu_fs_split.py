"""GPU parity of the split four-step pass B.  A four-step product with few
k2 tiles and channels would launch too few pass-B workgroups for the GPU,
each integrating a whole dump group; pass B then splits each group's spectra
into runs (one workgroup per run writes its sums, k_fs_b_reduce adds the
runs in order and applies pass B's finish).  Checked against the FP64 oracle
at the FP32 bar in every finish mode (dumps within a buffer; integrations
over several buffers one call, a batch or held buffers at a time), every
polarisation mode, one-pol input, and at engine level 1 (24-bit
intermediate) at its bar; and that the split really ran.
"""
import numpy as np
import pytest

from tests.spectro_cases import (MFMA_FAST, PLAIN_TOL, check_results, run_case, run_calls,
                                 unsplit_products)

pytestmark = pytest.mark.gpu

TONE = (1, 20.0, 300.0, 1024.0)


def _kernels(held=None, nbuf=2, **kw):
    """Names of the kernels nbuf buffers of this configuration launch, one
    call each (buffers may only hold their first-pass output until the
    group is full)."""
    import torch
    import rawspec_amd as R
    sp = R.Spectrometer(Nbps=8, **kw)
    if held is not None:
        assert sp.set_held_buffers(held) == 0
    d = torch.zeros(sp.device_input_size(), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    sp.profile(True)
    for _ in range(nbuf):
        assert sp.start_processing_device(d.data_ptr(), 1) == 0
    names = {k.split(".")[1] for k in sp.profile_report()}
    sp.close()
    return names


@pytest.mark.parametrize("nts,nas,held,nbuf", [(8192, 32, None, 2), (8192, 256, 0, 2),
                                               (131072, 32, None, 4)])
def test_split_selected(nts, nas, held, nbuf):
    """2 channels.  8192 points (4 k2 tiles of 16 per channel, 128 spectra
    per buffer): dump groups of 32 spectra within a buffer (4 runs each), or
    one buffer's 128 spectra per pass B of a two-buffer integration (16
    runs).  2^17 points (32 tiles of 8, 8 spectra per buffer): the pass B
    over 4 held buffers (4 runs of 8)."""
    kw = dict(Nc=2, Ntpb=1 << 20, Nts=[nts], Nas=[nas], Np=2, Npolout=[-4])
    names = _kernels(held=held, nbuf=nbuf, **kw)
    print(f"case: split selection N {nts} Nas {nas}: {sorted(names)}", flush=True)
    assert "k_fs_b" in names and "k_fs_b_reduce" in names


@pytest.mark.parametrize("pol", [1, 4, -4])
def test_split_dumps_within_a_buffer(pol):
    # 128 spectra per buffer, dumps of 32: 4 groups, each in 4 runs
    res, cfg, _ = run_case(Nc=2, Ntpb=1 << 20, Nts=[8192], Nas=[32], Npolout=[pol], nbuf=2,
                           tone=TONE)
    w = check_results(res, cfg, tol=PLAIN_TOL)
    print(f"  pol {pol}: worst {w}", flush=True)


@pytest.mark.parametrize("pattern,held", [([1, 1, 1, 1, 1, 1, 1, 1], None),
                                          ([1, 1, 1, 1, 1, 1, 1, 1], 0),
                                          ([4, 4], None), ([2, 1, 1, 3, 1], None)])
def test_split_integration_over_buffers(pattern, held):
    # Nis = 4: first / middle / last passes, held groups, batches, and a mix
    r, cfg, _, _ = run_calls(pattern, Nc=2, Ntpb=1 << 20, Nts=[8192, 16384], Nas=[512, 256],
                             Npolout=[-4, 4], held=held, tone=TONE)
    w = check_results(r, cfg, tol=PLAIN_TOL)
    print(f"  pattern {pattern} held {held}: worst {w}", flush=True)


def test_split_one_pol_and_larger_n1():
    # one-pol input (k_fs_b<N1, 1, 1>); 2^17 points (N1 = 512) over held
    # buffers
    res, cfg, _ = run_case(Nc=3, Ntpb=1 << 20, Nts=[8192], Nas=[64], Np=1, Npolout=[1], nbuf=2,
                           tone=(2, 20.0, 300.0, 1024.0))
    check_results(res, cfg, tol=PLAIN_TOL)
    r, cfg, _, _ = run_calls([1] * 16, Nc=2, Ntpb=1 << 20, Nts=[131072], Nas=[128],
                             Npolout=[-4], tone=TONE)
    check_results(r, cfg, tol=PLAIN_TOL)


def test_split_level1_24bit():
    kw = dict(Nc=2, Ntpb=1 << 20, Nts=[8192, 16384], Nas=[32, 256], Npolout=[-4, -4])
    res, cfg, _ = run_case(nbuf=4, tone=TONE, mfma=MFMA_FAST, **kw)
    w = check_results(res, cfg, tol=PLAIN_TOL,
                      tol_by_product=unsplit_products(kw["Nts"], kw["Nas"], mfma=MFMA_FAST))
    print(f"  level 1: worst {w}", flush=True)
    assert max(w.values()) > 0
