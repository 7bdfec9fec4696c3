"""GPU parity of batched processing (rawspec_amd_start_processing_device_batch):
the same dumps as buffer-by-buffer processing, within the plain FP32 bar
against the FP64 oracle, for batches that start, continue and complete
integrations, on both transform engines."""
import numpy as np
import pytest

from tests.spectro_cases import (MFMA_FAST, MFMA_FLOOR_MULT, MFMA_SPLIT, PLAIN_FLOOR_MULT, PLAIN_TOL,
                                 check_results,
                                 rel_err, run_device_case, unsplit_products)

pytestmark = pytest.mark.gpu


def _run(**kw):
    print(f"case: {kw}", flush=True)
    res, cfg = run_device_case(**kw)
    worst = check_results(res, cfg, tol=PLAIN_TOL, ics=kw.get("ics", False),
                          floor_mult=MFMA_FLOOR_MULT if kw.get("mfma") else PLAIN_FLOOR_MULT,
                          tol_by_product=unsplit_products(kw["Nts"], kw["Nas"], kw.get("Np", 2),
                                                          kw.get("Nbps", 8), kw.get("mfma", 0)))
    print(f"  dumps={len(res)} worst={ {k: f'{v:.2e}' for k, v in worst.items()} }", flush=True)
    return res, cfg


@pytest.mark.parametrize("mfma", [0, MFMA_SPLIT, MFMA_FAST])
def test_batches_across_integration_boundaries(mfma):
    # Nis = 3 for the 2^20 product: batches [2, 3, 1, 3] start, continue
    # and complete integrations inside and at the edges of each batch; the
    # small products dump every buffer (8) or every other buffer (1024)
    _run(Nc=2, Ntpb=1 << 19, Nts=[1 << 20, 8, 1024], Nas=[3, 128, 2048], Npolout=[-4, -4, -4],
         nbuf=9, batches=[2, 3, 1, 3], tone=(1, 20.0, 300.0, 1024.0), mfma=mfma)


@pytest.mark.parametrize("mfma", [0, MFMA_SPLIT])
def test_batch_covering_a_whole_integration(mfma):
    # one batch holding two whole integrations of 4 buffers (2 spectra each):
    # each sum goes straight to its dump
    _run(Nc=3, Ntpb=1 << 20, Nts=[1 << 19], Nas=[8], Npolout=[4], nbuf=8, batches=[8],
         fft_dir=-1, conj=1, tone=(0, 20.0, 300.0, 1024.0), mfma=mfma)


def test_long_batches_int16_ics():
    # batches longer than the per-launch maximum (8) and int16 with ICS
    w = [0.75, 1.25]
    _run(Nc=4, Nant=2, Ntpb=1 << 17, Nbps=16, sigma=1000.0, Nts=[1 << 17, 64], Nas=[12, 2048],
         Npolout=[1, -4], nbuf=12, batches=[11, 1], ics=True, Aws=w)


def test_batched_and_single_buffer_results_agree():
    kw = dict(Nc=2, Ntpb=1 << 19, Nts=[1 << 20], Nas=[4], Npolout=[-4], nbuf=8,
              tone=(1, 20.0, 300.0, 1024.0))
    r1, cfg = run_device_case(**kw)
    r2, _ = run_device_case(batches=[4, 4], **kw)
    assert [(n, i) for (n, i, *_x) in r1] == [(n, i) for (n, i, *_x) in r2]
    for a, b in zip(r1, r2):
        d = rel_err(b[2], a[2], cfg.Nds[0], 4, cfg.Nc, cfg.Nts[0], -4).max()
        assert d <= 2e-6, d
