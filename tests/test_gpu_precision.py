"""Per-product error of the transform engines against the FP64 oracle on
the full-Stokes product set {2^20, 8, 1024} (the 64-channel workload's
products, its int8 sigma-8 stream, blocks of 524288 samples): the FP32
vector path, the matrix-core path with split FP16 operands everywhere
(level 2) and the matrix-core engine at its 1e-3 bar (level 1: unsplit
operands for the 1024-point product, which integrates 3072 spectra) on the
same seeded input, max and median relative error per product (metric in
spectro_cases.py), with the error of a single-precision FFT (scipy.fft,
complex64) beside the 2^20 product.  Two channels: one carries
the tone, the other is noise only.  Three buffers: Nint 3 for the 2^20
product (held first-pass output across the one-buffer calls), 128 for the
8-point product (1024 dumps per buffer), 3072 for the 1024-point product
(one dump over all three buffers).  The 8-point product has no matrix-core
kernel and runs on the vector units under both engines.
"""
import numpy as np
import pytest

from tests.spectro_cases import (MFMA_BAR, MFMA_FAST, MFMA_FLOOR_MULT, MFMA_SPLIT, PLAIN_TOL,
                                 check_results, rel_err, run_case, unsplit_products)

pytestmark = pytest.mark.gpu


def _errors(res, cfg):
    """{product: (all relative errors, FP32 reference errors or None)}."""
    out = {}
    for (_n, i, g, r, _gi, _ri, f32) in res:
        q = abs(cfg.Npolout[i])
        e = rel_err(g, r, cfg.Nds[i], q, cfg.Nc, cfg.Nts[i], cfg.Npolout[i]).ravel()
        ef = (rel_err(f32, r, cfg.Nds[i], q, cfg.Nc, cfg.Nts[i], cfg.Npolout[i]).ravel()
              if f32 is not None else None)
        prev = out.get(i, (np.empty(0), None))
        out[i] = (np.concatenate([prev[0], e]),
                  ef if prev[1] is None else np.concatenate([prev[1], ef]))
    return out


def test_per_product_max_and_median_error_both_engines():
    kw = dict(Nc=2, Ntpb=524288, Nts=[1 << 20, 8, 1024], Nas=[3, 128, 3072],
              Npolout=[-4, -4, -4], nbuf=3, tone=(0, 20.0, 300.0, 1024.0))
    print(f"case: {kw}", flush=True)
    rv, cfg, _ = run_case(**kw)
    check_results(rv, cfg, tol=PLAIN_TOL)
    rm, _, _ = run_case(mfma=MFMA_SPLIT, **kw)
    check_results(rm, cfg, tol=PLAIN_TOL, floor_mult=MFMA_FLOOR_MULT)
    ru, _, _ = run_case(mfma=MFMA_FAST, **kw)
    check_results(ru, cfg, tol=PLAIN_TOL, floor_mult=MFMA_FLOOR_MULT,
                  tol_by_product=unsplit_products(kw["Nts"], kw["Nas"]))
    ev, em, eu = _errors(rv, cfg), _errors(rm, cfg), _errors(ru, cfg)
    assert sorted(ev) == sorted(em) == sorted(eu) == [0, 1, 2]
    print("  product   vector max / median      split max / median       "
          "level-1 max / median     FP32 FFT max / median", flush=True)
    for i in range(3):
        v, f = ev[i]
        m, _ = em[i]
        u, _ = eu[i]
        fs = f"{f.max():.2e} / {np.median(f):.2e}" if f is not None else "-"
        print(f"  N={cfg.Nts[i]:<8d} {v.max():.2e} / {np.median(v):.2e}   "
              f"{m.max():.2e} / {np.median(m):.2e}   {u.max():.2e} / {np.median(u):.2e}   {fs}",
              flush=True)
        assert v.size == m.size == u.size > 0
        # typical bins sit at FP32 rounding level on the vector and split paths
        assert np.median(v) <= PLAIN_TOL / 10 and np.median(m) <= PLAIN_TOL / 10
        assert m.max() <= MFMA_BAR and u.max() <= MFMA_BAR
    # the matrix-core kernels really ran for the 2^20 and 1024-point products
    for (_, i, g, *_r), (_, j, h, *_s) in zip(rv, rm):
        assert i == j
        if cfg.Nts[i] != 8:
            assert not np.array_equal(g, h), f"product {i}"
