"""GPU parity: librawspec (gfx950 HIP kernels, through the C ABI) against the
FP64 oracle on identical seeded input.  Tolerance: the plain FP32 bar,
max relative error <= 1e-5 (metric defined in spectro_cases.py); dump
counts, callback order and channel order must match exactly.
"""
import numpy as np
import pytest

from tests.spectro_cases import (MFMA_FAST, MFMA_FLOOR_MULT, MFMA_SPLIT, PLAIN_FLOOR_MULT, PLAIN_TOL,
                                 check_results,
                                 run_case, unsplit_products)

pytestmark = pytest.mark.gpu


def _run(tol=PLAIN_TOL, **kw):
    print(f"case: {kw}", flush=True)
    res, cfg, ev = run_case(**kw)
    worst = check_results(res, cfg, tol=tol, ics=kw.get("ics", False))
    print(f"  dumps={len(res)} worst={ {k: f'{v:.2e}' for k, v in worst.items()} }", flush=True)
    return res, cfg, ev


def test_8ch_int8_nfft1024_plain():
    # 8 ch, int8 sigma 8 + tone (ch 3, bin 300/1024), 4 blocks of 131072,
    # Nfft 1024, Nint 16, pols 1, forward (the CLI direction)
    res, cfg, ev = _run(Nc=8, Ntpb=131072, Nts=[1024], Nas=[16], nbuf=4,
                        tone=(3, 20.0, 300.0, 1024.0))
    assert len(res) == 4 and cfg.Nds[0] == 8
    # the tone lands in fine channel (300 + 512) % 1024 = 812 of channel 3
    g = res[0][2].reshape(8, 8, 1024)
    assert int(np.argmax(g[0, 3])) == 812


@pytest.mark.parametrize("nt", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("mode", [1, 4, -4])
def test_small_lengths_all_modes(nt, mode):
    na = max(1, 2048 // nt)
    _run(Nc=3, Ntpb=8192, Nts=[nt], Nas=[na], Npolout=[mode], nbuf=2, conj=1,
         fft_dir=-1 if mode == 4 else 1, tone=(1, 15.0, 37.0, 128.0))


@pytest.mark.parametrize("nt", [8192, 16384, 32768, 65536, 131072, 262144, 1 << 20])
def test_fourstep_lengths(nt):
    _run(Nc=2, Ntpb=max(nt, 65536), Nts=[nt], Nas=[2], Npolout=[-4], nbuf=2,
         tone=(1, 20.0, 300.0, 1024.0))


@pytest.mark.parametrize("nt", [1 << 21, 1 << 22, 1 << 23, 1 << 24])
def test_fourstep_lengths_above_2_20(nt):
    # N1 x N2 = 2048 x 1024 ... 4096 x 4096: pass-A teams of 128 and 256
    # threads, pass-B tiles of 2 and 1 columns, one spectrum per buffer,
    # integrated over two buffers (held first-pass output)
    _run(Nc=2, Ntpb=nt, Nts=[nt], Nas=[2], Npolout=[-4], nbuf=2,
         tone=(1, 20.0, 300.0, 1024.0))


def test_three_products_pols4_inverse_conj():
    # rawspectest-like product set at reduced size: 2^17, 8, 1024
    _run(Nc=2, Ntpb=65536, Nts=[131072, 8, 1024], Nas=[3, 128, 256],
         Npolout=[4, 4, 4], nbuf=6, fft_dir=-1, conj=1)


def test_three_products_stokes_forward():
    _run(Nc=2, Ntpb=65536, Nts=[131072, 8, 1024], Nas=[3, 128, 256],
         Npolout=[-4, -4, -4], nbuf=6, fft_dir=1, conj=0, tone=(0, 20.0, 300.0, 1024.0))


def test_int16_ics_multi_antenna():
    w = [0.5 + 0.0625 * a for a in range(4)]
    _run(Nc=8, Nant=4, Ntpb=16384, Nbps=16, sigma=1000.0, Nts=[64, 4096, 16384, 8],
         Nas=[256, 4, 2, 2048], Npolout=[1, -4, 4, 1], nbuf=4, ics=True, Aws=w,
         conj=1)


def test_sixteen_antennas_int16_four_products_ics():
    """The batched multi-antenna product set at full block length: 16
    antennas x 4 coarse channels each (the full geometry has 16 each; the
    channels are independent), 2 pols of complex int16 ~ round(N(0, 1000^2)),
    blocks of 262144 samples, Nfft {64, 4096, 131072, 8} x Nint {2048, 32, 4,
    16384}, pols 1, incoherent sum with weights uniform in [0.5, 1.5]; two
    buffers so the 131072-point product (two buffers per dump) dumps too."""
    import random
    rng = random.Random(16)
    w = [rng.uniform(0.5, 1.5) for _ in range(16)]
    res, cfg, ev = _run(Nc=64, Nant=16, Ntpb=262144, Nbps=16, sigma=1000.0,
                        Nts=[64, 4096, 131072, 8], Nas=[2048, 32, 4, 16384],
                        Npolout=[1, 1, 1, 1], nbuf=2, ics=True, Aws=w)
    assert [cfg.Nds[i] for i in range(4)] == [2, 2, 1, 2]
    assert [cfg.Nis[i] for i in range(4)] == [1, 1, 2, 1]
    assert sorted((n, i) for (n, i, *_r) in res) == [(0, 0), (0, 1), (0, 3), (1, 0), (1, 1),
                                                      (1, 2), (1, 3)]


@pytest.mark.parametrize("mode,mfma", [(1, 0), (4, 0), (-4, 0), (-4, MFMA_SPLIT), (-4, MFMA_FAST),
                                       (1, MFMA_FAST), (4, MFMA_FAST)],
                         ids=["power", "crosspol", "stokes", "stokes-mfma-split", "stokes-mfma",
                              "power-mfma", "crosspol-mfma"])
def test_three_products_full_integration_lengths(mode, mfma):
    """The 64-channel three-product set at its full integration lengths on
    one of its channels (channels are independent; the tone's): blocks of
    524288 samples, two per buffer, Nfft {2^20, 8, 1024} x Nint {51, 128,
    3072} over 51 buffers, so the 2^20 product integrates across 51 buffers
    (first-pass output held across the one-buffer calls) and the 1024-point
    product over 3072 spectra across 3 buffers; pols 1 (total power), 4
    (XX, YY, Re XY*, Im XY*) and -4 (I, Q, U, V); on the matrix cores, 9
    buffers with Nint {9, 128, 3072}, split operands everywhere, or at the
    engine's 1e-3 bar (the 1024-point product on unsplit operands, computed
    with the 8-point product by one kernel, in each detection mode).  FP32
    sums in blocks (registers, a
    fixed-order tree across slots, one add per buffer or held group) keep
    the 3072-spectrum integration within the plain bar without compensated
    summation.  Beside the bar's metric, the per-bin |dP|/P of the power
    components (no channel-mean floor) is printed."""
    from tests.spectro_cases import bin_rel_err, nonneg_components
    nbuf = 9 if mfma else 51
    print(f"case: pols {mode} mfma {mfma} nbuf {nbuf}", flush=True)
    res, cfg, ev = run_case(Nc=1, Ntpb=524288, Nts=[1 << 20, 8, 1024], Nas=[nbuf, 128, 3072],
                            Npolout=[mode] * 3, nbuf=nbuf, tone=(0, 20.0, 300.0, 1024.0),
                            mfma=mfma)
    worst = check_results(res, cfg, tol=PLAIN_TOL, floor_mult=MFMA_FLOOR_MULT if mfma else PLAIN_FLOOR_MULT,
                          tol_by_product=unsplit_products(cfg.Nts[:3], cfg.Nas[:3], mfma=mfma))
    print(f"  dumps={len(res)} worst={ {k: f'{v:.2e}' for k, v in worst.items()} }", flush=True)
    perbin = {}
    for (n, i, g, r, *_rest) in res:
        q = abs(cfg.Npolout[i])
        g4 = g.reshape(cfg.Nds[i], q, cfg.Nc, cfg.Nts[i])
        r4 = r.reshape(cfg.Nds[i], q, cfg.Nc, cfg.Nts[i])
        for k in nonneg_components(mode):
            perbin[i] = max(perbin.get(i, 0.0), bin_rel_err(g4[:, k], r4[:, k]))
    print(f"  per-bin |dP|/P of the power components: "
          f"{ {k: f'{v:.2e}' for k, v in perbin.items()} }", flush=True)
    assert [sum(1 for r in res if r[1] == i) for i in range(3)] == [1, nbuf, nbuf // 3]


def test_single_pol():
    _run(Np=1, Nc=4, Ntpb=32768, Nts=[32768, 256], Nas=[2, 16], nbuf=4)


def test_complex4_expansion():
    _run(Nc=4, Ntpb=16384, Nts=[1024, 16], Nas=[8, 64], Npolout=[-4, 1], nbuf=2,
         expand4=True)
