"""GPU parity of the fused pair at engine level 1.

An 8-point product whose dumps integrate 128 spectra (one dump per 1024
samples) and a 1024-point int8 product that the matrix-core engine runs on
unsplit FP16 operands are computed by one kernel from one read of the
samples (k_small_mfma_u with its 8-point section).  The 8-point product
keeps the FP32 vector engine's precision and is held to the plain 1e-5 bar;
the 1024-point product keeps its level-1 bar (1e-3).  Both run through the
reference API (host blocks, one buffer per call) and through the batched
device entry point, whose dumps then follow each buffer in product order.
"""
import pytest

from tests.spectro_cases import (MFMA_FAST, PLAIN_TOL, check_results, run_case,
                                 run_device_case, unsplit_products)

pytestmark = pytest.mark.gpu

TONE = (1, 20.0, 300.0, 1024.0)


def _kernels(level, **kw):
    """Full names (product.kernel.N) of the kernels one buffer launches."""
    import torch
    import rawspec_amd as R
    sp = R.Spectrometer(Np=2, Nbps=8, **kw)
    assert sp.set_mfma(level) == 0
    d = torch.zeros(sp.device_input_size(), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    sp.profile(True)
    assert sp.start_processing_device(d.data_ptr(), 1) == 0
    names = set(sp.profile_report())
    sp.close()
    return names


def _check(res, cfg, kw):
    worst = check_results(res, cfg, tol=PLAIN_TOL, ics=kw.get("ics", False),
                          tol_by_product=unsplit_products(kw["Nts"], kw["Nas"], mfma=MFMA_FAST))
    print(f"  dumps={len(res)} worst={ {k: f'{v:.2e}' for k, v in worst.items()} }", flush=True)
    assert {i for (_, i, *_r) in res} == set(range(len(kw["Nts"]))), "a product never dumped"


def test_fused_pair_kernel_selection():
    print("case: fused pair selection", flush=True)
    kw = dict(Nc=2, Ntpb=65536, Nts=[8, 1024], Npolout=[-4, -4])
    fused = _kernels(MFMA_FAST, Nas=[128, 64], **kw)
    assert "p1.k_small_mfma_u8.N1024" in fused, fused
    assert not any(k.startswith("p0.") for k in fused), fused
    # the vector engine, and shapes outside the pair: separate kernels
    assert {"p0.k_small.N8", "p1.k_small.N1024"} <= _kernels(0, Nas=[128, 64], **kw)
    for nas, pols in (([64, 64], [-4, -4]),     # 8-point dumps of 512 samples
                      ([128, 16], [-4, -4]),    # 1024-point integration too short: split
                      ([128, 64], [1, -4])):    # different detection
        names = _kernels(MFMA_FAST, Nas=nas, **dict(kw, Npolout=pols))
        assert "p0.k_small.N8" in names and not any("k_small_mfma_u8" in k for k in names), names


@pytest.mark.parametrize("order", ["8-first", "1024-first"])
@pytest.mark.parametrize("mode", [1, 4, -4])
def test_fused_pair_parity(mode, order):
    # the 1024-point product integrates over two buffers (its first-buffer
    # flag alternates); a 32-point product runs beside the pair
    nts, nas = ([8, 1024, 32], [128, 128, 16]) if order == "8-first" else ([1024, 32, 8], [128, 16, 128])
    kw = dict(Nc=3, Ntpb=65536, Nts=nts, Nas=nas, Npolout=[mode] * 3, nbuf=4, tone=TONE)
    print(f"case: fused pair pols {mode} {order}", flush=True)
    res, cfg, _ = run_case(mfma=MFMA_FAST, **kw)
    _check(res, cfg, kw)


@pytest.mark.parametrize("mode", [4, -4])
def test_fused_pair_inverse_conjugated_incoherent_sum(mode):
    kw = dict(Nc=4, Ntpb=65536, Nts=[8, 1024], Nas=[128, 64], Npolout=[mode] * 2, nbuf=2,
              fft_dir=-1, conj=1, Nant=2, ics=True, Aws=[0.5, 1.5], tone=TONE)
    print(f"case: fused pair inverse conj ICS pols {mode}", flush=True)
    res, cfg, _ = run_case(mfma=MFMA_FAST, **kw)
    _check(res, cfg, kw)


def test_fused_pair_batches_and_dump_order():
    # batches of 3, 1 and 4 buffers: the 1024-point product's integration
    # (three buffers) runs across batch boundaries; each product's dumps come
    # in order and the pair's dumps follow each buffer in product order
    kw = dict(Nc=2, Ntpb=65536, Nts=[1024, 8, 65536], Nas=[192, 128, 2],
              Npolout=[-4, -4, -4], nbuf=8, tone=TONE)
    print("case: fused pair, batched calls", flush=True)
    ev = []
    res, cfg = run_device_case(batches=[3, 1, 4], mfma=MFMA_FAST, events=ev, **kw)
    _check(res, cfg, kw)
    import rawspec_amd as R
    posts = [p for (p, t) in ev if t == R.RAWSPEC_CALLBACK_POST_DUMP]
    assert posts.count(1) == 8 and posts.count(0) == 2 and posts.count(2) == 4, posts
    # within the pair, product 0's dumps (after its third and sixth buffers)
    # precede product 1's dump of the same buffer
    pair = [p for p in posts if p in (0, 1)]
    assert pair == [1, 1, 0, 1, 1, 1, 0, 1, 1, 1], pair
