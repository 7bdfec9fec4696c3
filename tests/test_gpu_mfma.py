"""GPU parity of the matrix-core engine (rawspec_amd_set_mfma).

Level 2 (MFMA_SPLIT): split-FP16 MFMA transforms with FP32 accumulation
everywhere, checked against the FP64 oracle at nearly the bar of the FP32
vector path: max relative error <= 1e-5, or, for the four-step lengths, whose
rounding spurs of a strong tone exceed 1e-5 in any FP32 FFT, MFMA_FLOOR_MULT
times the error of an FP32 reference FFT (metric in spectro_cases.py).

Level 1 (MFMA_FAST, the engine at its own 1e-3 bar): 1024-point int8
products integrating >= 32 spectra run with unsplit FP16 operands and are
held to 1e-3 (their measured errors are printed); shorter 1024-point
integrations run the vector kernel, and everything else keeps the tight
bounds.

The matrix-core kernels cover the 1024-point small products and the
four-step lengths whose pass B is 1024 points (2^19, 2^20), with two pols;
the other shapes run on the vector units with the engine selected, which
these tests also exercise.
"""
import numpy as np
import pytest

from tests.spectro_cases import (MFMA_BAR, MFMA_FAST, MFMA_FLOOR_MULT, MFMA_SPLIT, PLAIN_TOL,
                                 check_results, rel_err, run_case, unsplit_products)

pytestmark = pytest.mark.gpu


def _run(**kw):
    print(f"case: {kw}", flush=True)
    res, cfg, ev = run_case(mfma=MFMA_SPLIT, **kw)
    worst = check_results(res, cfg, tol=PLAIN_TOL, ics=kw.get("ics", False),
                          floor_mult=MFMA_FLOOR_MULT)
    print(f"  dumps={len(res)} worst={ {k: f'{v:.2e}' for k, v in worst.items()} }", flush=True)
    return res, cfg, ev


@pytest.mark.parametrize("nt", [1 << 19, 1 << 20])
@pytest.mark.parametrize("mode", [1, 4, -4])
def test_mfma_fourstep_integrating_over_buffers(nt, mode):
    # Nis = 2: first buffer writes the internal accumulator, second adds and
    # writes the host layout
    _run(Nc=2, Ntpb=nt, Nts=[nt], Nas=[2], Npolout=[mode], nbuf=2,
         tone=(1, 20.0, 300.0, 1024.0))


def test_mfma_fourstep_three_buffer_integration_inverse_conj():
    _run(Nc=2, Ntpb=1 << 20, Nts=[1 << 20], Nas=[3], Npolout=[-4], nbuf=3, fft_dir=-1, conj=1,
         tone=(0, 20.0, 300.0, 1024.0))


def test_mfma_fourstep_several_dumps_per_buffer():
    # two 2^19 spectra per buffer: Nint 1 -> two dumps per buffer (one
    # spectrum per group), Nint 2 -> one dump of both spectra
    _run(Nc=3, Ntpb=1 << 20, Nts=[1 << 19, 1 << 19], Nas=[1, 2], Npolout=[-4, 4], nbuf=2,
         tone=(2, 20.0, 300.0, 1024.0))


def test_mfma_int16_and_small_products():
    # int16 operands (scaled into FP16 range) and small products that stay
    # on the vector units with the engine selected
    _run(Nc=2, Ntpb=1 << 20, Nbps=16, sigma=1000.0, Nts=[1 << 20, 1024, 8], Nas=[2, 8, 128],
         Npolout=[-4, -4, 1], nbuf=2)


def test_mfma_engine_switch_waits_for_the_integration_boundary():
    # switched on during an integration (buffer 1 of 3): the product keeps
    # the vector engine until its next integration starts (buffer 3)
    res, cfg, ev = run_case(Nc=2, Ntpb=1 << 20, Nts=[1 << 20], Nas=[3], Npolout=[-4], nbuf=6,
                            tone=(1, 20.0, 300.0, 1024.0), mfma_switch=(1, MFMA_SPLIT))
    check_results(res, cfg, tol=PLAIN_TOL, floor_mult=MFMA_FLOOR_MULT)
    assert len(res) == 2


def test_mfma_and_vector_engines_agree():
    kw = dict(Nc=2, Ntpb=1 << 20, Nts=[1 << 20], Nas=[1], Npolout=[-4], nbuf=1,
              tone=(1, 20.0, 300.0, 1024.0))
    rv, cfg, _ = run_case(**kw)
    rm, _, _ = run_case(mfma=MFMA_SPLIT, **kw)
    g_v, g_m, ref = rv[0][2], rm[0][2], rv[0][3]
    d = rel_err(g_m, g_v, cfg.Nds[0], 4, cfg.Nc, cfg.Nts[0], -4).max()
    e_v = rel_err(g_v, ref, cfg.Nds[0], 4, cfg.Nc, cfg.Nts[0], -4).max()
    e_m = rel_err(g_m, ref, cfg.Nds[0], 4, cfg.Nc, cfg.Nts[0], -4).max()
    print(f"  vector {e_v:.2e}  matrix {e_m:.2e}  between {d:.2e}", flush=True)
    assert not np.array_equal(g_v, g_m)      # the matrix-core kernel really ran
    assert e_m <= max(PLAIN_TOL, MFMA_FLOOR_MULT * e_v)


@pytest.mark.parametrize("mode", [1, 4, -4])
@pytest.mark.parametrize("nint", [8, 64, 256])
def test_mfma_small_1024(mode, nint):
    # Nint 8: one wavefront per dump group, written directly; 64: groups split
    # over wavefronts (partial sums + reduction); 256: integration over 4
    # buffers of 64 spectra
    _run(Nc=3, Ntpb=65536, Nts=[1024], Nas=[nint], Npolout=[mode], nbuf=4, conj=1,
         fft_dir=-1 if mode == 4 else 1, tone=(1, 15.0, 37.0, 128.0))


def test_mfma_small_1024_int16_with_other_lengths():
    _run(Nc=4, Ntpb=16384, Nbps=16, sigma=1000.0, Nts=[1024, 4096, 8], Nas=[16, 4, 512],
         Npolout=[-4, -4, 1], nbuf=3)


def test_mfma_small_engine_switch_mid_integration():
    # the small kernels of both engines share the host layout: switching in
    # the middle of an integration is exact bookkeeping
    res, cfg, ev = run_case(Nc=2, Ntpb=32768, Nts=[1024], Nas=[96], Npolout=[-4], nbuf=6,
                            tone=(0, 20.0, 300.0, 1024.0), mfma_switch=(2, MFMA_SPLIT))
    check_results(res, cfg, tol=PLAIN_TOL)
    assert len(res) == 2


# -- level 1: unsplit FP16 operands where the integration allows ------------

def _kernels(level, **kw):
    """Names of the kernels one buffer of this configuration launches."""
    import torch
    import rawspec_amd as R
    sp = R.Spectrometer(Np=2, Nbps=8, **kw)
    assert sp.set_mfma(level) == 0 and sp.get_mfma() == level
    d = torch.zeros(sp.device_input_size(), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    sp.profile(True)
    assert sp.start_processing_device(d.data_ptr(), 1) == 0
    names = {k.split(".")[1] for k in sp.profile_report()}
    sp.close()
    return names


def test_unsplit_kernel_selection():
    kw = dict(Nc=2, Ntpb=65536, Nts=[1024], Npolout=[-4])
    print("case: kernel selection by level and integration length", flush=True)
    assert "k_small_mfma_u" in _kernels(MFMA_FAST, Nas=[64], **kw)
    # too short for unsplit operands: the vector kernel (faster than split)
    short = _kernels(MFMA_FAST, Nas=[16], **kw)
    assert "k_small" in short and not any(k.startswith("k_small_mfma") for k in short), short
    split = _kernels(MFMA_SPLIT, Nas=[64], **kw)
    assert "k_small_mfma" in split and "k_small_mfma_u" not in split, split
    assert "k_small" in _kernels(0, Nas=[64], **kw)


@pytest.mark.parametrize("mode", [1, 4, -4])
@pytest.mark.parametrize("nint", [32, 256, 3072])
def test_unsplit_small_1024(mode, nint):
    """1024-point products on unsplit FP16 operands against the FP64 oracle:
    within the engine's 1e-3 bar; forward and inverse, conjugated input,
    dump groups written directly, split over wavefronts and integrated over
    buffers (3072 spectra: three buffers of 1024)."""
    nbuf = 3 if nint == 3072 else 4
    kw = dict(Nc=3, Ntpb=1 << 20 if nint == 3072 else 65536, Nts=[1024], Nas=[nint],
              Npolout=[mode], nbuf=nbuf, conj=1, fft_dir=-1 if mode == 4 else 1,
              tone=(1, 15.0, 37.0, 128.0))
    print(f"case: unsplit {kw}", flush=True)
    res, cfg, ev = run_case(mfma=MFMA_FAST, **kw)
    worst = check_results(res, cfg, tol_by_product={0: MFMA_BAR})
    print(f"  dumps={len(res)} worst={worst[0]:.2e} (bar {MFMA_BAR:.0e})", flush=True)


def test_unsplit_with_other_products_and_switch():
    """Level 1 on a product set mixing unsplit (1024 points, 64 spectra),
    split (1024 points, 8 spectra: a second context) and vector (8 points)
    products, switched on in the middle of an integration."""
    kw = dict(Nc=2, Ntpb=65536, Nts=[1024, 8], Nas=[64, 128], Npolout=[-4, -4], nbuf=4,
              tone=(0, 20.0, 300.0, 1024.0))
    res, cfg, ev = run_case(mfma_switch=(1, MFMA_FAST), **kw)
    w = check_results(res, cfg, tol_by_product=unsplit_products(kw["Nts"], kw["Nas"]))
    print(f"  worst={ {k: f'{v:.2e}' for k, v in w.items()} }", flush=True)
    res, cfg, ev = run_case(mfma=MFMA_FAST, Nc=2, Ntpb=65536, Nts=[1024], Nas=[8], Npolout=[-4],
                            nbuf=2, tone=(0, 20.0, 300.0, 1024.0))
    check_results(res, cfg, tol=PLAIN_TOL)


@pytest.mark.parametrize("nt", [1 << 17, 1 << 20])
@pytest.mark.parametrize("mode", [1, 4, -4])
def test_level1_fourstep_24bit_intermediate(nt, mode):
    """Level 1 keeps four-step products' intermediate in 24-bit floats: one
    spectrum per dump (the worst case: nothing averages) and an integration
    over three buffers (held first-pass output), forward and inverse, with
    the tone; within the 1e-3 bar of the engine against the FP64 oracle, and
    within 1e-4 of the FP32 path on the same input (a numpy model of the
    rounding gives at most 2e-5)."""
    for nas, nbuf, fdir in [(1, 1, 1), (3, 3, -1)]:
        kw = dict(Nc=2, Ntpb=nt, Nts=[nt], Nas=[nas], Npolout=[mode], nbuf=nbuf, fft_dir=fdir,
                  conj=1, tone=(1, 20.0, 300.0, 1024.0))
        print(f"case: level 1 {kw}", flush=True)
        r1, cfg, _ = run_case(mfma=MFMA_FAST, **kw)
        w = check_results(r1, cfg, tol_by_product={0: MFMA_BAR})
        rv, _, _ = run_case(**kw)
        d = max(rel_err(a[2], b[2], cfg.Nds[0], abs(mode), cfg.Nc, nt, mode).max()
                for a, b in zip(r1, rv))
        print(f"  vs oracle {w[0]:.2e}, vs the FP32 path {d:.2e}", flush=True)
        assert d <= 1e-4
        assert not all(np.array_equal(a[2], b[2]) for a, b in zip(r1, rv))   # 24-bit really ran


@pytest.mark.parametrize("level", [MFMA_SPLIT, MFMA_FAST])
@pytest.mark.parametrize("nt", [1 << 20, 1 << 21])
def test_matrix_core_pass_a_selection(level, nt):
    """Level 2 runs pass A on the matrix cores for 2^20 points (N1 = N2 =
    1024) and on the vector units for 2^21 (N1 = 2048, whose pass B reads
    2-wide k2 tiles); level 1 keeps the faster vector pass A (with the
    24-bit intermediate).  Every case integrates over two buffers within its
    level's bar."""
    kw = dict(Nc=2, Ntpb=nt, Nts=[nt], Nas=[2], Npolout=[-4])
    names = _kernels(level, **kw)
    print(f"case: level {level} N {nt}: {sorted(names)}", flush=True)
    on_mfma = level == MFMA_SPLIT and nt == 1 << 20
    assert ("k_fs_a_mfma" in names) == on_mfma
    assert ("k_fs_a" in names) == (not on_mfma)
    res, cfg, _ = run_case(mfma=level, nbuf=2, tone=(1, 20.0, 300.0, 1024.0), **kw)
    w = check_results(res, cfg, tol=PLAIN_TOL, floor_mult=MFMA_FLOOR_MULT,
                      tol_by_product=unsplit_products(kw["Nts"], kw["Nas"], mfma=level))
    print(f"  worst {w[0]:.2e}", flush=True)
