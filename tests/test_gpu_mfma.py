"""GPU parity of the matrix-core engine (rawspec_amd_set_mfma): split-FP16
MFMA transforms with FP32 accumulation, checked against the FP64 oracle at
nearly the bar of the FP32 vector path: max relative error <= 1e-5, or, for
the four-step lengths, whose rounding spurs of a strong tone exceed 1e-5 in
any FP32 FFT, MFMA_FLOOR_MULT times the error of an FP32 reference FFT
(metric in spectro_cases.py).  The engine's own bar is 1e-3.  The
matrix-core kernels cover the four-step lengths whose pass B is 1024 points
(2^19, 2^20) with two pols; the other shapes run on the vector units with
the engine selected, which these tests also exercise.
"""
import numpy as np
import pytest

from tests.spectro_cases import MFMA_FLOOR_MULT, PLAIN_TOL, check_results, rel_err, run_case

pytestmark = pytest.mark.gpu


def _run(**kw):
    print(f"case: {kw}", flush=True)
    res, cfg, ev = run_case(mfma=True, **kw)
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
                            tone=(1, 20.0, 300.0, 1024.0), mfma_switch=(1, True))
    check_results(res, cfg, tol=PLAIN_TOL, floor_mult=MFMA_FLOOR_MULT)
    assert len(res) == 2


def test_mfma_and_vector_engines_agree():
    kw = dict(Nc=2, Ntpb=1 << 20, Nts=[1 << 20], Nas=[1], Npolout=[-4], nbuf=1,
              tone=(1, 20.0, 300.0, 1024.0))
    rv, cfg, _ = run_case(**kw)
    rm, _, _ = run_case(mfma=True, **kw)
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
                            tone=(0, 20.0, 300.0, 1024.0), mfma_switch=(2, True))
    check_results(res, cfg, tol=PLAIN_TOL)
    assert len(res) == 2
