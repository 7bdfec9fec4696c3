"""GPU parity of transform lengths that are not powers of two (the reference
accepts any length cuFFT plans, as long as every product's length divides
the longest, rawspec_gpu.cu:519-539).  Up to 2048 points the small kernel
computes them by Bluestein's identity at a padded power-of-two length
(chirp multiply, FFT, multiply by the chirp kernel's FFT, inverse FFT,
chirp multiply), with FP64-generated tables; the FFT shift of an odd length
puts bin k at (k + N/2) mod N as the reference's two-copy shift does
(rawspec_gpu.cu:1659-1691).  Checked against the FP64 oracle (whose own FFT
handles any length) at the FP32 bar: odd and even lengths, lengths beside
power-of-two products of the same context, every polarisation mode, one-pol
and int16 input, integration within and over buffers, inverse transforms
with conjugated input.  Longer lengths that are not powers of two are
refused at initialisation.
"""
import pytest

from tests.spectro_cases import PLAIN_TOL, check_results, run_case, run_calls

pytestmark = pytest.mark.gpu


def _kernels(**kw):
    import torch
    import rawspec_amd as R
    sp = R.Spectrometer(Nbps=8, **kw)
    d = torch.zeros(sp.device_input_size(), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    sp.profile(True)
    assert sp.start_processing_device(d.data_ptr(), 1) == 0
    names = set(sp.profile_report())
    sp.close()
    return names


def test_lengths_run_the_small_kernel():
    names = _kernels(Nc=2, Ntpb=96000, Nts=[1000, 100, 10], Nas=[8, 16, 32], Np=2,
                     Npolout=[-4, -4, -4])
    print("case: kernels", sorted(names), flush=True)
    assert {"p0.k_small.N1000", "p1.k_small.N100", "p2.k_small.N10"} <= names


@pytest.mark.parametrize("pol", [1, 4, -4])
def test_decimal_lengths(pol):
    # 1000 (padded to 2048), 100 (256), 10 (32): dumps of 8, 16, 32 spectra
    # within 96000-sample buffers, and over two buffers for the longest
    res, cfg, _ = run_case(Nc=3, Ntpb=96000, Nts=[1000, 100, 10], Nas=[192, 16, 32],
                           Npolout=[pol] * 3, nbuf=4, tone=(1, 20.0, 300.0, 1000.0))
    w = check_results(res, cfg, tol=PLAIN_TOL)
    print(f"  pol {pol}: worst {w}", flush=True)


def test_odd_and_mixed_lengths():
    # Ntmax 1536 (padded to 4096): 3 * 512 beside the powers of two 512 and
    # 8 of the same context, and 3 (padded to 8, one thread per transform)
    res, cfg, _ = run_case(Nc=2, Ntpb=1536 * 32, Nts=[1536, 512, 3, 8], Nas=[16, 8, 64, 128],
                           Npolout=[-4, 4, 1, -4], nbuf=3, tone=(0, 20.0, 77.0, 1536.0))
    w = check_results(res, cfg, tol=PLAIN_TOL)
    print(f"  mixed: worst {w}", flush=True)
    res, cfg, _ = run_case(Nc=2, Ntpb=2047 * 16, Nts=[2047, 89], Nas=[4, 23], Npolout=[-4, 1],
                           nbuf=2, tone=(1, 20.0, 5.0, 89.0))
    w = check_results(res, cfg, tol=PLAIN_TOL)
    print(f"  odd: worst {w}", flush=True)


def test_one_pol_int16_inverse_conj():
    res, cfg, _ = run_case(Nc=2, Ntpb=1200 * 16, Nts=[1200, 600], Nas=[4, 32], Np=1,
                           Npolout=[1, 1], nbuf=2, fft_dir=-1, tone=(1, 20.0, 30.0, 600.0))
    check_results(res, cfg, tol=PLAIN_TOL)
    res, cfg, _ = run_case(Nc=2, Ntpb=1200 * 16, Nts=[1200, 600], Nas=[4, 32], Nbps=16,
                           Npolout=[4, -4], nbuf=2, fft_dir=-1, conj=1,
                           tone=(0, 20.0, 30.0, 600.0))
    check_results(res, cfg, tol=PLAIN_TOL)


def test_integration_over_buffers_one_call_and_batches():
    r, cfg, _, _ = run_calls([1, 1, 3, 1, 2], Nc=2, Ntpb=1000 * 10, Nts=[1000, 50], Nas=[40, 200],
                             Npolout=[-4, 4], tone=(1, 20.0, 300.0, 1000.0))
    w = check_results(r, cfg, tol=PLAIN_TOL)
    print(f"  over buffers: worst {w}", flush=True)


def test_longer_lengths_refused():
    import rawspec_amd as R
    with pytest.raises(RuntimeError):
        R.Spectrometer(Nc=1, Ntpb=3000 * 4, Nts=[3000], Nas=[1], Np=2, Nbps=8, Npolout=[1])


def test_incoherent_sum_and_complex4():
    # 4 antennas x 2 channels, weighted incoherent sum of 768- and 24-point
    # products; then 4-bit complex input expanded on the device
    w = [0.5 + 0.25 * a for a in range(4)]
    res, cfg, _ = run_case(Nc=8, Nant=4, Ntpb=768 * 16, Nts=[768, 24], Nas=[8, 64],
                           Npolout=[1, -4], nbuf=2, ics=True, Aws=w,
                           tone=(2, 20.0, 50.0, 768.0))
    check_results(res, cfg, tol=PLAIN_TOL, ics=True)
    res, cfg, _ = run_case(Nc=4, Ntpb=1000 * 16, Nts=[1000, 40], Nas=[8, 100],
                           Npolout=[-4, 1], nbuf=2, expand4=True)
    check_results(res, cfg, tol=PLAIN_TOL)
