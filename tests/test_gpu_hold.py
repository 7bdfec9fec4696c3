"""GPU parity of held first-pass output (rawspec_amd_set_held_buffers): a
four-step product whose integration spans several buffers, called one buffer
at a time, keeps each buffer's first-pass output and runs the second pass once
per group of held buffers.  Checked against the FP64 oracle at the FP32 bar of
spectro_cases.py, bit for bit against the batch call over the same buffers
(the same groups through the same kernels), and across integration resets
and mixed one-buffer / batch calls.
"""
import numpy as np
import pytest

from tests.spectro_cases import MFMA_SPLIT, MFMA_FLOOR_MULT, PLAIN_TOL, check_results
from tests.spectro_cases import run_calls as run_pattern

pytestmark = pytest.mark.gpu


CASE = dict(Nc=2, Ntpb=65536, Nts=[131072, 1024], Nas=[11, 256], Npolout=[-4, -4],
            tone=(1, 20.0, 300.0, 1024.0))


def test_held_one_buffer_calls_match_oracle_and_the_batch_call_bit_for_bit():
    # Nis = 11: held groups of 8 + 3 per integration, two integrations; the
    # batch call over each integration runs the same groups
    r1, cfg, g1, _ = run_pattern([1] * 22, **CASE)
    worst = check_results(r1, cfg, tol=PLAIN_TOL)
    print(f"  worst={worst}", flush=True)
    _, _, gb, _ = run_pattern([11, 11], **CASE)
    for i in range(2):
        assert len(g1[i]) == len(gb[i])
        for a, b in zip(g1[i], gb[i]):
            assert np.array_equal(a, b), f"product {i}"


@pytest.mark.parametrize("held", [0, 2, 3, 16])
def test_held_group_sizes(held):
    # 0: off (pass B per buffer); 3: groups 3+3+3+2; 16: capped at Nis = 11
    r, cfg, _, _ = run_pattern([1] * 11, held=held, **CASE)
    check_results(r, cfg, tol=PLAIN_TOL)


def test_held_mixed_with_batches_and_resets():
    # one-buffer calls around batch calls inside integrations (the batch call
    # runs the held buffers' second pass first), and resets that drop held
    # buffers mid-integration
    pattern = [1, 1, 1, 4, 1, 1, 1, 1, 1, "reset", 1, 1, 1, "reset", 2, 1, 1, 1, 1, 1, 1, 1, 1,
               3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1]
    r, cfg, got, _ = run_pattern(pattern, **CASE)
    check_results(r, cfg, tol=PLAIN_TOL)
    assert len(got[0]) == 3


def test_held_inverse_conj_full_pols():
    r, cfg, _, _ = run_pattern([1] * 8, Nc=3, Ntpb=65536, Nts=[65536, 8], Nas=[4, 128],
                            Npolout=[4, 4], fft_dir=-1, conj=1, tone=(2, 20.0, 300.0, 1024.0))
    check_results(r, cfg, tol=PLAIN_TOL)


def test_held_several_spectra_per_buffer():
    # 2^19 points in 2^20-sample buffers: two spectra per buffer, Nis = 3
    r, cfg, _, _ = run_pattern([1] * 6, Nc=2, Ntpb=1 << 20, Nts=[1 << 19], Nas=[6], Npolout=[1],
                            tone=(0, 20.0, 300.0, 1024.0))
    check_results(r, cfg, tol=PLAIN_TOL)


def test_held_matrix_core_pass_b():
    kw = dict(Nc=2, Ntpb=1 << 20, Nts=[1 << 20], Nas=[5], Npolout=[-4], mfma=MFMA_SPLIT,
              tone=(1, 20.0, 300.0, 1024.0))
    r1, cfg, g1, _ = run_pattern([1] * 10, **kw)
    check_results(r1, cfg, tol=PLAIN_TOL, floor_mult=MFMA_FLOOR_MULT)
    _, _, gb, _ = run_pattern([5, 5], **kw)
    for a, b in zip(g1[0], gb[0]):
        assert np.array_equal(a, b)
