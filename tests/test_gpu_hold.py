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

import oracle as O
from tests.spectro_cases import (FP32_FLOOR_NT, MFMA_FLOOR_MULT, PLAIN_TOL, Fp32Ref,
                                 check_results, gen_blocks)

pytestmark = pytest.mark.gpu


def run_pattern(pattern, *, Nc, Ntpb, Nts, Nas, Npolout, held=None, mfma=False, tone=None,
                fft_dir=1, conj=0, seed=1):
    """pattern: a list of calls, each k > 0 (k buffers: 1 through
    rawspec_amd_start_processing_device, more through the batch call) or
    "reset" (rawspec_reset_integration).  Returns (results as in
    spectro_cases.run_case, cfg, the GPU's dumps per product in order)."""
    import torch
    import rawspec_amd as R

    print(f"case: {pattern} held={held} mfma={mfma} Nts={Nts} Nas={Nas}", flush=True)
    No = len(Nts)
    got = {i: [] for i in range(No)}

    def on_dump(sp, product, cbtype):
        if cbtype == R.RAWSPEC_CALLBACK_POST_DUMP:
            got[product].append(sp.pwrbuf(product).copy())

    sp = R.Spectrometer(Nc=Nc, Ntpb=Ntpb, Nts=Nts, Nas=Nas, Np=2, Nbps=8, Npolout=Npolout,
                        input_conjugated=conj, callback=on_dump)
    assert sp.get_held_buffers() == 8
    if held is not None:
        assert sp.set_held_buffers(held) == 0 and sp.get_held_buffers() == held
    if mfma:
        assert sp.set_mfma(True) == 0
    cfg = O.make_cfg(No=No, Np=2, Nant=1, Nc=Nc, Ntpb=Ntpb, Nbps=8, Npolout=Npolout, Nts=Nts,
                     Nas=Nas, input_conjugated=conj)
    assert O.derive(cfg) == 0
    nbuf = sum(k for k in pattern if k != "reset")
    Nb_ = sp.ctx.Nb
    host = [O.device_layout(gen_blocks(Nc, Ntpb, 2, 8, Nb_, seed=seed, tone=tone,
                                       first_blk=n * Nb_), Nc, Ntpb, 2) for n in range(nbuf)]
    dev = [torch.from_numpy(h.view(np.uint8).copy()).to("cuda") for h in host]
    torch.cuda.synchronize()
    st = O.State(cfg)

    def f32_refs():
        return {i: Fp32Ref(cfg, i) for i in range(No) if Nts[i] >= FP32_FLOOR_NT}

    f32 = f32_refs()
    refs = {i: [] for i in range(No)}
    n = 0
    for k in pattern:
        if k == "reset":
            assert sp.reset_integration() == 0
            st.reset()
            f32 = f32_refs()
            continue
        if k == 1:
            assert sp.start_processing_device(dev[n].data_ptr(), fft_dir) == 0
        else:
            assert sp.start_processing_device_batch([d.data_ptr() for d in dev[n:n + k]],
                                                    fft_dir) == 0
        for h in host[n:n + k]:
            _, ref, _ = st.process(h, fft_dir)
            f32out = {i: r.process(h, fft_dir) for i, r in f32.items()}
            for i in ref:
                refs[i].append((ref[i], f32out.get(i)))
        n += k
    assert sp.wait_for_completion() == 0
    sp.close()
    st.close()
    results = []
    for i in range(No):
        assert len(got[i]) == len(refs[i]), (i, len(got[i]), len(refs[i]))
        for g, (r, f) in zip(got[i], refs[i]):
            results.append((0, i, g, r, None, None, f))
    del dev
    return results, cfg, got


CASE = dict(Nc=2, Ntpb=65536, Nts=[131072, 1024], Nas=[11, 256], Npolout=[-4, -4],
            tone=(1, 20.0, 300.0, 1024.0))


def test_held_one_buffer_calls_match_oracle_and_the_batch_call_bit_for_bit():
    # Nis = 11: held groups of 8 + 3 per integration, two integrations; the
    # batch call over each integration runs the same groups
    r1, cfg, g1 = run_pattern([1] * 22, **CASE)
    worst = check_results(r1, cfg, tol=PLAIN_TOL)
    print(f"  worst={worst}", flush=True)
    _, _, gb = run_pattern([11, 11], **CASE)
    for i in range(2):
        assert len(g1[i]) == len(gb[i])
        for a, b in zip(g1[i], gb[i]):
            assert np.array_equal(a, b), f"product {i}"


@pytest.mark.parametrize("held", [0, 2, 3, 16])
def test_held_group_sizes(held):
    # 0: off (pass B per buffer); 3: groups 3+3+3+2; 16: capped at Nis = 11
    r, cfg, _ = run_pattern([1] * 11, held=held, **CASE)
    check_results(r, cfg, tol=PLAIN_TOL)


def test_held_mixed_with_batches_and_resets():
    # one-buffer calls around batch calls inside integrations (the batch call
    # runs the held buffers' second pass first), and resets that drop held
    # buffers mid-integration
    pattern = [1, 1, 1, 4, 1, 1, 1, 1, 1, "reset", 1, 1, 1, "reset", 2, 1, 1, 1, 1, 1, 1, 1, 1,
               3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1]
    r, cfg, got = run_pattern(pattern, **CASE)
    check_results(r, cfg, tol=PLAIN_TOL)
    assert len(got[0]) == 3


def test_held_inverse_conj_full_pols():
    r, cfg, _ = run_pattern([1] * 8, Nc=3, Ntpb=65536, Nts=[65536, 8], Nas=[4, 128],
                            Npolout=[4, 4], fft_dir=-1, conj=1, tone=(2, 20.0, 300.0, 1024.0))
    check_results(r, cfg, tol=PLAIN_TOL)


def test_held_several_spectra_per_buffer():
    # 2^19 points in 2^20-sample buffers: two spectra per buffer, Nis = 3
    r, cfg, _ = run_pattern([1] * 6, Nc=2, Ntpb=1 << 20, Nts=[1 << 19], Nas=[6], Npolout=[1],
                            tone=(0, 20.0, 300.0, 1024.0))
    check_results(r, cfg, tol=PLAIN_TOL)


def test_held_matrix_core_pass_b():
    kw = dict(Nc=2, Ntpb=1 << 20, Nts=[1 << 20], Nas=[5], Npolout=[-4], mfma=True,
              tone=(1, 20.0, 300.0, 1024.0))
    r1, cfg, g1 = run_pattern([1] * 10, **kw)
    check_results(r1, cfg, tol=PLAIN_TOL, floor_mult=MFMA_FLOOR_MULT)
    _, _, gb = run_pattern([5, 5], **kw)
    for a, b in zip(g1[0], gb[0]):
        assert np.array_equal(a, b)
