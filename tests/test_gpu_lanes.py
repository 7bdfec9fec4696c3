"""Two-lane execution (four-step products on one stream, small products on
another, dumps on the context's stream): the same dumps, bit for bit, and
the same callback order as with everything on one stream, single buffers
and batches, both engines; and within the oracle's bound."""
from __future__ import annotations

import numpy as np
import pytest

from tests.spectro_cases import MFMA_SPLIT, check_results, run_device_case

pytestmark = pytest.mark.gpu

CASES = {
    # four-step 2^15 integrating over 2 buffers, small 1024 (3 per buffer)
    # and 8 (over 4 buffers): dumps at different buffers per product
    "mixed": dict(Nc=4, Ntpb=32768, Nts=[32768, 1024, 8], Nas=[2, 8, 16384],
                  Npolout=[-4, 4, 1], nbuf=8),
    # int16, ICS over 2 antennas, a four-step product dumping every buffer
    "ics16": dict(Nc=4, Ntpb=16384, Nbps=16, sigma=1000.0, Nts=[16384, 64], Nas=[1, 64],
                  Npolout=[1, -4], Nant=2, ics=True, Aws=[0.75, 1.25], nbuf=4),
}


def _run(case, lanes, batches=None, mfma=False):
    ev = []
    kw = dict(CASES[case])
    res, cfg = run_device_case(**kw, batches=batches, lanes=lanes, events=ev, mfma=mfma,
                               tone=(1, 20.0, 300.0, 1024.0))
    return res, cfg, ev


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("batches", [None, [3, 5], [8]])
def test_lanes_match_one_stream(case, batches):
    nbuf = CASES[case]["nbuf"]
    if batches is not None and sum(batches) != nbuf:
        batches = [nbuf]
    r1, cfg, e1 = _run(case, True, batches)
    r0, _, e0 = _run(case, False, batches)
    assert e1 == e0 and len(e1) > 0
    assert len(r1) == len(r0)
    for a, b in zip(r1, r0):
        assert (a[0], a[1]) == (b[0], b[1])
        assert np.array_equal(a[2], b[2])
        if a[4] is not None:
            assert np.array_equal(a[4], b[4])
    check_results(r1, cfg, ics=CASES[case].get("ics", False))


def test_lanes_matrix_engine():
    r1, cfg, e1 = _run("mixed", True, [4, 4], mfma=MFMA_SPLIT)
    r0, _, e0 = _run("mixed", False, [4, 4], mfma=MFMA_SPLIT)
    assert e1 == e0
    for a, b in zip(r1, r0):
        assert np.array_equal(a[2], b[2])


def _device_outputs(sp, n):
    """Each product's device-resident output, copied to the host."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    out = []
    for i in range(n):
        a = np.empty(sp.ctx.h_pwrbuf_size[i] // 4, dtype=np.float32)
        assert hip.hipMemcpy(a.ctypes.data, sp.device_pwrbuf(i), a.nbytes, 2) == 0
        out.append(a)
    return out


def test_device_output_without_callbacks():
    """Device-resident output with no dump callback: a batch call's dumps
    queue nothing, so the lanes are not joined at them.  After every call
    each product's device output equals the one-stream run's bit for bit:
    a 2^15 four-step product dumping inside the batches and the fused 8 +
    1024-point pair of the matrix-core engine's 1e-3 level, every mode."""
    import torch
    import rawspec_amd as R
    from tests.spectro_cases import MFMA_FAST
    kw = dict(Nc=4, Ntpb=65536, Nts=[32768, 8, 1024], Nas=[6, 128, 64], Np=2, Nbps=8)
    for mode in (1, 4, -4):
        print(f"case: device output, no callback, pols {mode}", flush=True)
        sps = []
        for lanes in (True, False):
            sp = R.Spectrometer(**kw, Npolout=[mode] * 3)
            assert sp.set_device_output(True) == 0 and sp.set_mfma(MFMA_FAST) == 0
            assert sp.set_lanes(lanes) == 0 and sp.get_lanes() == int(lanes)
            sps.append(sp)
        nbytes = sps[0].device_input_size()
        g = torch.Generator(device="cuda").manual_seed(7)
        bufs = [torch.randint(0, 256, (nbytes,), dtype=torch.uint8, device="cuda", generator=g)
                for _ in range(8)]
        torch.cuda.synchronize()
        sps[0].profile(True)
        for k in (0, 4):
            outs = []
            for sp in sps:
                assert sp.start_processing_device_batch([b.data_ptr() for b in bufs[k:k + 4]]) == 0
                assert sp.wait_for_completion() == 0
                outs.append(_device_outputs(sp, 3))
            if k == 0:
                names = set(sps[0].profile_report())
                assert "p1.k_small_mfma_u8.N1024" in names or "p2.k_small_mfma_u8.N1024" in names, names
                sps[0].profile(False)
            for i, (x, y) in enumerate(zip(*outs)):
                assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (mode, k, i)
                assert np.isfinite(x).all() and x.any(), (mode, k, i)
        for sp in sps:
            sp.close()
