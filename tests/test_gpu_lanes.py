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
