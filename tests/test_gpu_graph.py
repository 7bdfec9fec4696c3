"""A start call captured into a HIP graph and replayed.

When every product's integration completes within one buffer (as in the
16-antenna, four-product workload), one buffer's work has the same shape
every time, so a caller may capture rawspec_amd_start_processing_device on
the context's stream (rawspec_amd_set_stream) and replay the graph for each
new buffer written into the captured input.  The capture holds the two-lane
fork and join, the kernels of all products, the incoherent sums and the
dump copies into the pinned host buffers.  Every replay is checked against
an eagerly run context on the same data (bit-identical: the same kernels in
the same order) and against the FP64 oracle; eager work and synchronisations
run between the replays.
"""
import numpy as np
import pytest

import oracle as O
from tests.spectro_cases import PLAIN_TOL, gen_blocks, rel_err

pytestmark = pytest.mark.gpu

KW = dict(Nc=4, Ntpb=65536, Nts=[64, 4096, 32768, 8], Nas=[1024, 16, 2, 8192], Np=2, Nbps=16,
          Npolout=[1, -4, 4, 1], Nant=2, incoherently_sum=1, Aws=[0.5, 1.5])


def _outputs(sp, n):
    return [sp.pwrbuf(i).copy() for i in range(n)] + [sp.icsbuf(i).copy() for i in range(n)]


def test_captured_start_call_replays_like_eager_calls():
    import torch
    import rawspec_amd as R
    n = len(KW["Nts"])
    eager = R.Spectrometer(**KW)
    graphed = R.Spectrometer(**KW)
    for sp in (eager, graphed):
        assert sp.set_lanes(1) == 0
    cfg = O.make_cfg(No=n, Np=2, Nant=2, Nc=KW["Nc"], Ntpb=KW["Ntpb"], Nbps=16,
                     Npolout=KW["Npolout"], Nts=KW["Nts"], Nas=KW["Nas"], incoherently_sum=1,
                     Naws=2)
    assert O.derive(cfg) == 0
    assert [cfg.Nis[i] for i in range(n)] == [1] * n, "every integration within one buffer"
    Nb = graphed.ctx.Nb
    host = [O.device_layout(gen_blocks(KW["Nc"], KW["Ntpb"], 2, 16, Nb, seed=21, sigma=1000.0,
                                       tone=(1, 20.0, 300.0, 4096.0), first_blk=k * Nb),
                            KW["Nc"], KW["Ntpb"], 2) for k in range(4)]
    data = [torch.from_numpy(h.view(np.uint8).copy()).to("cuda") for h in host]
    s = torch.cuda.Stream()
    assert graphed.set_stream(s.cuda_stream) == 0
    d = torch.empty_like(data[0])
    torch.cuda.synchronize()
    print("case: warm-up call on the capture stream", flush=True)
    with torch.cuda.stream(s):
        d.copy_(data[0])
    assert graphed.start_processing_device(d.data_ptr(), 1) == 0
    assert graphed.wait_for_completion() == 0
    print("case: capture", flush=True)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        assert graphed.start_processing_device(d.data_ptr(), 1) == 0
    st = O.State(cfg, weights=KW["Aws"])
    for k in range(6):
        b = k % len(data)
        print(f"case: replay {k} on buffer {b}", flush=True)
        with torch.cuda.stream(s):
            d.copy_(data[b])
            g.replay()
        s.synchronize()
        got = _outputs(graphed, n)
        assert eager.start_processing_device(data[b].data_ptr(), 1) == 0
        assert eager.wait_for_completion() == 0
        want = _outputs(eager, n)
        for i, (x, y) in enumerate(zip(got, want)):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"output {i} of replay {k}"
        if k < len(data):
            _, ref, refics = st.process(host[b], 1)
            for i in range(n):
                q = abs(KW["Npolout"][i])
                e = rel_err(got[i], ref[i], cfg.Nds[i], q, KW["Nc"], KW["Nts"][i], KW["Npolout"][i])
                ei = rel_err(got[n + i], refics[i], cfg.Nds[i], q, KW["Nc"] // 2, KW["Nts"][i],
                             KW["Npolout"][i])
                bound = PLAIN_TOL if KW["Nts"][i] < 8192 else 1e-4
                assert e.max() <= bound and ei.max() <= bound, (k, i, e.max(), ei.max())
    st.close()
    del g
    assert graphed.set_stream(0) == 0
    graphed.close()
    eager.close()
