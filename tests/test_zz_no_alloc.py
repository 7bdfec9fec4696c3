"""The allocation-free entry point, rawspec_amd_start_processing_device_noalloc,
against rawspec_amd_start_processing_device on the same device buffers.

Both run the same kernels; the allocation-free call keeps its partial sums and
the four-step intermediate in the library's static per-GPU scratch and never
allocates held first-pass output, so the reference context runs with held
output off and every dump must be bit-identical.  Two calls in a row on two
configurations, the second smaller (fewer channels, shorter blocks), so the
scratch is reused at a different size; dumps are also checked against the
FP64 oracle at the plain bar (at engine level 1, the matrix-core engine's own
bar for the products it runs below FP32 precision).  Ends with a
configuration whose intermediate exceeds the scratch, which must be refused
before anything is enqueued.
"""
import numpy as np
import pytest

import oracle as O
from tests.spectro_cases import PLAIN_TOL, gen_blocks, rel_err, unsplit_products

pytestmark = pytest.mark.gpu


def _pair(R, kw, got, mfma):
    def cb(name):
        def on_dump(sp, product, cbtype):
            if cbtype == R.RAWSPEC_CALLBACK_POST_DUMP:
                got[name].append((product, sp.pwrbuf(product).copy()))
        return on_dump
    ref = R.Spectrometer(callback=cb("ref"), **kw)
    na = R.Spectrometer(callback=cb("noalloc"), **kw)
    assert ref.set_held_buffers(0) == 0
    assert ref.set_mfma(mfma) == 0 and na.set_mfma(mfma) == 0
    return ref, na


def _case(R, torch, *, Nc, Ntpb, Nts, Nas, Npolout, nbuf, seed, mfma=0):
    kw = dict(Nc=Nc, Ntpb=Ntpb, Nts=Nts, Nas=Nas, Np=2, Nbps=8, Npolout=Npolout)
    print(f"case: noalloc vs device call {kw} nbuf {nbuf} engine level {mfma}", flush=True)
    got = {"ref": [], "noalloc": []}
    ref, na = _pair(R, kw, got, mfma)
    Nb = ref.ctx.Nb
    host = [O.device_layout(gen_blocks(Nc, Ntpb, 2, 8, Nb, seed=seed, tone=(0, 20.0, 300.0, 1024.0),
                                       first_blk=n * Nb), Nc, Ntpb, 2) for n in range(nbuf)]
    dev = [torch.from_numpy(h.view(np.uint8).copy()).to("cuda") for h in host]
    torch.cuda.synchronize()
    for d in dev:
        assert ref.start_processing_device(d.data_ptr(), 1) == 0
        assert na.start_processing_device_noalloc(d.data_ptr(), 1) == 0
    assert ref.wait_for_completion() == 0 and na.wait_for_completion() == 0
    ref.close()
    na.close()
    assert len(got["ref"]) == len(got["noalloc"]) > 0
    for (pa, a), (pb, b) in zip(got["ref"], got["noalloc"]):
        assert pa == pb
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"product {pa} differs"
    # and both against the oracle
    cfg = O.make_cfg(No=len(Nts), Np=2, Nant=1, Nc=Nc, Ntpb=Ntpb, Nbps=8, Npolout=Npolout,
                     Nts=Nts, Nas=Nas)
    assert O.derive(cfg) == 0
    st = O.State(cfg)
    refs = []
    for h in host:
        _, r, _ = st.process(h, 1)
        refs += [(i, r[i]) for i in sorted(r)]
    st.close()
    gpu = sorted(got["noalloc"], key=lambda t: t[0])
    refs = sorted(refs, key=lambda t: t[0])
    assert [p for p, _ in gpu] == [p for p, _ in refs]
    level1 = unsplit_products(Nts, Nas, mfma=mfma) if mfma else {}
    for (i, g), (_, r) in zip(gpu, refs):
        e = rel_err(g, r, cfg.Nds[i], abs(Npolout[i]), Nc, Nts[i], Npolout[i]).max()
        bound = level1.get(i, PLAIN_TOL if Nts[i] < 8192 else 5e-5)
        assert e <= bound, f"product {i}: {e:.2e}"
    del dev


def test_noalloc_matches_device_call():
    import torch
    import rawspec_amd as R
    # a four-step product integrating over two buffers, a 1024-point product
    # whose dump groups split over workgroups (partial sums in the scratch)
    # and an 8-point product
    _case(R, torch, Nc=8, Ntpb=131072, Nts=[131072, 1024, 8], Nas=[2, 256, 64],
          Npolout=[-4, -4, 1], nbuf=4, seed=11)
    # smaller: fewer channels, shorter blocks, other products
    _case(R, torch, Nc=2, Ntpb=32768, Nts=[16384, 512], Nas=[4, 64],
          Npolout=[4, 1], nbuf=4, seed=12)
    # four-step products whose pass B splits each dump group into runs of
    # spectra (partial sums in the scratch): dumps within a buffer (8192
    # points, 2 groups of 16 spectra) and across four buffers (16384 points)
    _case(R, torch, Nc=2, Ntpb=1 << 18, Nts=[8192, 16384, 8], Nas=[16, 64, 128],
          Npolout=[-4, 4, 1], nbuf=4, seed=13)
    # the matrix-core engine at its 1e-3 bar with a fused pair: one kernel
    # computes the 8-point and the unsplit 1024-point product, whose dump
    # groups split over wavefronts (partial sums in the scratch) and span two
    # buffers
    _case(R, torch, Nc=4, Ntpb=65536, Nts=[8, 1024], Nas=[128, 128], Npolout=[-4, -4],
          nbuf=4, seed=14, mfma=1)


def test_noalloc_refuses_oversized_scratch():
    import torch
    import rawspec_amd as R
    # one channel of a 2^24-point product needs 256 MiB of intermediate, and
    # the 1024-point product's partial sums take scratch space beside it
    sp = R.Spectrometer(Nc=1, Ntpb=1 << 24, Nts=[1 << 24, 1024], Nas=[1, 16384], Np=2, Nbps=8,
                        Npolout=[1, 1])
    d = torch.zeros(sp.device_input_size(), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    print("case: oversized configuration refused", flush=True)
    assert sp.start_processing_device_noalloc(d.data_ptr(), 1) == 1
    # the regular entry point still works on the same context afterwards
    assert sp.start_processing_device(d.data_ptr(), 1) == 0
    assert sp.wait_for_completion() == 0
    sp.close()
