"""Device memory of a context against rawspec_amd_plan(), the figure the
serving loop admits streams by: once its integrations have started (the
held first-pass store and every lazily sized buffer allocated), a context
holds what the plan reports (input buffer, integration and internal
accumulators, partial sums of split groups and split pass-B launches, the
four-step intermediate, Bluestein and twiddle tables), to the allocator's
page rounding; closing it returns all of it, and contexts created and
closed over and over leak nothing."""
import pytest

import rawspec_amd.serve as S

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _free():
    import torch
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


def _run(shape, nbuf):
    """Creates the context, runs nbuf buffers (device input), returns it."""
    import torch
    import rawspec_amd as R
    sp = R.Spectrometer(**shape.spectrometer_kwargs())
    d = torch.zeros(sp.device_input_size(), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    for _ in range(nbuf):
        assert sp.start_processing_device(d.data_ptr(), 1) == 0
    assert sp.wait_for_completion() == 0
    del d
    torch.cuda.empty_cache()
    return sp


SHAPES = {
    # 2^20 points over 4 channels integrating 8 buffers: the held store
    # (8 buffers of first-pass output) beside a 1024-point product
    "held": (S.StreamShape(Nc=4, Ntpb=1 << 20, Nts=[1 << 20, 1024], Nas=[8, 16],
                           Npolout=[-4, 1]), 9),
    # 8192 points over 2 channels: pass B split into runs (d_partb)
    "split": (S.StreamShape(Nc=2, Ntpb=1 << 20, Nts=[8192, 16], Nas=[512, 64],
                            Npolout=[4, -4]), 5),
    # lengths that are not powers of two (Bluestein tables), incoherent sum
    "bluestein": (S.StreamShape(Nc=4, Ntpb=1000 * 64, Nts=[1000, 100], Nas=[64, 8],
                                Npolout=[-4, 1], Nant=2, Aws=[1.0, 0.5]), 2),
}


@pytest.mark.parametrize("name", list(SHAPES))
def test_context_holds_what_the_plan_reports(name):
    shape, nbuf = SHAPES[name]
    _run(*SHAPES["split"]).close()            # code objects loaded before measuring
    plan = S.plan(shape)
    assert plan is not None
    f0 = _free()
    sp = _run(shape, nbuf)
    used = f0 - _free()
    print(f"case: {name} plan {plan.device_bytes / MiB:.1f} MiB, held by the context "
          f"{used / MiB:.1f} MiB", flush=True)
    assert abs(used - plan.device_bytes) <= 32 * MiB + plan.device_bytes // 100
    sp.close()
    assert f0 - _free() <= 8 * MiB


def test_create_close_cycles_leak_nothing():
    shape, nbuf = SHAPES["split"]
    _run(shape, nbuf).close()
    f0 = _free()
    for k in range(20):
        _run(shape, nbuf).close()
    print(f"case: 20 create/close cycles, {(f0 - _free()) / MiB:.1f} MiB not returned", flush=True)
    assert f0 - _free() <= 8 * MiB
