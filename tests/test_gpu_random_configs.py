"""Seeded random configurations against the FP64 oracle, through the
reference API (host blocks, rawspec_start_processing) and through the
batched device entry point (random batch sizes, two-lane execution forced on
or off, or left to the library): transform lengths that
are powers of two and other lengths (each a divisor of the longest), one or
two pols, int8 or int16, every polarisation mode, forward and inverse,
conjugated input, several channels, the incoherent sum over antennas,
integrations within and over buffers, and the three engine levels, each at
its own bar.  The seed names the case; every case must pass."""
import random

import pytest

import oracle as O
from tests.spectro_cases import (MFMA_FLOOR_MULT, PLAIN_FLOOR_MULT, PLAIN_TOL, check_results, run_case,
                                 run_device_case, unsplit_products)

pytestmark = pytest.mark.gpu


def _config(seed):
    rng = random.Random(1000 + seed)
    for _ in range(1000):
        No = rng.randint(1, 3)
        if rng.random() < 0.25:
            ntmax = rng.choice([12, 96, 100, 240, 1000, 1536, 2000])
        else:
            ntmax = 1 << rng.randint(3, 20)
        divs = [d for d in range(2, ntmax + 1) if ntmax % d == 0]
        nts = [ntmax] + [rng.choice(divs) for _ in range(No - 1)]
        Np = rng.choice([1, 2, 2])
        Nbps = rng.choice([8, 8, 16])
        Nant = rng.choice([1, 1, 2])
        Nc = Nant * rng.choice([1, 2, 3])
        Ntpb = ntmax * rng.choice([1, 2, 4, 8])
        if Nc * Ntpb > 1 << 21:          # keeps the oracle to seconds
            continue
        nas = []
        for nt in nts:
            nss = Ntpb // nt
            nas.append(rng.choice([1, 2, 4, 16, nss, 2 * nss, 3 * nss]))
        pols = [rng.choice([1, 4, -4]) if Np == 2 else 1 for _ in nts]
        cfg = O.make_cfg(No=No, Np=Np, Nant=Nant, Nc=Nc, Ntpb=Ntpb, Nbps=Nbps, Npolout=pols,
                         Nts=nts, Nas=nas, incoherently_sum=1 if Nant > 1 else 0,
                         Naws=Nant if Nant > 1 else 0)
        if O.derive(cfg) != 0 or max(cfg.Nis[i] for i in range(No)) > 6:
            continue
        kw = dict(Nc=Nc, Ntpb=Ntpb, Nts=nts, Nas=nas, Np=Np, Nbps=Nbps, Npolout=pols,
                  nbuf=max(cfg.Nis[i] for i in range(No)), fft_dir=rng.choice([1, -1]),
                  conj=rng.choice([0, 1]), seed=seed, tone=(0, 20.0, 3.0, float(ntmax)),
                  mfma=rng.choice([0, 0, 1, 2]))
        if Nant > 1:
            kw.update(Nant=Nant, ics=True, Aws=[0.5 + 0.25 * a for a in range(Nant)])
        return kw
    raise AssertionError("no valid configuration drawn")


@pytest.mark.parametrize("seed", range(200))
def test_random_configuration(seed):
    kw = _config(seed)
    print(f"case: seed {seed} {kw}", flush=True)
    res, cfg, _ = run_case(**kw)
    level = kw["mfma"]
    check_results(res, cfg, tol=PLAIN_TOL, ics=kw.get("ics", False),
                  floor_mult=MFMA_FLOOR_MULT if level == 2 else PLAIN_FLOOR_MULT,
                  tol_by_product=unsplit_products(kw["Nts"], kw["Nas"], kw["Np"], kw["Nbps"],
                                                  mfma=level))
    assert res, "no dump"


@pytest.mark.parametrize("seed", range(100))
def test_random_configuration_batched(seed):
    kw = _config(seed)
    rng = random.Random(seed)
    nbuf = max(kw.pop("nbuf"), rng.choice([2, 4, 6, 8]))
    batches, left = [], nbuf
    while left:
        batches.append(rng.randint(1, left))
        left -= batches[-1]
    lanes = rng.choice([None, True, False])
    print(f"case: seed {seed} batches {batches} lanes {lanes} {kw}", flush=True)
    res, cfg = run_device_case(nbuf=nbuf, batches=batches, lanes=lanes, **kw)
    level = kw["mfma"]
    check_results(res, cfg, tol=PLAIN_TOL, ics=kw.get("ics", False),
                  floor_mult=MFMA_FLOOR_MULT if level == 2 else PLAIN_FLOOR_MULT,
                  tol_by_product=unsplit_products(kw["Nts"], kw["Nas"], kw["Np"], kw["Nbps"],
                                                  mfma=level))
    assert res, "no dump"
