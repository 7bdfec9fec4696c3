"""Serving loop (rawspec_amd.serve): seeded shape mix, admission by HBM and
host budgets, and stream isolation (each stream's dumps are the same alone
and in the mix, and match the FP64 oracle).

The scheduler is exercised on the CPU with a stand-in stream class (it
touches no device itself); the GPU tests drive real contexts through the
C ABI.
"""
from __future__ import annotations

import random

import numpy as np
import pytest

import oracle as O
import rawspec_amd.serve as S

MiB = 1 << 20


def test_random_shapes_are_valid_and_in_range():
    rng = random.Random(3)
    seen_ffts, seen_pols, seen_bits, seen_blocks = set(), set(), set(), set()
    for _ in range(300):
        s = S.random_shape(rng)
        assert 1 <= s.Nc <= 64 and 1 <= s.Nant <= 8 and s.Nc % s.Nant == 0
        assert s.Nbps in (8, 16) and s.Np == 2
        blocsize = s.Nc * s.Ntpb * 2 * 2 * s.Nbps // 8
        assert blocsize in (16 * MiB, 64 * MiB, 128 * MiB)
        assert 1 <= len(s.Nts) <= 4
        for nt, na, po in zip(s.Nts, s.Nas, s.Npolout):
            assert nt & (nt - 1) == 0 and 8 <= nt <= 1 << 20
            assert 1 <= na <= 3072 and po in (1, 4, -4)
        assert (s.Aws is not None) == (s.Nant > 1)
        if s.Aws is not None:
            assert len(s.Aws) == s.Nant and all(0.5 <= w <= 1.5 for w in s.Aws)
        # the oracle's restatement of rawspec_initialize accepts it too, with
        # the same derived counts
        p = S.plan(s)
        cfg = O.make_cfg(No=len(s.Nts), Np=2, Nant=s.Nant, Nc=s.Nc, Ntpb=s.Ntpb, Nbps=s.Nbps,
                         Npolout=s.Npolout, Nts=s.Nts, Nas=s.Nas,
                         incoherently_sum=1 if s.Aws else 0, Naws=len(s.Aws or []))
        assert O.derive(cfg) == 0
        assert cfg.Nb == p.Nb
        assert [cfg.Nis[i] for i in range(len(s.Nts))] == p.Nis
        seen_ffts.update(s.Nts)
        seen_pols.update(s.Npolout)
        seen_bits.add(s.Nbps)
        seen_blocks.add(blocsize)
    assert len(seen_ffts) == 18 and seen_pols == {1, 4, -4} and seen_bits == {8, 16}
    assert len(seen_blocks) == 3


def test_requests_are_seeded_with_poisson_arrivals():
    a = S.make_requests(n=32, seed=9)
    b = S.make_requests(n=32, seed=9)
    assert [(r.arrival, r.shape, r.nbuf) for r in a] == [(r.arrival, r.shape, r.nbuf) for r in b]
    fixed = S.StreamShape(Nc=2, Ntpb=1 << 16, Nts=[1024], Nas=[4], Npolout=[1])
    reqs = S.make_requests(n=4000, rate=4.0, seed=1, shape_fn=lambda rng: fixed)
    t = np.array([r.arrival for r in reqs])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 0.25) < 0.02          # mean 4 arrivals per second
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.08    # exponential: cv = 1


class _FakeStream:
    """Stand-in for one admitted stream: finishes after nbuf pumps, records
    how many streams were live at once."""
    live = 0
    max_live = 0

    def __init__(self, req, srv, plan, t_arrival, t_admit):
        self.req, self.plan = req, plan
        self.sent, self.lat, self.dumps = 0, [], []
        self.t_arrival = t_arrival
        _FakeStream.live += 1
        _FakeStream.max_live = max(_FakeStream.max_live, _FakeStream.live)

    def pump(self):
        if self.sent < self.req.nbuf:
            self.sent += 1
            if self.sent % self.plan.Nis[0] == 0:
                self.lat.append(1e-3)
                self.dumps.append(self.sent)
            return True
        return False

    @property
    def finished(self):
        return self.sent == self.req.nbuf

    def close(self):
        _FakeStream.live -= 1


def _fake_reqs(n, seed=4):
    rng = random.Random(seed)
    return S.make_requests(n=n, rate=1e4, seed=seed, stream_bytes=64 * MiB,
                           shape_fn=lambda r: S.random_shape(r, block_mib=(16,), max_fft_log2=14,
                                                             max_nint=4))


def test_admission_holds_the_hbm_budget_and_retires_streams():
    reqs = _fake_reqs(24)
    srv = S.Server(stream_cls=_FakeStream, record_dumps=True)
    foot = sorted(srv.footprint(r)[0] for r in reqs)
    budget = sum(foot[-3:])                 # room for the three largest at most
    srv.hbm_budget = budget
    _FakeStream.live = _FakeStream.max_live = 0
    st = srv.run(reqs, arrival_scale=0.0)
    assert st["peak_hbm_bytes"] <= budget
    assert 1 < st["peak_active"] and _FakeStream.max_live == st["peak_active"]
    assert _FakeStream.live == 0
    assert sorted(st["results"]) == [r.sid for r in reqs]        # every stream completed
    assert st["ingest_bytes"] == sum(r.nbuf * S.plan(r.shape).buf_bytes for r in reqs)
    for r in reqs:
        nis = S.plan(r.shape).Nis[0]
        assert st["results"][r.sid] == list(range(nis, r.nbuf + 1, nis))


def test_admission_slots_and_oversized_requests():
    reqs = _fake_reqs(10, seed=6)
    _FakeStream.live = _FakeStream.max_live = 0
    st = S.Server(stream_cls=_FakeStream, max_active=1).run(reqs, arrival_scale=0.0)
    assert st["peak_active"] == 1 and _FakeStream.max_live == 1
    _FakeStream.live = _FakeStream.max_live = 0
    st = S.Server(stream_cls=_FakeStream, max_active=4).run(reqs, arrival_scale=0.0)
    assert st["peak_active"] == 4
    small = S.Server(stream_cls=_FakeStream, hbm_budget=1 * MiB)
    with pytest.raises(ValueError):
        small.run(reqs[:1])
    small = S.Server(stream_cls=_FakeStream, host_budget=1 * MiB)
    with pytest.raises(ValueError):
        small.run(reqs[:1])


def test_later_arrivals_join_running_streams():
    """Continuous batching: a request arriving while others run is admitted
    without waiting for them to finish."""
    fixed = S.StreamShape(Nc=4, Ntpb=1 << 20, Nts=[1024], Nas=[1], Npolout=[1])
    reqs = [S.Request(i, fixed, arrival=0.002 * i, nbuf=4000, seed=i) for i in range(3)]

    class Slow(_FakeStream):
        def pump(self):
            import time
            time.sleep(2e-5)
            return super().pump()

    _FakeStream.live = _FakeStream.max_live = 0
    st = S.Server(stream_cls=Slow).run(reqs, arrival_scale=1.0)
    assert st["peak_active"] == 3


# -- GPU --------------------------------------------------------------------

def _small_mix(n=8, seed=21):
    return S.make_requests(
        n=n, rate=1e4, seed=seed, stream_bytes=4 * MiB, max_stream_bytes=12 * MiB,
        shape_fn=lambda r: S.random_shape(r, max_fft_log2=15, block_mib=(1, 2), max_chans=8,
                                          max_nint=8))


def _by_key(dumps):
    return {(k, i): (p, q) for (k, i, p, q) in dumps}


@pytest.mark.gpu
@pytest.mark.parametrize("batch,mfma", [(1, False), (4, True)])
def test_stream_results_same_alone_and_in_the_mix(batch, mfma):
    reqs = _small_mix()
    print("mix:", [(r.sid, r.shape.Nc, r.shape.Nbps, r.shape.Nts, r.shape.Nas, r.shape.Npolout,
                    r.nbuf) for r in reqs], f"batch {batch} mfma {mfma}", flush=True)
    mix = S.Server(record_dumps=True, max_active=len(reqs), batch=batch, mfma=mfma).run(
        reqs, arrival_scale=0.0)
    assert mix["peak_active"] > 1
    assert _check_alone_vs_mix(reqs, mix, batch=batch, mfma=mfma) >= len(reqs)


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [1, 3])
def test_event_timed_latency_counts_every_dump(batch):
    """Without recorded dumps the contexts get no callback and each dump is
    timed from the completion event of the call that completes it: one
    latency per dump a product completes (nbuf // Nis each), as many as the
    callback-timed run of the same requests counts, all positive and no
    longer than that stream's whole run."""
    reqs = _small_mix(n=6, seed=23)
    want = sum(r.nbuf // nis for r in reqs for nis in S.plan(r.shape).Nis)
    print("case: event-timed dumps, batch", batch, "expected", want, flush=True)
    ev = S.Server(max_active=len(reqs), batch=batch).run(reqs, arrival_scale=0.0)
    cb = S.Server(record_dumps=True, max_active=len(reqs), batch=batch).run(reqs, arrival_scale=0.0)
    assert ev["dumps"] == cb["dumps"] == want > 0
    assert 0.0 < ev["p50_ms"] <= ev["p99_ms"] <= ev["wall_s"] * 1e3


def _check_alone_vs_mix(reqs, mix, **kw):
    """Every stream's dumps in the mix are bit-identical to the stream run
    alone (same server options kw, one slot)."""
    nd = 0
    for r in reqs:
        alone = S.Server(record_dumps=True, max_active=1, **kw).run([r], arrival_scale=0.0)
        a, m = _by_key(alone["results"][r.sid]), _by_key(mix["results"][r.sid])
        p = S.plan(r.shape)
        expect = {(k - 1, i) for i in range(len(r.shape.Nts))
                  for k in range(p.Nis[i], r.nbuf + 1, p.Nis[i])}
        assert set(a) == set(m) == expect, r.sid
        for key in a:
            assert np.array_equal(a[key][0].view(np.uint32), m[key][0].view(np.uint32)), (r.sid, key)
            if a[key][1] is not None:
                assert np.array_equal(a[key][1].view(np.uint32), m[key][1].view(np.uint32))
            nd += 1
    return nd


def _check_oracle(srv, reqs, mix, mfma):
    """Every stream's dumps against the FP64 oracle fed the same buffers."""
    import torch
    from tests.spectro_cases import (MFMA_FLOOR_MULT, PLAIN_TOL, FP32_FLOOR_NT, Fp32Ref,
                                     rel_err)
    for r in reqs:
        s, p = r.shape, S.plan(r.shape)
        ring = [b.cpu().numpy() for b in srv.make_ring(r, p)]   # same seed: same buffers
        torch.cuda.synchronize()
        No = len(s.Nts)
        cfg = O.make_cfg(No=No, Np=2, Nant=s.Nant, Nc=s.Nc, Ntpb=s.Ntpb, Nbps=s.Nbps,
                         Npolout=s.Npolout, Nts=s.Nts, Nas=s.Nas,
                         incoherently_sum=1 if s.Aws else 0, Naws=len(s.Aws or []))
        assert O.derive(cfg) == 0
        st = O.State(cfg, weights=s.Aws)
        f32 = {i: Fp32Ref(cfg, i) for i in range(No) if s.Nts[i] >= FP32_FLOOR_NT}
        got = _by_key(mix["results"][r.sid])
        for k in range(r.nbuf):
            dev = ring[k % len(ring)]
            _, ref, refics = st.process(dev, 1)
            fl = {i: f.process(dev, 1) for i, f in f32.items()}
            for i in ref:
                q = abs(cfg.Npolout[i])
                g, gi = got[(k, i)]
                e = rel_err(g, ref[i], cfg.Nds[i], q, cfg.Nc, cfg.Nts[i], cfg.Npolout[i]).max()
                tol = PLAIN_TOL
                if i in fl:
                    floor = rel_err(fl[i], ref[i], cfg.Nds[i], q, cfg.Nc, cfg.Nts[i],
                                    cfg.Npolout[i]).max()
                    tol = max(tol, (MFMA_FLOOR_MULT if mfma else 2.0) * floor)
                assert e <= tol, (r.sid, k, i, e, tol)
                if s.Aws:
                    ei = rel_err(gi, refics[i], cfg.Nds[i], q, cfg.Nc // cfg.Nant, cfg.Nts[i],
                                 cfg.Npolout[i]).max()
                    assert ei <= PLAIN_TOL, (r.sid, k, i, ei)
        st.close()


@pytest.mark.gpu
@pytest.mark.parametrize("batch,mfma", [(1, False), (4, True)])
def test_mixed_streams_match_the_oracle(batch, mfma):
    reqs = _small_mix(n=6, seed=33)
    srv = S.Server(record_dumps=True, max_active=len(reqs), batch=batch, mfma=mfma)
    mix = srv.run(reqs, arrival_scale=0.0)
    _check_oracle(srv, reqs, mix, mfma)


@pytest.mark.gpu
def test_serving_shape_ranges_shared_streams():
    """The serving workload's shape ranges at reduced channel counts: Nfft up
    to 2^20, blocks of 16, 64 or 128 MiB, Nint up to 3072, 8 or 16 bits,
    1-4 products, pols 1, 4 or -4, up to 4 channels (and antennas); this
    seeded mix holds products of 2^20 (64 spectra), 65536, 8192, 1024 and
    8 to 32 points, int8 and int16, 64 and 128 MiB buffers and a two-antenna
    stream with the incoherent sum.  The mix runs its contexts on two shared
    HIP streams; every stream's dumps are bit-identical to the stream run
    alone on its own stream and match the FP64 oracle."""
    def shape(rng):
        # drawn from the ranges, redrawn until every product dumps within
        # three buffers (so every product's output is checked)
        while True:
            sh = S.random_shape(rng, max_fft_log2=20, block_mib=(16, 64, 128), max_chans=4,
                                max_nint=3072)
            if max(S.plan(sh).Nis) <= 3:
                return sh
    reqs = S.make_requests(n=4, rate=1e4, seed=47, stream_bytes=384 * MiB,
                           max_stream_bytes=512 * MiB, shape_fn=shape)
    print("mix:", [(r.sid, r.shape.Nc, r.shape.Nbps, r.shape.Ntpb, r.shape.Nts, r.shape.Nas,
                    r.shape.Npolout, r.shape.Nant, r.nbuf) for r in reqs], flush=True)
    assert max(max(r.shape.Nts) for r in reqs) == 1 << 20
    srv = S.Server(record_dumps=True, max_active=len(reqs), streams=2)
    mix = srv.run(reqs, arrival_scale=0.0)
    assert mix["peak_active"] > 1
    assert _check_alone_vs_mix(reqs, mix) >= len(reqs)
    _check_oracle(srv, reqs, mix, False)
