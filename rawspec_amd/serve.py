"""Serving loop: concurrent voltage streams of different shapes on one GPU.

Each admitted stream owns one librawspec context, which keeps its partial
integrations resident in device memory and runs on its own HIP stream, so
the device work of all admitted streams overlaps (contexts keep every piece
of state in their own buffers and kernel arguments, so any number of them
coexist in one process).  The loop does continuous batching, not queueing:

  * requests arrive over time (seeded Poisson arrivals);
  * a request is admitted as soon as a slot is free and its footprint fits
    next to the streams already running: the device bytes
    rawspec_amd_plan() reports plus its input ring against the HBM budget,
    and the pinned host bytes against the host budget; requests wait in
    arrival order otherwise;
  * every admitted stream keeps up to `depth` submissions (each of `batch`
    consecutive buffers) queued on its HIP stream, refilled as soon as one
    completes, interleaved with all others;
  * a finished stream is retired at once and its memory returns to the
    budgets for the next arrival.

Block-to-dump latency is measured per dump: from the call that submitted the
buffer completing an integration to the moment the loop sees that call
complete (its event, polled every sweep; queueing in front of the GPU
included, admission wait excluded and reported separately).  Contexts get no
dump callback then: each callback is a host function that holds the
context's stream, and the hardware queue it shares with other contexts,
until it returns, and enqueueing one costs about 25 us of host time per
call.  With record_dumps the POST dump callback copies each dump (the host
buffers are reused by the next dump) and times it instead.

There is no reference counterpart: the reference serves one stream per
process (its texture object and callback pointers are process globals).
Every stream here still goes through the reference's C API semantics
(rawspec_initialize, start processing, dump callbacks,
rawspec_check_for_completion / rawspec_wait_for_completion, rawspec_cleanup).
"""
from __future__ import annotations

import collections
import ctypes as C
import dataclasses
import random
import time

import numpy as np

from . import RAWSPEC_CALLBACK_POST_DUMP, RawspecContext, Spectrometer, lib

__all__ = ["StreamShape", "Request", "plan", "random_shape", "make_requests", "Server"]

MiB = 1 << 20
GiB = 1 << 30


@dataclasses.dataclass
class StreamShape:
    Nc: int
    Ntpb: int
    Nts: list
    Nas: list
    Npolout: list
    Np: int = 2
    Nbps: int = 8
    Nant: int = 1
    Aws: list = None          # antenna weights; incoherent sum when Nant > 1

    def spectrometer_kwargs(self):
        kw = dict(Nc=self.Nc, Ntpb=self.Ntpb, Nts=list(self.Nts), Nas=list(self.Nas),
                  Npolout=list(self.Npolout), Np=self.Np, Nbps=self.Nbps, Nant=self.Nant)
        if self.Aws is not None:
            kw.update(incoherently_sum=1, Aws=list(self.Aws))
        return kw


@dataclasses.dataclass
class Request:
    sid: int
    shape: StreamShape
    arrival: float            # seconds after the start of the run
    nbuf: int                 # device buffers the stream sends
    seed: int


@dataclasses.dataclass
class Plan:
    device_bytes: int         # rawspec_amd_plan(): what rawspec_initialize allocates on the GPU
    host_bytes: int           # pinned host bytes rawspec_initialize allocates
    buf_bytes: int            # one device input buffer (Nb blocks)
    Nb: int
    Nis: list                 # buffers per dump, per product


def plan(shape: StreamShape):
    """Plan for a shape, or None when rawspec_initialize would reject it."""
    c = RawspecContext()
    c.No, c.Np, c.Nant, c.Nc, c.Ntpb, c.Nbps = (len(shape.Nts), shape.Np, shape.Nant, shape.Nc,
                                                shape.Ntpb, shape.Nbps)
    for i, (nt, na, po) in enumerate(zip(shape.Nts, shape.Nas, shape.Npolout)):
        c.Nts[i], c.Nas[i], c.Npolout[i] = nt, na, po
    aws = None
    if shape.Aws is not None:
        aws = (C.c_float * len(shape.Aws))(*shape.Aws)
        c.incoherently_sum, c.Naws, c.Aws = 1, len(shape.Aws), C.cast(aws, C.POINTER(C.c_float))
    dev, host = C.c_size_t(), C.c_size_t()
    if lib().rawspec_amd_plan(C.byref(c), C.byref(dev), C.byref(host)) != 0:
        return None
    nsamp = c.Nb * c.Ntpb
    nis = [max(1, na // (nsamp // nt)) for nt, na in zip(shape.Nts, shape.Nas)]
    return Plan(dev.value, host.value, c.Nc * nsamp * c.Np * 2 * c.Nbps // 8, c.Nb, nis)


def random_shape(rng: random.Random, max_fft_log2=20, block_mib=(16, 64, 128), max_chans=64,
                 max_nint=3072) -> StreamShape:
    """One seeded shape of the serving mix: coarse channels uniform in
    1..max_chans, antennas in 1..8, 8 or 16 bits, 2 pols, blocks of one of
    block_mib MiB, 1-4 products with Nfft from 2^3..2^max_fft_log2, Nint
    uniform in 1..max_nint and pols from {1, 4, -4}.  Draws that rawspec_initialize would reject are
    redrawn (the whole geometry when the channels, antennas, block size and
    Nffts are incompatible; Nint alone when only Nfft*Nint fails to divide or
    be divided by the samples per buffer).  Streams with several antennas
    also form the incoherent sum, with weights uniform in [0.5, 1.5]."""
    while True:
        nbps = rng.choice([8, 16])
        blocsize = rng.choice(block_mib) * MiB
        nc = rng.randint(1, max_chans)
        nant = rng.randint(1, 8)
        per = nc * 2 * 2 * nbps // 8
        if nc % nant or blocsize % per:
            continue
        ntpb = blocsize // per
        no = rng.randint(1, 4)
        nts = [1 << rng.randint(3, max_fft_log2) for _ in range(no)]
        ntmax = max(nts)
        if ntmax <= ntpb:
            if ntpb % ntmax:
                continue
            nsamp = ntpb
        else:
            if ntmax % ntpb:
                continue
            nsamp = ntmax
        nas = []
        for nt in nts:
            nss = nsamp // nt
            while True:
                na = rng.randint(1, max_nint)
                if (nss % na == 0) if na <= nss else (na % nss == 0):
                    break
            nas.append(na)
        pols = [rng.choice([1, 4, -4]) for _ in range(no)]
        aws = [rng.uniform(0.5, 1.5) for _ in range(nant)] if nant > 1 else None
        s = StreamShape(Nc=nc, Ntpb=ntpb, Nts=nts, Nas=nas, Npolout=pols, Nbps=nbps, Nant=nant,
                        Aws=aws)
        if plan(s) is not None:
            return s


def make_requests(n=32, rate=4.0, seed=5, stream_bytes=8 * GiB, max_stream_bytes=16 * GiB,
                  shape_fn=None):
    """n requests with Poisson arrivals (mean `rate` per second).  Each stream
    sends about stream_bytes of input, more when two dumps of its
    fastest-dumping product need more, up to max_stream_bytes (at least one
    buffer)."""
    rng = random.Random(seed)
    t = 0.0
    reqs = []
    for sid in range(n):
        t += rng.expovariate(rate)
        shape = shape_fn(rng) if shape_fn else random_shape(rng)
        p = plan(shape)
        nbuf = max(stream_bytes // p.buf_bytes,
                   min(2 * min(p.Nis), max_stream_bytes // p.buf_bytes), 1)
        reqs.append(Request(sid, shape, t, int(nbuf), seed * 1000 + sid))
    return reqs


class _Stream:
    """One admitted request: its context, input ring and in-flight buffers."""

    def __init__(self, req, srv, plan_, t_arrival, t_admit):
        import torch
        self.req, self.plan = req, plan_
        self.t_arrival, self.t_admit = t_arrival, t_admit
        self.depth, self.batch = srv.depth, srv.batch
        self.sent = 0
        self.t_sub = []
        self.inflight = collections.deque()
        self.ndump = [0] * len(req.shape.Nts)
        self.lat = []
        self.dumps = [] if srv.record_dumps else None
        self.sp = Spectrometer(**req.shape.spectrometer_kwargs(), gpu_index=srv.gpu,
                               callback=self._on_dump if srv.record_dumps else None,
                               record_events=False)
        if self.sp.set_device_output(not srv.record_dumps):
            raise RuntimeError("rawspec_amd_set_device_output failed")
        if self.sp.set_mfma(srv.mfma):
            raise RuntimeError("rawspec_amd_set_mfma failed")
        # one stream per context: dozens of concurrent contexts with three
        # streams each would share the few hardware queues badly
        if self.sp.set_lanes(srv.lanes):
            raise RuntimeError("rawspec_amd_set_lanes failed")
        assert self.sp.device_input_size() == plan_.buf_bytes
        pool = srv.stream_pool()
        if pool:
            # contexts share the server's few streams (rawspec_amd_set_stream)
            self.hs = pool[srv.take_slot() % len(pool)]
            if self.sp.set_stream(self.hs.cuda_stream):
                raise RuntimeError("rawspec_amd_set_stream failed")
        else:
            self.hs = torch.cuda.ExternalStream(self.sp.stream(), device=f"cuda:{srv.gpu}")
        self.ring = srv.make_ring(req, plan_)

    def _on_dump(self, sp, product, cbtype):
        # runs on the HIP runtime's callback thread, in stream order
        if cbtype != RAWSPEC_CALLBACK_POST_DUMP:
            return
        now = time.perf_counter()
        self.ndump[product] += 1
        k = self.ndump[product] * self.plan.Nis[product]        # 1-based buffer index
        self.lat.append(now - self.t_sub[k - 1])
        if self.dumps is not None:
            ics = sp.icsbuf(product).copy() if self.req.shape.Aws is not None else None
            self.dumps.append((k - 1, product, sp.pwrbuf(product).copy(), ics))

    def pump(self):
        """Retires completed buffers and queues new ones; True if it did either."""
        import torch
        moved = False
        while self.inflight and self.inflight[0][0].query():
            _ev, s0, nb = self.inflight.popleft()
            if self.dumps is None:
                # dumps completed by buffers s0 + 1 .. s0 + nb (1-based)
                now = time.perf_counter()
                for p, nis in enumerate(self.plan.Nis):
                    for k in range((s0 // nis + 1) * nis, s0 + nb + 1, nis):
                        self.ndump[p] += 1
                        self.lat.append(now - self.t_sub[k - 1])
            moved = True
        while len(self.inflight) < self.depth and self.sent < self.req.nbuf:
            nb = min(self.batch, self.req.nbuf - self.sent)
            t = time.perf_counter()
            self.t_sub.extend([t] * nb)
            ptrs = [self.ring[(self.sent + j) % len(self.ring)].data_ptr() for j in range(nb)]
            rc = (self.sp.start_processing_device(ptrs[0], 1) if nb == 1
                  else self.sp.start_processing_device_batch(ptrs, 1))
            if rc:
                raise RuntimeError(f"stream {self.req.sid}: start processing failed")
            ev = torch.cuda.Event()
            ev.record(self.hs)
            self.inflight.append((ev, self.sent, nb))
            self.sent += nb
            moved = True
        return moved

    @property
    def finished(self):
        return self.sent == self.req.nbuf and not self.inflight

    def close(self):
        if self.sp.wait_for_completion():
            raise RuntimeError(f"stream {self.req.sid}: wait_for_completion failed")
        self.sp.close()
        self.ring = None


class Server:
    """Continuous-batching scheduler over librawspec contexts on one GPU.

    hbm_budget / host_budget: bytes the admitted streams may hold together
    on the device (contexts plus input rings) and in pinned host memory.
    max_active: concurrent streams (slots); 1 runs the streams one after the
    other.  depth: submissions queued per stream; batch: buffers per
    submission (1: rawspec_amd_start_processing_device, more: the batched
    entry point).  mfma: transform engine of every stream's context.  ring:
    distinct device input buffers per stream (buffer k reads
    ring[k % ring]).  streams: HIP streams the contexts share (assigned in
    turn at admission, rawspec_amd_set_stream); 0 gives every context its
    own stream.  stream_cls: the
    class that serves one admitted request (its constructor, pump(),
    finished, close(), sent, lat, dumps, plan); the scheduler itself touches
    no device.
    """

    def __init__(self, hbm_budget=200e9, host_budget=96 * GiB, max_active=32, depth=2, ring=2,
                 gpu=0, record_dumps=False, poll_s=20e-6, stream_cls=None, batch=1, mfma=False,
                 lanes=False, streams=0):
        self.hbm_budget, self.host_budget = hbm_budget, host_budget
        self.max_active, self.depth, self.nring = max_active, depth, ring
        self.batch, self.mfma = max(1, int(batch)), int(mfma)
        self.nstreams, self._pool, self._slot = int(streams), None, 0
        self.lanes = bool(lanes)
        self.gpu, self.record_dumps, self.poll_s = gpu, record_dumps, poll_s
        self.stream_cls = stream_cls or _Stream
        self._gen = None

    def stream_pool(self):
        """The shared streams (default priority), created on first use."""
        if self.nstreams <= 0:
            return None
        if self._pool is None:
            import torch
            self._pool = [torch.cuda.Stream(device=f"cuda:{self.gpu}")
                          for _ in range(self.nstreams)]
        return self._pool

    def take_slot(self):
        self._slot += 1
        return self._slot - 1

    def footprint(self, req, p=None):
        p = p or plan(req.shape)
        return p.device_bytes + min(self.nring, req.nbuf) * p.buf_bytes, p.host_bytes

    def make_ring(self, req, p):
        """Seeded synthetic input buffers in the device layout
        [Nc][Nb*Ntpb][Np][re,im]: int8 ~ round(N(0, 8^2)) clipped to +-127, int16
        ~ round(N(0, 1000^2)) clipped to +-32767.  Generated on a side stream,
        so admitting a stream never waits for the other streams' work."""
        import torch
        dev = f"cuda:{self.gpu}"
        if self._gen is None:
            self._gen = torch.cuda.Stream(device=dev)
        g = torch.Generator(device=dev)
        g.manual_seed(req.seed)
        n16 = req.shape.Nbps == 16
        sigma, lim = (1000.0, 32767) if n16 else (8.0, 127)
        out = []
        with torch.cuda.stream(self._gen):
            for _ in range(min(self.nring, req.nbuf)):
                v = torch.randn(p.buf_bytes // (2 if n16 else 1), generator=g, device=dev)
                v = v.mul_(sigma).round_().clamp_(-lim, lim)
                out.append(v.to(torch.int16 if n16 else torch.int8))
                del v
        self._gen.synchronize()
        return out

    def run(self, reqs, arrival_scale=1.0):
        """Serves the requests; arrival_scale multiplies the arrival times
        (0: all present at the start).  Returns statistics and, with
        record_dumps, every stream's dumps {sid: [(buffer, product, pwr, ics)]}."""
        pending = collections.deque(sorted(reqs, key=lambda r: r.arrival))
        plans = {r.sid: plan(r.shape) for r in reqs}
        for r in reqs:
            d, h = self.footprint(r, plans[r.sid])
            if d > self.hbm_budget or h > self.host_budget:
                raise ValueError(f"stream {r.sid} needs {d} device / {h} host bytes, "
                                 "more than the budget")
        active, done = [], {}
        used_d = used_h = 0
        peak_d, peak_active = 0, 0
        lat, waits, admit_s = [], [], []
        total_bytes = 0
        t0 = time.perf_counter()
        t_first = None
        while pending or active:
            now = time.perf_counter() - t0
            while (pending and pending[0].arrival * arrival_scale <= now
                   and len(active) < self.max_active):
                r = pending[0]
                d, h = self.footprint(r, plans[r.sid])
                if used_d + d > self.hbm_budget or used_h + h > self.host_budget:
                    break
                pending.popleft()
                ta = time.perf_counter()
                s = self.stream_cls(r, self, plans[r.sid], t0 + r.arrival * arrival_scale, ta)
                admit_s.append(time.perf_counter() - ta)
                waits.append(ta - s.t_arrival)
                s.need = (d, h)
                used_d += d
                used_h += h
                peak_d = max(peak_d, used_d)
                active.append(s)
                peak_active = max(peak_active, len(active))
                if t_first is None:
                    t_first = ta
            moved = False
            for s in list(active):
                moved |= s.pump()
                if s.finished:
                    s.close()
                    active.remove(s)
                    used_d -= s.need[0]
                    used_h -= s.need[1]
                    lat += s.lat
                    total_bytes += s.sent * s.plan.buf_bytes
                    done[s.req.sid] = s.dumps
                    moved = True
            if not moved:
                if not active and pending:
                    idle = pending[0].arrival * arrival_scale - (time.perf_counter() - t0)
                    if idle > 0:
                        time.sleep(min(idle, 0.01))
                else:
                    time.sleep(self.poll_s)
        t_end = time.perf_counter()     # every stream has drained (close() waits)
        span = t_end - (t_first if t_first is not None else t0)
        la = np.array(lat) * 1e3 if lat else np.zeros(1)
        wa = np.array(waits) * 1e3 if waits else np.zeros(1)
        return {
            "streams": len(reqs),
            "ingest_bytes": int(total_bytes),
            "wall_s": span,
            "GBps": total_bytes / span / 1e9 if span > 0 else 0.0,
            "dumps": int(len(lat)),
            "p50_ms": float(np.percentile(la, 50)),
            "p99_ms": float(np.percentile(la, 99)),
            "admit_wait_p50_ms": float(np.percentile(wa, 50)),
            "admit_wait_p99_ms": float(np.percentile(wa, 99)),
            "admit_setup_ms_mean": float(np.mean(admit_s) * 1e3) if admit_s else 0.0,
            "peak_active": peak_active,
            "peak_hbm_bytes": int(peak_d),
            "results": done if self.record_dumps else None,
        }
