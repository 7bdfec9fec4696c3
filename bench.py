#!/usr/bin/env python3
"""RAW ingest throughput of librawspec (gfx950 HIP spectrometer core) on one GPU.

Headline workload ("stokes_64ch_int8_3products"): 1 antenna, 64 coarse
channels, 2 pols, complex int8 samples round(N(0, 8^2)) clipped to +-127
with a complex tone of amplitude 20 at fine bin 300/1024 of channel 3
(pol 0); blocks of 524288 time samples (128 MiB), so Nb = 2 blocks (256 MiB)
per device buffer; three products Nfft = {1048576, 8, 1024} integrating
Nint = {51, 128, 3072} spectra, full Stokes (pols = -4), forward FFT.

One step = one device buffer, with the input already resident in HBM: a
ring of 64 buffers (16 GiB, the whole 128-block run) is generated on the
device before timing and cycled.  Dumps stay in device memory
(rawspec_amd_set_device_output), so PCIe is outside the timed region; the
PCIe-inclusive figure is reported separately (--pcie).  Four paths are
timed: "plain", one buffer per rawspec_amd_start_processing_device() call
(the granularity of the reference's rawspec_start_processing()) on the FP32
vector units; "plain_mfma", the same calls on the matrix-core engine;
"served", batches of 8 consecutive buffers per
rawspec_amd_start_processing_device_batch() call; and "served_mfma", the
same on the matrix-core engine.  `value` is the faster served path (both
keep the plain path's error bar).  paths.served_autotuned is the served
path once more after rawspec_amd_autotune picked launch plans for batches
of 8 (reported beside `value`, not in it).  The library runs this context's
four-step and small products on two streams that overlap (rawspec_amd_set_
lanes; --lanes off for one stream); the per-kernel times, and the roofline
of the dominant kernel, come from a pass with HIP events around every
kernel, in which the products run one after another.  Beside the 8 TB/s
vendor peak, roofline.measured_peaks gives what a plain 1 GiB device copy
and fill reach on the same GPU, and roofline.traffic_GBps the dominant
kernel's measured HBM bytes (committed rocprofv3 counter summary) over its
launch time.

Also measured, reported under "workloads": the other single-stream
workloads (8 channels with 4 MiB buffers and Nfft 1024; the headline geometry
with pols 1 and pols 4; 16 antennas x 16 channels of complex int16 with four
products and the incoherent sum), one buffer per call, 8 per call, and 8 per
call after rawspec_amd_autotune picked launch plans for that mode; under
"serving", the
serving loop (32 Poisson-arriving streams of random shapes, concurrent vs
one at a time); under "file_ingest", the rawspec program reading a GUPPI RAW
file of the headline geometry from the page cache and writing filterbank
files, with the reference program's read-copy-start order and with the
pipelined ingest.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "RAW ingest GB/s on 1 GPU, MFMA/served vs plain FP32, ≤1e-3 rel err; p99 latency"
HBM_PEAK_GBS = 8000.0

C3 = dict(name="stokes_64ch_int8_3products", Nc=64, Np=2, Nbps=8, Ntpb=524288, nblocks=128, Nts=[1 << 20, 8, 1024],
          Nas=[51, 128, 3072], Npolout=[-4, -4, -4], sigma=8.0, tone=(3, 20.0, 300, 1024),
          seed=3, Nant=1)
OTHER = [
    dict(name="power_8ch_int8_nfft1024", Nc=8, Np=2, Nbps=8, Ntpb=131072, nblocks=4, Nts=[1024], Nas=[16],
         Npolout=[1], sigma=8.0, tone=(3, 20.0, 300, 1024), seed=1, Nant=1),
    dict(name="power_64ch_int8_3products", Nc=64, Np=2, Nbps=8, Ntpb=524288, nblocks=128,
         Nts=[1 << 20, 8, 1024], Nas=[51, 128, 3072], Npolout=[1, 1, 1], sigma=8.0, tone=None,
         seed=2, Nant=1),
    dict(name="crosspol_64ch_int8_3products", Nc=64, Np=2, Nbps=8, Ntpb=524288, nblocks=128,
         Nts=[1 << 20, 8, 1024], Nas=[51, 128, 3072], Npolout=[4, 4, 4], sigma=8.0, tone=None,
         seed=2, Nant=1),
    dict(name="power_16ant_16ch_int16_4products_ics", Nc=256, Np=2, Nbps=16, Ntpb=262144, nblocks=64,
         Nts=[64, 4096, 131072, 8], Nas=[2048, 32, 4, 16384], Npolout=[1, 1, 1, 1],
         sigma=1000.0, tone=None, seed=4, Nant=16),
]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def under_profiler():
    """True inside a rocprofv3 (rocprofiler-sdk) run."""
    return ("rocprofiler" in os.environ.get("LD_PRELOAD", "") or
            any(k.startswith("ROCPROF") for k in os.environ))


def weights(W):
    if W["Nant"] <= 1:
        return None
    rng = random.Random(W["seed"])
    return [rng.uniform(0.5, 1.5) for _ in range(W["Nant"])]


def make_sp(R, W):
    aws = weights(W)
    kw = dict(incoherently_sum=1, Aws=aws) if aws else {}
    return R.Spectrometer(Nc=W["Nc"], Ntpb=W["Ntpb"], Nts=W["Nts"], Nas=W["Nas"], Np=W["Np"],
                          Nbps=W["Nbps"], Npolout=W["Npolout"], Nant=W["Nant"],
                          record_events=False, **kw)


def gen_ring(torch, sp, W, nring):
    """Seeded synthetic device buffers [Nc][Nb*Ntpb][Np][re,im] (int8 or
    int16): round(N(0, sigma^2)) plus the optional tone (pol 0, phase
    continuous in the sample index), clipped to the signed range."""
    c = sp.ctx
    S = c.Nb * c.Ntpb
    lim = 32767 if W["Nbps"] == 16 else 127
    ring = []
    g = torch.Generator(device="cuda")
    g.manual_seed(W["seed"])
    t = torch.arange(S, device="cuda", dtype=torch.int64)
    for r in range(nring):
        x = torch.randn((c.Nc, S, c.Np, 2), generator=g, device="cuda", dtype=torch.float32)
        x.mul_(W["sigma"])
        if W["tone"]:
            ch, amp, k, n = W["tone"]
            ph = (((t + r * S) * k) % n).to(torch.float64) * (2.0 * math.pi / n)
            x[ch, :, 0, 0] += (amp * torch.cos(ph)).to(torch.float32)
            x[ch, :, 0, 1] += (amp * torch.sin(ph)).to(torch.float32)
        x = x.round_().clamp_(-lim, lim)
        ring.append(x.to(torch.int16 if W["Nbps"] == 16 else torch.int8).contiguous())
        del x
    torch.cuda.synchronize()
    return ring


def alg_bytes(sp, buf_bytes):
    """Algorithmic bytes per buffer: RAW read once + each product's dump
    bytes (and ICS) amortised over its Nis buffers."""
    c = sp.ctx
    out = 0.0
    per_prod = []
    for i in range(c.No):
        nss = c.Nb * c.Ntpb // c.Nts[i]
        nis = max(1, c.Nas[i] // nss)
        b = c.h_pwrbuf_size[i] * (1 + (1.0 / c.Nant if c.incoherently_sum else 0.0)) / nis
        per_prod.append((b * nis, nis))
        out += b
    return buf_bytes + out, per_prod


def measured_traffic(kernel, batch, engine):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 --pmc summary of the same leg (FETCH_SIZE and WRITE_SIZE runs,
    each scaled by the counter's reading for a 1 GiB copy), or None when no
    summary of that engine and batch size holds the kernel."""
    name = "stokes_64ch_int8_3products_hbm_traffic" + ("_mfma" if engine == "mfma" else "") + ".json"
    try:
        rec = json.load(open(os.path.join(ROOT, "profiles", name)))
    except (OSError, ValueError):
        return None
    k = rec.get("kernels", {}).get(kernel)
    if not k or rec.get("batch") != batch or rec.get("engine") != engine:
        return None
    return k.get("bytes_per_launch")


def measured_peaks(torch, nbytes=1 << 30, reps=5):
    """HBM rates this GPU reaches with plain device kernels, for comparison
    with the 8 TB/s vendor peak: a device-to-device copy of nbytes (read +
    write bytes counted) and a fill (write only), median of reps, timed with
    events on torch's current stream (the one these kernels run on)."""
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").random_(0, 255)
    dst = torch.empty_like(src)

    def rate(fn, moved):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        ts = []
        for _ in range(reps):
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return moved / sorted(ts)[len(ts) // 2] / 1e9

    out = {"copy_GBps": round(rate(lambda: dst.copy_(src), 2 * nbytes), 1),
           "fill_GBps": round(rate(lambda: dst.fill_(7), nbytes), 1),
           "bytes": nbytes, "method": "median of 5 event-timed launches"}
    del src, dst
    torch.cuda.synchronize()
    return out


def submitter(sp, ring, batch):
    """go(k0, n): submits buffers k0 .. k0+n-1 of the ring (n a multiple of
    batch), one call per buffer (batch 1, rawspec_amd_start_processing_device)
    or per `batch` consecutive buffers (rawspec_amd_start_processing_device_batch)."""
    ptrs = [r.data_ptr() for r in ring]

    def go(k0, n):
        for k in range(k0, k0 + n, batch):
            if batch == 1:
                rc = sp.start_processing_device(ptrs[k % len(ptrs)], 1)
            else:
                rc = sp.start_processing_device_batch(
                    [ptrs[(k + j) % len(ptrs)] for j in range(batch)], 1)
            if rc:
                raise RuntimeError("start processing failed")
    return go


def timed(torch, sp, ring, steps, warmup, batch=1):
    go = submitter(sp, ring, batch)
    go(0, warmup)
    if sp.wait_for_completion():
        raise RuntimeError("wait failed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    go(warmup, steps)
    if sp.wait_for_completion():
        raise RuntimeError("wait failed")
    torch.cuda.synchronize()
    return time.perf_counter() - t0, go


def cpu_baseline(sp, ring, nthreads, nbuf, chans, runs=3):
    """The oracle's C code timed on the host: the FP64 build and the same
    algorithm built in single precision, each over nbuf buffers of the
    workload (channels chans), runs times; returns {"fp64": ..., "fp32": ...}
    with the median GB/s and every run's seconds."""
    import oracle as O
    c = sp.ctx
    cfg = O.make_cfg(No=c.No, Np=c.Np, Nc=c.Nc, Ntpb=c.Ntpb, Nbps=8,
                     Npolout=list(c.Npolout)[:c.No], Nts=list(c.Nts)[:c.No],
                     Nas=list(c.Nas)[:c.No])
    assert O.derive(cfg) == 0
    host = [ring[k % len(ring)].cpu().numpy().reshape(-1) for k in range(nbuf)]
    nbytes = len(chans) * c.Nb * c.Ntpb * c.Np * 2 * nbuf
    out = {}
    for name, fp32 in (("fp64", False), ("fp32", True)):
        ts = []
        for _ in range(runs):
            st = O.State(cfg, threads=nthreads, fp32=fp32)
            st.set_channels(chans)
            t0 = time.perf_counter()
            for h in host:
                st.process(h, 1)
            ts.append(time.perf_counter() - t0)
            st.close()
        med = sorted(ts)[len(ts) // 2]
        out[name] = {"GBps": round(nbytes / med / 1e9, 4), "seconds": [round(t, 2) for t in ts]}
    return out


def other_config(torch, R, W, nring=8):
    sp = make_sp(R, W)
    sp.set_device_output(True)
    buf_bytes = sp.device_input_size()
    ring = gen_ring(torch, sp, W, nring)
    nbuf_run = max(1, W["nblocks"] // sp.ctx.Nb)          # the whole run of blocks
    steps = max(nbuf_run, 32)
    dt, _ = timed(torch, sp, ring, steps, warmup=max(4, steps // 8))
    B_alg, _ = alg_bytes(sp, buf_bytes)
    ms = dt / steps * 1e3
    out = {"GBps": round(steps * buf_bytes / dt / 1e9, 2), "ms_per_step": round(ms, 4),
           "step_bytes": buf_bytes, "steps": steps,
           "pipeline_frac_of_hbm": round(B_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # the same buffers 8 per call (rawspec_amd_start_processing_device_batch)
    sp.reset_integration()
    sb = (steps + 7) // 8 * 8
    dt8, _ = timed(torch, sp, ring, sb, warmup=8, batch=8)
    out["served_GBps"] = round(sb * buf_bytes / dt8 / 1e9, 2)
    out["served_ms_per_step"] = round(dt8 / sb * 1e3, 4)
    # the same on the matrix-core engine at its 1e-3 bar (integrations
    # restarted so the leg runs only its own engine), then back to FP32
    assert sp.set_mfma(1) == 0 and sp.reset_integration() == 0
    dtm, _ = timed(torch, sp, ring, sb, warmup=8, batch=8)
    out["served_mfma_GBps"] = round(sb * buf_bytes / dtm / 1e9, 2)
    assert sp.set_mfma(0) == 0 and sp.reset_integration() == 0
    # the launch-plan autotuner for batches of 8, then the batches again
    t1 = time.perf_counter()
    rep = sp.autotune(ring[0].data_ptr(), 8)
    out["autotune_s"] = round(time.perf_counter() - t1, 2)
    dt8t, _ = timed(torch, sp, ring, sb, warmup=8, batch=8)
    out["served_autotuned_GBps"] = round(sb * buf_bytes / dt8t / 1e9, 2)
    out["autotune_plans"] = [ln.split(":")[0] for ln in rep.strip().splitlines()]
    sp.close()
    del ring
    torch.cuda.empty_cache()
    return out


def file_ingest(torch, W, nblocks):
    """The rawspec program on a GUPPI RAW file of W's geometry (nblocks
    blocks, DIRECTIO headers, samples from the seeded device generator, read
    back from the page cache), in three ingest modes; plus the pinned
    host-to-device copy rate, the PCIe bound of any file ingest.  Modes:
    the reference program's order (one reader; read, wait, copy, start),
    the same order with 16 read threads, and the pipelined ingest."""
    import shutil
    import subprocess
    import tempfile

    from rawspec_amd.rawfile import header, observation_cards
    cli = os.path.join(ROOT, "rawspec_amd", "bin", "rawspec")
    if not os.path.exists(cli):
        log("bench: rawspec program not built; no file ingest figures")
        return None
    if under_profiler():
        # A traced process that spawned traced child programs (the rawspec
        # runs below inherit the profiler and write to its output path) has
        # crashed in the profiler's exit handlers; profile the program itself
        # instead (rocprofv3 ... -- rawspec_amd/bin/rawspec ...).
        log("bench: running under a profiler: file ingest leg skipped")
        return None
    d = tempfile.mkdtemp(prefix="rawspec_ingest_")
    out = {"blocks": nblocks, "block_bytes": W["Nc"] * W["Ntpb"] * W["Np"] * 2 * W["Nbps"] // 8,
           "source": "one GUPPI RAW file in the page cache", "output": "SIGPROC filterbank files"}
    try:
        stem = os.path.join(d, "stem")
        g = torch.Generator(device="cuda")
        g.manual_seed(W["seed"])
        lim = 32767 if W["Nbps"] == 16 else 127
        dt = torch.int16 if W["Nbps"] == 16 else torch.int8
        with open(stem + ".0000.raw", "wb") as f:
            for b in range(nblocks):
                cards = observation_cards(obsnchan=W["Nc"], npol=W["Np"], nbits=W["Nbps"],
                                          ntpb=W["Ntpb"], pktidx=32 * b, directio=1)
                f.write(header(cards, 1))
                x = torch.randn((W["Nc"], W["Ntpb"], W["Np"], 2), generator=g, device="cuda")
                f.write(x.mul_(W["sigma"]).round_().clamp_(-lim, lim).to(dt).cpu().numpy().tobytes())
                del x
        dest = os.path.join(d, "out")
        os.mkdir(dest)
        args = ["-f", ",".join(map(str, W["Nts"])), "-t", ",".join(map(str, W["Nas"])),
                "-p", ",".join(map(str, W["Npolout"])), "-d", dest, "--stats"]
        # the stem four times over (integrations restart per stem): 4 x nblocks blocks
        for name, extra in [("reference_order", ["--sync-ingest", "--readers", "1"]),
                            ("sync_parallel_reads", ["--sync-ingest", "--readers", "16"]),
                            ("pipelined", [])]:
            r = subprocess.run([cli, *args, *extra, stem, stem, stem, stem], capture_output=True,
                               text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(f"rawspec failed: {r.stderr[-2000:]}")
            st = json.loads([x for x in r.stderr.splitlines() if x.startswith("{")][-1])
            out[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()}
            log(f"bench: file ingest {name}: {st['GBps']:.2f} GB/s "
                f"({st['buffers']} buffers, read wait {st['read_wait_seconds']:.3f} s of "
                f"{st['seconds']:.3f} s)")
        n = out["block_bytes"] * 2
        h = torch.empty(n, dtype=torch.uint8, pin_memory=True)
        dv = torch.empty(n, dtype=torch.uint8, device="cuda")
        dv.copy_(h, non_blocking=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(8):
            dv.copy_(h, non_blocking=True)
        torch.cuda.synchronize()
        out["h2d_pinned_GBps"] = round(8 * n / (time.perf_counter() - t1) / 1e9, 2)
        out["pipelined_vs_reference_order"] = round(out["pipelined"]["GBps"] /
                                                    out["reference_order"]["GBps"], 2)
        # release the pinned block now rather than at process exit
        torch.cuda.synchronize()
        del h, dv
        torch._C._host_emptyCache()
        torch.cuda.empty_cache()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def serving(torch):
    import rawspec_amd.serve as S
    reqs = S.make_requests(n=32, rate=4.0, seed=5)
    out = {"streams": len(reqs), "arrivals": "Poisson, mean 4/s, seed 5",
           "ingest_bytes": sum(r.nbuf * S.plan(r.shape).buf_bytes for r in reqs)}
    for name, kw, scale in [("concurrent", {}, 1.0), ("sequential", {"max_active": 1}, 1.0),
                            ("concurrent_batched", {"batch": 8}, 1.0),
                            ("concurrent_all_present", {}, 0.0),
                            ("concurrent_batched_all_present", {"batch": 8}, 0.0),
                            ("sequential_all_present", {"max_active": 1}, 0.0)]:
        st = S.Server(**kw).run(reqs, arrival_scale=scale)
        st.pop("results")
        out[name] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()
                     if k not in ("streams", "ingest_bytes")}
        log(f"bench: serving {name}: {st['GBps']:.1f} GB/s, p50 {st['p50_ms']:.2f} ms, "
            f"p99 {st['p99_ms']:.2f} ms, admission wait p99 {st['admit_wait_p99_ms']:.1f} ms")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=512)
    ap.add_argument("--warmup", type=int, default=104)
    ap.add_argument("--batch", type=int, default=8,
                    help="buffers per call on the served path (rawspec_amd_start_processing_device_batch)")
    ap.add_argument("--ring", type=int, default=64, help="device-resident input buffers")
    ap.add_argument("--latency-steps", type=int, default=102)
    ap.add_argument("--cpu-buffers", type=int, default=8, help="buffers per CPU-baseline run")
    ap.add_argument("--cpu-chans", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-others", action="store_true", help="headline workload only")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving loop")
    ap.add_argument("--ingest-blocks", type=int, default=32,
                    help="blocks of the GUPPI RAW file the rawspec program ingests (0: skip)")
    ap.add_argument("--pcie", action="store_true", help="also time the host-buffer API path")
    ap.add_argument("--legs", default="plain,plain_mfma,served,served_mfma,served_autotuned",
                    help="paths to time (comma-separated subset of the default)")
    ap.add_argument("--lanes", choices=["auto", "on", "off"], default="auto",
                    help="two-lane execution of the headline context (auto: the library's "
                         "choice); per-kernel times always come from a pass without lanes")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run a plain 1 GiB device copy 4 times (calibrates the HBM "
                         "byte counters of a rocprofv3 --pmc run of this command)")
    a = ap.parse_args(argv)

    a.steps = max(a.batch, (a.steps + a.batch - 1) // a.batch * a.batch)
    a.warmup = (a.warmup + a.batch - 1) // a.batch * a.batch
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        return 2

    import torch
    import rawspec_amd as R

    W = C3
    log(f"bench: workload {W['name']} Nc={W['Nc']} Ntpb={W['Ntpb']} Nts={W['Nts']} Nas={W['Nas']} "
        f"pols={W['Npolout']}")
    sp = make_sp(R, W)
    c = sp.ctx
    buf_bytes = sp.device_input_size()
    assert buf_bytes == c.Nc * c.Nb * c.Ntpb * c.Np * 2
    sp.set_device_output(True)
    if a.lanes != "auto":
        assert sp.set_lanes(a.lanes == "on") == 0
    lanes = sp.get_lanes()
    ring = gen_ring(torch, sp, W, a.ring)
    log(f"bench: ring of {len(ring)} device buffers x {buf_bytes / 2**20:.0f} MiB ready; "
        f"lanes {'on' if lanes else 'off'}")

    def measure(batch, mfma):
        """Throughput of one leg, then its per-kernel HIP-event timing.  The
        engine applies from each product's next integration start, so every
        integration is restarted first: the leg times only its own engine."""
        assert sp.set_mfma(mfma) == 0
        assert sp.reset_integration() == 0
        dt, go = timed(torch, sp, ring, a.steps, a.warmup, batch=batch)
        sp.profile(True)
        nprof = max(batch, (max(51, min(a.steps, 152)) // batch) * batch)
        go(0, nprof)
        prof = sp.profile_report()
        sp.profile(False)
        kern = {k: {"launches": n, "avg_ms": ms / n, "total_ms": ms, "ms_per_step": ms / nprof}
                for k, (n, ms) in prof.items()}
        if mfma:
            # the matrix-core engine at its 1e-3 bar runs the 1024-point
            # product (3072 spectra per dump) on unsplit FP16 operands, with
            # the 8-point product (128 spectra per dump) computed in FP32 by
            # the same kernel from the same read of the samples, and keeps
            # the 2^20 product's intermediate in 24-bit floats; both
            # four-step passes stay on the vector units, where they are
            # faster
            names = {k.split(".")[1] for k in kern if k.endswith(".N1024")}
            assert names == {"k_small_mfma_u8", "k_reduce"} or names == {"k_small_mfma_u8"}, \
                f"matrix-core leg ran {sorted(names)} for the 1024-point product"
            assert not any(k.endswith(".N8") for k in kern), \
                f"matrix-core leg ran {sorted(kern)}: the 8-point product has a kernel of its own"
            assert "p0.k_fs_a.N1048576" in kern and "p0.k_fs_b.N1048576" in kern, \
                f"matrix-core leg ran {sorted(kern)}"
        gbs = a.steps * buf_bytes / dt / 1e9
        log(f"bench: batch {batch} mfma {int(mfma)}: {a.steps} steps in {dt:.3f} s -> {gbs:.1f} GB/s "
            f"({dt / a.steps * 1e3:.3f} ms/step); " +
            ", ".join(f"{k}={v['ms_per_step']:.3f}ms/step" for k, v in sorted(kern.items())))
        return {"GBps": gbs, "ms_per_step": dt / a.steps * 1e3, "kernels": kern, "nprof": nprof,
                "go": go}

    if a.calib_copy:
        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda").random_(0, 255)
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    want = [x for x in a.legs.split(",") if x]
    spec = {"plain": (1, False), "plain_mfma": (1, True), "served": (a.batch, False),
            "served_mfma": (a.batch, True)}
    legs = {k: measure(*spec[k]) for k in spec if k in want}
    if "plain" not in legs:
        legs["plain"] = legs[want[0]]
    assert sp.set_mfma(False) == 0 and sp.reset_integration() == 0
    head_name = max([k for k in ("served", "served_mfma") if k in legs] or ["plain"],
                    key=lambda k: legs[k]["GBps"])
    head = legs[head_name]
    value, ms_per_step = head["GBps"], head["ms_per_step"]
    kernels = head["kernels"]

    # roofline of the headline leg's dominant kernel: algorithmic bytes per
    # launch = its share of the RAW bytes read once and of the dump bytes
    # written once per integration, split evenly over its launches per step
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    B_alg, per_prod = alg_bytes(sp, buf_bytes)
    prod = int(dom.split(".")[0][1:])
    kname = dom.split(".")[1]
    out_b, nis = per_prod[prod]                     # dump bytes of that product, buffers per dump
    if kname.startswith("k_small"):
        kb_step = buf_bytes + out_b / nis
    elif kname.startswith("k_fs_a"):
        kb_step = buf_bytes
    else:
        kb_step = out_b / nis
    kb = kb_step / (kernels[dom]["launches"] / head["nprof"])
    dom_ms = kernels[dom]["avg_ms"]
    achieved = kb / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(dom, a.batch if head_name != "plain" else 1,
                                            "mfma" if head_name == "served_mfma" else "vector"),
                "kernel": dom, "alg_bytes_per_launch": kb, "avg_launch_ms": dom_ms}
    peaks = measured_peaks(torch)
    roofline["measured_peaks"] = peaks
    if roofline["traffic"]:
        # the dominant kernel's measured HBM bytes per launch over its launch time
        t_rate = roofline["traffic"] / (dom_ms * 1e-3) / 1e9
        roofline["traffic_GBps"] = round(t_rate, 1)
        roofline["traffic_frac_of_copy"] = round(t_rate / peaks["copy_GBps"], 3)
    log(f"bench: device copy {peaks['copy_GBps']} GB/s, fill {peaks['fill_GBps']} GB/s")
    pipeline = {"alg_bytes_per_step": B_alg, "achieved_GBs": round(B_alg / (ms_per_step * 1e-3) / 1e9, 1),
                "frac": round(B_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"bench: headline {head_name}; dominant kernel {dom} avg {dom_ms:.3f} ms")

    # synchronous latency: submit -> every product complete, one buffer
    # (plain) and one batch (served)
    def latency(go, batch):
        lat = []
        for k in range(0, a.latency_steps * batch, batch):
            t1 = time.perf_counter()
            go(k, batch)
            sp.wait_for_completion()
            lat.append((time.perf_counter() - t1) * 1e3)
        lat.sort()
        return {"p50": round(lat[len(lat) // 2], 3),
                "p99": round(lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))], 3)}

    lat_plain = latency(submitter(sp, ring, 1), 1)
    assert sp.set_mfma(head_name == "served_mfma") == 0 and sp.reset_integration() == 0
    lat_served = latency(submitter(sp, ring, a.batch), a.batch)
    assert sp.set_mfma(False) == 0 and sp.reset_integration() == 0
    p50, p99 = lat_plain["p50"], lat_plain["p99"]

    # the served path after rawspec_amd_autotune picked launch plans for
    # batches of this size (last: the plans stay with the context)
    tuned = None
    if "served_autotuned" in want:
        sp.reset_integration()
        t1 = time.perf_counter()
        rep = sp.autotune(ring[0].data_ptr(), a.batch)
        tune_s = time.perf_counter() - t1
        dtt, _ = timed(torch, sp, ring, a.steps, a.warmup, batch=a.batch)
        tuned = {"GBps": round(a.steps * buf_bytes / dtt / 1e9, 2),
                 "ms_per_step": round(dtt / a.steps * 1e3, 4),
                 "vs_plain": round(a.steps * buf_bytes / dtt / 1e9 / legs["plain"]["GBps"], 3),
                 "autotune_s": round(tune_s, 2),
                 "plans": [ln.split(":")[0] for ln in rep.strip().splitlines()]}
        log(f"bench: served after autotune ({tune_s:.2f} s): {tuned['GBps']} GB/s; {tuned['plans']}")

    extra = {}
    if a.pcie:
        # host-buffer API path: copy blocks from pinned host memory, dumps to h_pwrbuf
        sp.set_device_output(False)
        hb = ring[0].cpu().numpy().reshape(c.Nc, c.Nb, -1)
        for b in range(c.Nb):
            sp.block(b)[:] = hb[:, b, :].reshape(-1).view("uint8")
        n = 51
        t1 = time.perf_counter()
        for k in range(n):
            sp.copy_and_start(c.Nb, fft_dir=1)
            sp.wait_for_completion()
        extra["pcie_inclusive_GBs"] = round(n * buf_bytes / (time.perf_counter() - t1) / 1e9, 2)
        sp.set_device_output(True)

    cpu = None
    if not a.no_cpu:
        nth = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        chans = (list(range(c.Nc)) if a.cpu_chans >= c.Nc else
                 sorted({0, 3, c.Nc - 1} | set(range(5, c.Nc, max(1, c.Nc // a.cpu_chans))))[:a.cpu_chans])
        cb = cpu_baseline(sp, ring, nth, a.cpu_buffers, chans)
        best = max(cb, key=lambda k: cb[k]["GBps"])
        cpu = {"value": cb[best]["GBps"], "unit": "GB/s", "cores": nth, "kind": "port",
               "sample": f"the oracle's C code (OpenMP) in FP64 and built in FP32, {len(chans)} of "
                         f"{c.Nc} channels x {a.cpu_buffers} buffers of this workload, all 3 "
                         f"products, median of 3 runs each; value: the faster ({best})",
               "variants": cb}
        log(f"bench: cpu baseline fp64 {cb['fp64']['GBps']:.4f} GB/s {cb['fp64']['seconds']} s, "
            f"fp32 {cb['fp32']['GBps']:.4f} GB/s {cb['fp32']['seconds']} s on {nth} threads")

    sp.close()
    del ring
    torch.cuda.empty_cache()

    configs = {}
    if not a.no_others:
        for Wo in OTHER:
            configs[Wo["name"]] = other_config(torch, R, Wo)
            log(f"bench: {Wo['name']}: {configs[Wo['name']]}")
    ingest = file_ingest(torch, W, a.ingest_blocks) if a.ingest_blocks > 0 else None
    serve = None if a.no_serve else serving(torch)

    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        # level 1: FP16 matrix products (FP32 accumulate) for the 1024-point
        # product, FP32 vector transforms elsewhere, the 2^20 product's
        # intermediate in 24-bit floats
        "dtype": ("fp16-mfma (fp32 accumulate) + fp32; 24-bit four-step intermediate"
                  if head_name == "served_mfma" else "fp32"),
        "data": "synthetic",
        "config": {"workload": "stokes_64ch_int8_3products: 64 ch x 2 pol complex int8 (sigma 8, tone ch3 bin 300/1024), "
                               "blocks of 524288 samples, Nb=2 (256 MiB/step), "
                               "Nfft {1048576,8,1024} x Nint {51,128,3072}, pols -4",
                   "path": f"{head_name}: batches of {a.batch} consecutive buffers per call, "
                           f"{'matrix-core' if head_name == 'served_mfma' else 'FP32 vector'} engine",
                   "step_bytes": buf_bytes, "ring_buffers": a.ring, "input": "HBM-resident",
                   "dumps": "device-resident",
                   "lanes": "four-step and small products on two streams" if lanes else "off",
                   "kernel_times": "HIP events from a pass with the products in sequence"},
        "paths": dict({k: {"GBps": round(v["GBps"], 2), "ms_per_step": round(v["ms_per_step"], 4),
                           "vs_plain": round(v["GBps"] / legs["plain"]["GBps"], 3)}
                       for k, v in legs.items()},
                      **({"served_autotuned": tuned} if tuned else {})),
        "roofline": roofline,
        "pipeline_roofline": pipeline,
        "kernels": kernels,
        "kernels_plain": legs["plain"]["kernels"],
        "latency_ms": {"plain": dict(lat_plain, mode="submit->complete, one buffer"),
                       "served": dict(lat_served, mode=f"submit->complete, one batch of {a.batch}")},
        "cpu_baseline": cpu,
        "workloads": configs,
        "file_ingest": ingest,
        "serving": serve,
    }
    res.update(extra)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
