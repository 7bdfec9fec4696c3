"""Shared helpers: run one configuration through librawspec (GPU, via the C
ABI) and through the FP64 oracle on identical seeded input, and compare.

Error metric (same for every test): for each dump d and coarse channel c,
let tot[j] be the total-power spectrum of that (d, c) (P for pols=1,
XX+YY for +4, I for -4).  The error of component q at fine channel j is
    |gpu[q][j] - ref[q][j]| / max(tot[j], mean_j tot[j])
i.e. a relative error against the bin's own power, floored at the channel's
mean power so that Q, U, V and cross terms (which cross zero) and bins of
near-zero power are judged against the power that carries them.
"""
from __future__ import annotations

import numpy as np

import oracle as O

PLAIN_TOL = 1e-5      # plain FP32 path vs FP64 oracle (max over all bins)
MFMA_BAR = 1e-3       # the matrix-core engine's own bar (rawspec_amd_set_mfma level 1)
# rawspec_amd_set_mfma levels: matrix cores at the 1e-3 bar (unsplit FP16
# operands for int8 products integrating >= 32 spectra), and split operands
# everywhere (FP32-level error)
MFMA_FAST = 1
MFMA_SPLIT = 2
# Four-step products (N >= 8192) are held to max(PLAIN_TOL, k x the error of
# an FP32 FFT on the same input) -- a strong tone puts FP32 rounding spurs
# above 1e-5 in any single-precision FFT of that length -- and never above
# this fixed ceiling, so an error that grows with the input's FP32 floor
# still fails.  The ceiling is about twice the largest error of one
# full-length FP32 FFT at 2^20 with one spectrum per dump (4.8e-5,
# test_tolerance_basis.py).
FOURSTEP_CEIL = 1e-4
# k for the FP32 engines.  The reference is one full-length FP32 FFT; both
# errors are maxima over all bins, and a four-step FFT in FP32, even with
# exact twiddles and scipy's own FP32 transforms for both passes, measures
# a median 1.0 but up to about 4.4 x the reference's error on the same data
# (test_tolerance_basis.py; the GPU's four-step products: median 0.98, at
# most 4.1 over 1722 random cases), so 5.
PLAIN_FLOOR_MULT = 5.0


def gen_blocks(Nc, Ntpb, Np, Nbps, nblocks, seed=1, sigma=8.0, tone=None, first_blk=0):
    """tone = (chan, amp, bin, nfft) or None."""
    clip = 32767 if Nbps == 16 else 127
    tc, ta, tb, tn = tone if tone else (-1, 0.0, 0.0, 1.0)
    return [O.gen_block(Nc, Ntpb, Np, Nbps, seed, sigma, clip, first_blk + b, tc, ta, tb, tn)
            for b in range(nblocks)]


def rel_err(gpu: np.ndarray, ref: np.ndarray, nds: int, q: int, nc: int, nt: int,
            mode: int) -> np.ndarray:
    g = np.asarray(gpu, dtype=np.float64).reshape(nds, q, nc, nt)
    r = np.asarray(ref, dtype=np.float64).reshape(nds, q, nc, nt)
    if mode == 1:
        tot = r[:, 0]
    elif mode == 4:
        tot = r[:, 0] + r[:, 1]
    else:
        tot = r[:, 0]
    floor = tot.mean(axis=-1, keepdims=True)
    den = np.maximum(np.maximum(tot, floor), 1e-30)
    return np.abs(g - r) / den[:, None]


FP32_FLOOR_NT = 8192   # four-step lengths: judged against the FP32 spur floor


class Fp32Ref:
    """Single-precision reference pipeline (scipy.fft on complex64, float32
    detection and integration) for one product.  It measures the error any
    FP32 FFT makes on the same input; for long transforms with a strong tone
    that floor exceeds 1e-5 (rounding spurs of the tone), so four-step
    products are held to max(PLAIN_TOL, PLAIN_FLOOR_MULT x this floor)."""

    def __init__(self, cfg, i):
        self.cfg, self.i = cfg, i
        self.acc = np.zeros(cfg.pwr_len[i], dtype=np.float32)
        self.count = 0

    def process(self, dev: np.ndarray, fft_dir: int):
        import scipy.fft as sf
        c, i = self.cfg, self.i
        nt, nss, na = c.Nts[i], c.Nss[i], c.Nas[i]
        q = abs(c.Npolout[i])
        nab = min(na, nss)
        full = 32767.0 if c.Nbps == 16 else 127.0
        v = dev.reshape(c.Nc, nss, nt, c.Np, 2).astype(np.float32)
        v = np.maximum(v, -full)
        x = (v[..., 0] + 1j * v[..., 1]).astype(np.complex64)
        X = sf.fft(x, axis=2) if fft_dir > 0 else sf.ifft(x, axis=2) * np.float32(nt)
        X = np.fft.fftshift(X.astype(np.complex64), axes=2)
        s = np.float32(1.0 / (full * full))
        g = np.float32(-1.0 if c.input_conjugated else 1.0)
        X0 = X[..., 0]
        p0 = (X0.real * X0.real + X0.imag * X0.imag) * s
        if q == 1:
            comps = [p0 + ((X[..., 1].real ** 2 + X[..., 1].imag ** 2) * s if c.Np == 2 else 0)]
        else:
            X1 = X[..., 1]
            p1 = (X1.real * X1.real + X1.imag * X1.imag) * s
            re = (X0.real * X1.real + X0.imag * X1.imag) * s
            im = g * (X0.imag * X1.real - X0.real * X1.imag) * s
            comps = [p0, p1, re, im] if c.Npolout[i] == 4 else [p0 + p1, p0 - p1, re, im]
        P = np.stack(comps, axis=0).astype(np.float32)          # [q][Nc][nss][nt]
        P = P.reshape(q, c.Nc, nss // nab, nab, nt).sum(axis=3, dtype=np.float32)
        P = P.transpose(2, 0, 1, 3).reshape(-1)                 # [d][q][Nc][nt]
        self.acc += P
        self.count += 1
        if self.count % c.Nis[i] == 0:
            out = self.acc.copy()
            self.acc[:] = 0
            return out
        return None


def run_case(*, Nc, Ntpb, Nts, Nas, Np=2, Nbps=8, Npolout=None, nbuf=1, Nb=0,
             fft_dir=1, conj=0, Nant=1, ics=False, Aws=None, seed=1, sigma=8.0,
             tone=None, expand4=False, Nbc=0, mfma=False, mfma_switch=None):
    """Runs nbuf buffers through GPU and oracle.  Returns a list of
    (buffer, product, gpu_pwr, ref_pwr, gpu_ics, ref_ics) per dump event and
    the Spectrometer (closed).  mfma: the matrix-core engine level
    (MFMA_FAST, MFMA_SPLIT; 0 / False: vector units); mfma_switch = (buffer,
    level) changes the engine before that buffer."""
    import rawspec_amd as R

    No = len(Nts)
    Npolout = list(Npolout) if Npolout is not None else [1] * No
    sp = R.Spectrometer(Nc=Nc, Ntpb=Ntpb, Nts=Nts, Nas=Nas, Np=Np,
                        Nbps=4 if expand4 else Nbps, Npolout=Npolout, Nant=Nant,
                        Nbc=Nbc, Nb=Nb, input_conjugated=conj,
                        incoherently_sum=1 if ics else 0, Aws=Aws)
    ctx = sp.ctx
    if mfma:
        assert sp.set_mfma(int(mfma)) == 0 and sp.get_mfma() == int(mfma)
    cfg = O.make_cfg(No=No, Np=Np, Nant=Nant, Nc=Nc, Ntpb=Ntpb,
                     Nbps=4 if expand4 else Nbps, Nbc=Nbc, Nb=Nb, Npolout=Npolout,
                     Nts=Nts, Nas=Nas, input_conjugated=conj,
                     incoherently_sum=1 if ics else 0,
                     Naws=len(Aws) if Aws is not None else 0)
    assert O.derive(cfg) == 0
    assert cfg.Nb == ctx.Nb and list(cfg.Nds)[:No] == list(ctx.Nds)[:No]
    st = O.State(cfg, weights=Aws)
    f32 = {i: Fp32Ref(cfg, i) for i in range(No) if Nts[i] >= FP32_FLOOR_NT}
    Nb_ = ctx.Nb
    dump_bits = 8 if expand4 else Nbps
    results = []
    for n in range(nbuf):
        if mfma_switch is not None and n == mfma_switch[0]:
            assert sp.set_mfma(mfma_switch[1]) == 0
        blocks = gen_blocks(Nc, Ntpb, Np, dump_bits, Nb_, seed=seed, sigma=sigma if not expand4 else 3.0,
                            tone=tone, first_blk=n * Nb_)
        if expand4:
            # complex4 host blocks: clip to the 4-bit range and pack nibbles
            packed = []
            for b in blocks:
                v = np.clip(b.astype(np.int16), -8, 7)
                re, im = v[0::2], v[1::2]
                byte = ((re.astype(np.int16) & 0xF) << 4) | (im.astype(np.int16) & 0xF)
                packed.append(byte.astype(np.uint8))
                b[:] = v.astype(b.dtype)      # what the expansion must reproduce
            for bi, pb in enumerate(packed):
                sp.block(bi)[: pb.size] = pb
        else:
            for bi, b in enumerate(blocks):
                sp.block(bi)[:] = b.view(np.uint8)
        for i in range(No):
            sp.pwrbuf(i)[:] = np.frombuffer(b"\x55\x55\x55\x55", dtype=np.float32)[0]
        ev0 = len(sp.events)
        rc = sp.copy_and_start(Nb_, expand4=expand4, fft_dir=fft_dir)
        assert rc == 0
        assert sp.wait_for_completion() == 0
        dev = O.device_layout(blocks, Nc, Ntpb, Np)
        mask, ref, refics = st.process(dev, fft_dir)
        f32out = {i: r.process(dev, fft_dir) for i, r in f32.items()}
        posts = [p for (p, t) in sp.events[ev0:] if t == 1]
        assert sorted(posts) == sorted(ref.keys()), (posts, list(ref.keys()))
        for i in ref:
            g = sp.pwrbuf(i).copy()
            gi = sp.icsbuf(i).copy() if ics else None
            results.append((n, i, g, ref[i], gi, refics.get(i), f32out.get(i)))
    events = list(sp.events)
    sp.close()
    st.close()
    return results, cfg, events


# Four-step products on the matrix cores: split-FP16 operands carry 22
# significant bits against FP32's 24, so with a strong tone their rounding
# spurs reach about 1.3 x those of the FP32 engine (2.0 against 1.5 x the
# FP32 reference at 2^20; still far below the 1e-3 bar of the matrix-core
# engine): the FP32 engine's factor, one more.
MFMA_FLOOR_MULT = PLAIN_FLOOR_MULT + 1.0


def bin_rel_err(gpu: np.ndarray, ref: np.ndarray) -> float:
    """max |gpu - ref| / ref over the bins of a non-negative component (the
    per-bin metric without the channel-mean floor), for reporting."""
    g = np.asarray(gpu, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    return float((np.abs(g - r) / np.maximum(r, 1e-30)).max())


def nonneg_components(mode: int):
    """Components that are powers (never negative): P; XX, YY; I."""
    return [0] if mode in (1, -4) else [0, 1]


def check_results(results, cfg, tol=PLAIN_TOL, ics=False, floor_mult=PLAIN_FLOOR_MULT, ceil=FOURSTEP_CEIL,
                  tol_by_product=None):
    """Asserts every dump is within tol (four-step products: within
    min(ceil, max(tol, floor_mult x the FP32 reference's error))), and with
    ics its incoherent sum within the same bound; returns max error per
    product.  tol_by_product {product: bar} overrides tol (and the four-step
    rule) for the products it names."""
    worst = {}
    for (n, i, g, r, gi, ri, f32) in results:
        q = abs(cfg.Npolout[i])
        assert not np.any(g.view(np.uint32) == 0x55555555), f"unwritten output, product {i}"
        e = rel_err(g, r, cfg.Nds[i], q, cfg.Nc, cfg.Nts[i], cfg.Npolout[i])
        worst[i] = max(worst.get(i, 0.0), float(e.max()))
        if tol_by_product and i in tol_by_product:
            t = tol_by_product[i]
        else:
            t = tol
            if f32 is not None:
                floor = float(rel_err(f32, r, cfg.Nds[i], q, cfg.Nc, cfg.Nts[i], cfg.Npolout[i]).max())
                t = min(ceil, max(tol, floor_mult * floor))
                print(f"  product {i}: gpu {e.max():.2e}  fp32-reference floor {floor:.2e}  bound {t:.2e}")
        assert e.max() <= t, f"product {i} buffer {n}: max rel err {e.max():.3e} > {t}"
        if ics:
            # the incoherent sum is a positive-weighted sum of the antennas'
            # spectra: held to its product's bound
            ncpa = cfg.Nc // cfg.Nant
            ei = rel_err(gi, ri, cfg.Nds[i], q, ncpa, cfg.Nts[i], cfg.Npolout[i])
            assert ei.max() <= t, f"ICS product {i}: max rel err {ei.max():.3e} > {t}"
    return worst


def run_device_case(*, Nc, Ntpb, Nts, Nas, Np=2, Nbps=8, Npolout=None, nbuf=1, batches=None,
                    fft_dir=1, conj=0, Nant=1, ics=False, Aws=None, seed=1, sigma=8.0,
                    tone=None, mfma=False, lanes=None, events=None):
    """Runs nbuf device-resident buffers through rawspec_amd_start_processing_device
    (batches None) or rawspec_amd_start_processing_device_batch (batches: list
    of batch sizes summing to nbuf), dumps copied to the host, and the same
    buffers through the oracle.  Returns (results, cfg) with results as in
    run_case: (buffer, product, gpu, ref, gpu_ics, ref_ics, fp32_ref) per dump,
    matched per product in dump order.  lanes: force two-lane execution on or
    off (None: the library's choice); events: a list that receives every
    (product, callback type) in call order."""
    import torch
    import rawspec_amd as R

    No = len(Nts)
    Npolout = list(Npolout) if Npolout is not None else [1] * No
    got = {i: [] for i in range(No)}

    def on_dump(sp, product, cbtype):
        if events is not None:
            events.append((product, cbtype))
        if cbtype == R.RAWSPEC_CALLBACK_POST_DUMP:
            got[product].append((sp.pwrbuf(product).copy(),
                                 sp.icsbuf(product).copy() if ics else None))

    sp = R.Spectrometer(Nc=Nc, Ntpb=Ntpb, Nts=Nts, Nas=Nas, Np=Np, Nbps=Nbps, Npolout=Npolout,
                        Nant=Nant, input_conjugated=conj, incoherently_sum=1 if ics else 0,
                        Aws=Aws, callback=on_dump)
    if mfma:
        assert sp.set_mfma(int(mfma)) == 0
    if lanes is not None:
        assert sp.set_lanes(lanes) == 0 and sp.get_lanes() == int(bool(lanes))
    cfg = O.make_cfg(No=No, Np=Np, Nant=Nant, Nc=Nc, Ntpb=Ntpb, Nbps=Nbps, Npolout=Npolout,
                     Nts=Nts, Nas=Nas, input_conjugated=conj,
                     incoherently_sum=1 if ics else 0, Naws=len(Aws) if Aws is not None else 0)
    assert O.derive(cfg) == 0
    Nb_ = sp.ctx.Nb
    host = []
    for n in range(nbuf):
        blocks = gen_blocks(Nc, Ntpb, Np, Nbps, Nb_, seed=seed, sigma=sigma, tone=tone,
                            first_blk=n * Nb_)
        host.append(O.device_layout(blocks, Nc, Ntpb, Np))
    dev = [torch.from_numpy(h.view(np.uint8).copy()).to("cuda") for h in host]
    torch.cuda.synchronize()
    if batches is None:
        for d in dev:
            assert sp.start_processing_device(d.data_ptr(), fft_dir) == 0
            assert sp.wait_for_completion() == 0
    else:
        assert sum(batches) == nbuf
        k = 0
        for b in batches:
            assert sp.start_processing_device_batch([d.data_ptr() for d in dev[k:k + b]],
                                                    fft_dir) == 0
            k += b
        assert sp.wait_for_completion() == 0
    sp.close()
    st = O.State(cfg, weights=Aws)
    f32 = {i: Fp32Ref(cfg, i) for i in range(No) if Nts[i] >= FP32_FLOOR_NT}
    seen = {i: 0 for i in range(No)}
    results = []
    for n, h in enumerate(host):
        _, ref, refics = st.process(h, fft_dir)
        f32out = {i: r.process(h, fft_dir) for i, r in f32.items()}
        for i in ref:
            g, gi = got[i][seen[i]]
            seen[i] += 1
            results.append((n, i, g, ref[i], gi, refics.get(i), f32out.get(i)))
    st.close()
    assert all(seen[i] == len(got[i]) for i in range(No)), (seen, {i: len(v) for i, v in got.items()})
    del dev
    return results, cfg


def run_calls(pattern, *, Nc, Ntpb, Nts, Nas, Npolout, held=None, mfma=False, tone=None,
                fft_dir=1, conj=0, seed=1):
    """pattern: a list of calls, each k > 0 (k buffers: 1 through
    rawspec_amd_start_processing_device, more through the batch call),
    "reset" (rawspec_reset_integration) or "autotune" / "autotune:<k>"
    (rawspec_amd_autotune on the next buffer for k buffers per call, which
    restarts the integrations the same way).
    held: rawspec_amd_set_held_buffers before the first call.  Returns
    (results as in run_case, cfg, the GPU's dumps per product in order,
    the autotune reports)."""
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
        assert sp.set_mfma(int(mfma)) == 0
    cfg = O.make_cfg(No=No, Np=2, Nant=1, Nc=Nc, Ntpb=Ntpb, Nbps=8, Npolout=Npolout, Nts=Nts,
                     Nas=Nas, input_conjugated=conj)
    assert O.derive(cfg) == 0
    nbuf = sum(k for k in pattern if isinstance(k, int))
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
    reports = []
    n = 0
    for k in pattern:
        if isinstance(k, str):
            if k == "reset":
                assert sp.reset_integration() == 0
            else:
                per_call = int(k.split(":")[1]) if ":" in k else 1
                reports.append(sp.autotune(dev[min(n, nbuf - 1)].data_ptr(), per_call))
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
    return results, cfg, got, reports


def unsplit_products(Nts, Nas, Np=2, Nbps=8, mfma=MFMA_FAST):
    """Products the engine at level MFMA_FAST runs below FP32 precision, held
    to MFMA_BAR (every other product to the plain rules): 1024-point int8
    products integrating >= 32 spectra (unsplit FP16 operands) and, with two
    pols, four-step products (24-bit intermediate)."""
    if mfma != MFMA_FAST or Np != 2:
        return {}
    return {i: MFMA_BAR for i, (n, a) in enumerate(zip(Nts, Nas))
            if (n == 1024 and a >= 32 and Nbps == 8) or n >= FP32_FLOOR_NT}
