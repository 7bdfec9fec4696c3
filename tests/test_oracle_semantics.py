"""Pins the oracle's restatement of the spectrometer semantics on the CPU:
sample conversion, complex4 expansion, detection and channel order (the
rawspectest impulse known answers), tone placement, integration counts
(the README's dump counts) and the validation rules.
"""
import numpy as np
import pytest

import oracle as O

A = 64.0 / 127.0


# ---- sample conversion (rawspec_gpu.cu:845-863 texture SNORM read) -------
def test_snorm_scale_and_clamp():
    # rawspectest.c:98-103: int8 127 -> 1.0; int16 127 -> 127/32767
    assert O.snorm8(127) == 1.0
    assert O.snorm8(64) == 64.0 / 127.0
    assert O.snorm8(-127) == -1.0
    assert O.snorm8(-128) == -1.0          # most negative code clamps
    assert O.snorm16(127) == 127.0 / 32767.0
    assert O.snorm16(32767) == 1.0
    assert O.snorm16(-32768) == -1.0
    assert O.snorm8(0) == 0.0


def test_complex4_expansion_all_bytes():
    # rawspec_gpu.cu:342-350: high nibble real, low nibble imaginary, signed
    for b in range(256):
        hi, lo = b >> 4, b & 15
        want = (hi - 16 if hi >= 8 else hi, lo - 16 if lo >= 8 else lo)
        assert O.expand4(b) == want


def test_splitmix64_known_answer():
    # first output of SplitMix64 from state 0 (Vigna's reference generator)
    assert O.splitmix64(0) == 0xE220A8397B1DCDAF


def test_generator_is_counter_based():
    a = O.gen_block(4, 256, 2, 8, 9, 8.0, 127, 3)
    b = O.gen_block(4, 512, 2, 8, 9, 8.0, 127, 1)
    # block 3 of 256 samples == second half of block 1 of 512 samples
    a = a.reshape(4, 256, 4)
    b = b.reshape(4, 512, 4)
    assert np.array_equal(a, b[:, 256:, :])
    x = O.gen_block(1, 1 << 16, 2, 8, 1, 8.0, 127, 0).astype(float)
    assert abs(x.std() - 8.0) < 0.1 and abs(x.mean()) < 0.1


# ---- rawspectest impulse known answers (rawspectest.c:36-169) -------------
def _rawspectest_cfg(nbps):
    blocsize = 92274688
    Nc = 88
    Ntpb = blocsize // (2 * 2 * Nc * nbps // 8)
    c = O.make_cfg(No=4, Np=2, Nc=Nc, Ntpb=Ntpb, Nbps=nbps, Nbc=1,
                   Npolout=(1, 1, 4, -4), Nts=(1 << 20, 8, 1024, 1024),
                   Nas=(1, 1 << 17, 1 << 10, 1 << 10), input_conjugated=1)
    assert O.derive(c) == 0
    return c


@pytest.mark.parametrize("nbps", [8, 16])
def test_rawspectest_impulse_kat(nbps):
    c88 = _rawspectest_cfg(nbps)
    full = 127.0 if nbps == 8 else 32767.0
    assert c88.Nb == (4 if nbps == 8 else 8) and c88.Nbc == 11   # auto batching of 88
    assert list(c88.Nds) == [1, 1, 1, 1] and list(c88.Nis) == [1, 1, 1, 1]
    # same geometry on 2 channels (results per channel do not depend on Nc)
    c = O.make_cfg(No=4, Np=2, Nc=2, Ntpb=c88.Ntpb, Nbps=nbps, Npolout=(1, 1, 4, -4),
                   Nts=(1 << 20, 8, 1024, 1024), Nas=(1, 1 << 17, 1 << 10, 1 << 10),
                   input_conjugated=1)
    assert O.derive(c) == 0 and c.Nb == c88.Nb
    dt = np.int8 if nbps == 8 else np.int16
    blk = [np.zeros(c.Nc * c.Ntpb * 4, dtype=dt) for _ in range(c.Nb)]
    # sample 8 pol 0 re = 127 and sample 9 pol 1 im = 64, block Nb_host-1
    blk[-1][8 * 4 + 0] = 127
    blk[-1][9 * 4 + 3] = 64
    dev = O.device_layout(blk, c.Nc, c.Ntpb, 2)
    st = O.State(c)
    mask, out, _ = st.process(dev, -1)       # rawspectest uses fft_dir = -1
    assert mask == 0b1111
    a = 64.0 / full
    one = (127.0 / full) ** 2
    p0 = out[0].reshape(c.Nc, 1 << 20)
    assert np.allclose(p0[0], one + a * a, rtol=1e-12) and not p0[1:].any()
    p1 = out[1].reshape(c.Nc, 8)
    assert np.allclose(p1[0], one + a * a, rtol=1e-12) and not p1[1:].any()
    j = np.arange(1024)
    s, co = np.sin(2 * np.pi * j / 1024), np.cos(2 * np.pi * j / 1024)
    amp = a * 127.0 / full
    p2 = out[2].reshape(4, c.Nc, 1024)[:, 0]      # +4: XX, YY, Re XY*, Im XY*
    assert np.allclose(p2[0], one) and np.allclose(p2[1], a * a)
    assert np.allclose(p2[2], amp * s, atol=1e-12) and np.allclose(p2[3], -amp * co, atol=1e-12)
    p3 = out[3].reshape(4, c.Nc, 1024)[:, 0]      # -4: I, Q, U, V
    assert np.allclose(p3[0], one + a * a) and np.allclose(p3[1], one - a * a)
    assert np.allclose(p3[2], amp * s, atol=1e-12) and np.allclose(p3[3], -amp * co, atol=1e-12)


def test_conjugation_flips_only_imaginary_cross_term():
    outs = {}
    for conj in (0, 1):
        c = O.make_cfg(No=2, Np=2, Nc=2, Ntpb=4096, Npolout=(4, -4), Nts=(256, 256),
                       Nas=(4, 4), input_conjugated=conj)
        assert O.derive(c) == 0
        blk = O.gen_block(2, 4096, 2, 8, 5, 8.0, 127, 0)
        _, o, _ = O.State(c).process(blk, 1)
        outs[conj] = o
    for i in (0, 1):
        a = outs[0][i].reshape(c.Nds[i], 4, -1)
        b = outs[1][i].reshape(c.Nds[i], 4, -1)
        assert np.array_equal(a[:, :3], b[:, :3]) and np.array_equal(a[:, 3], -b[:, 3])


def test_stokes_identities_and_modes():
    c = O.make_cfg(No=3, Np=2, Nc=3, Ntpb=8192, Npolout=(1, 4, -4), Nts=(512,) * 3,
                   Nas=(1,) * 3)
    assert O.derive(c) == 0
    blk = O.gen_block(3, 8192, 2, 8, 11, 8.0, 127, 0)
    _, o, _ = O.State(c).process(blk, 1)
    n = c.Nds[0]
    P = o[0].reshape(n, 1, 3, 512)[:, 0]
    X = o[1].reshape(n, 4, 3, 512)
    S = o[2].reshape(n, 4, 3, 512)
    assert np.allclose(P, X[:, 0] + X[:, 1]) and np.allclose(P, S[:, 0])
    assert np.allclose(S[:, 1], X[:, 0] - X[:, 1])
    assert np.allclose(S[:, 2], X[:, 2]) and np.allclose(S[:, 3], X[:, 3])
    # Na = 1: I^2 = Q^2 + 4 (U^2 + V^2) (no factor 2 on U, V in rawspec)
    assert np.allclose(S[:, 0] ** 2, S[:, 1] ** 2 + 4 * (S[:, 2] ** 2 + S[:, 3] ** 2), rtol=1e-9)


@pytest.mark.parametrize("fft_dir,expect", [(1, 300), (-1, -300)])
def test_tone_channel_and_power(fft_dir, expect):
    # tone amplitude A at +300/1024 cycles/sample: forward -> bin 300, shown at
    # fine channel (300 + 512) % 1024 = 812; inverse mirrors to bin -300.
    N = 1024
    c = O.make_cfg(No=1, Np=1, Nc=1, Ntpb=N, Npolout=(1,), Nts=(N,), Nas=(1,))
    assert O.derive(c) == 0
    m = np.arange(N)
    z = 20 * np.exp(2j * np.pi * 300 * m / N)
    blk = np.empty(2 * N, np.int8)
    blk[0::2] = np.round(z.real)
    blk[1::2] = np.round(z.imag)
    _, o, _ = O.State(c).process(blk, fft_dir)
    P = o[0]
    j = (expect + N // 2) % N
    assert int(np.argmax(P)) == j
    assert abs(P[j] / (20 * N / 127) ** 2 - 1) < 2e-3      # integer rounding of the tone


def test_white_noise_mean_power():
    # E[P] = Nt * 2 sigma^2 / 127^2 per pol; sigma 8, Nt 1024 -> 8.13
    c = O.make_cfg(No=1, Np=1, Nc=1, Ntpb=1 << 18, Npolout=(1,), Nts=(1024,), Nas=(256,))
    assert O.derive(c) == 0
    blk = O.gen_block(1, 1 << 18, 1, 8, 2, 8.0, 127, 0)
    _, o, _ = O.State(c).process(blk, 1)
    mean = o[0].mean() / 256
    assert abs(mean - 1024 * 2 * 64 / 127 ** 2) < 0.05


# ---- integration schedule (rawspec_gpu.cu:1543,1572; README.md:86-88) ------
def test_readme_dump_counts():
    # CLI defaults -f 1048576,8,1024 -t 51,128,3072 over 1001 buffers of 2^20
    c = O.make_cfg(No=3, Np=2, Nc=44, Ntpb=1 << 20, Npolout=(1, 1, 1),
                   Nts=(1 << 20, 8, 1024), Nas=(51, 128, 3072))
    assert O.derive(c) == 0
    counts = [O.lib().ro_dump_events(c, i, 1001) * c.Nds[i] for i in range(3)]
    assert counts == [19, 1025024, 333]


def test_integration_across_buffers_equals_single_long_buffer():
    # Nis > 1: two buffers of 4096 with Na = 4 spectra of 2048 equal one
    # buffer of 8192 with the same integration.
    blocks = [O.gen_block(2, 4096, 2, 8, 4, 8.0, 127, b) for b in range(2)]
    c1 = O.make_cfg(No=1, Np=2, Nc=2, Ntpb=4096, Npolout=(-4,), Nts=(2048,), Nas=(4,))
    c2 = O.make_cfg(No=1, Np=2, Nc=2, Ntpb=4096, Nb=2, Npolout=(-4,), Nts=(2048,), Nas=(4,))
    assert O.derive(c1) == 0 and O.derive(c2) == 0 and c1.Nis[0] == 2 and c2.Nis[0] == 1
    s1 = O.State(c1)
    m0, _, _ = s1.process(blocks[0], 1)
    m1, o1, _ = s1.process(blocks[1], 1)
    _, o2, _ = O.State(c2).process(O.device_layout(blocks, 2, 4096, 2), 1)
    assert m0 == 0 and m1 == 1
    assert np.allclose(o1[0], o2[0], rtol=1e-12)


def test_ics_weighted_sum():
    w = [0.5, 1.25, 0.75]
    c = O.make_cfg(No=1, Np=2, Nant=3, Nc=6, Ntpb=2048, Npolout=(-4,), Nts=(256,),
                   Nas=(8,), incoherently_sum=1, Naws=3)
    assert O.derive(c) == 0
    blk = O.gen_block(6, 2048, 2, 8, 1, 8.0, 127, 0)
    _, o, ics = O.State(c, weights=w).process(blk, 1)
    p = o[0].reshape(1, 4, 3, 2, 256)        # [d][q][ant][cc][j]
    want = np.einsum("a,dqacj->dqcj", np.array(w), p).reshape(-1)
    assert np.allclose(ics[0], want, rtol=1e-12)


# ---- validation rules (rawspec_gpu.cu:471-641) -----------------------------
BAD = [
    (dict(No=0), "No"), (dict(No=5), "No"), (dict(Np=3), "Np"), (dict(Ntpb=0), "Ntpb"),
    (dict(Nbps=12), "Nbps"), (dict(Nts=(0,)), "Nt0"),
    (dict(No=2, Nts=(1024, 96), Nas=(1, 1), Npolout=(1, 1)), "NtDiv"),
    (dict(Nb=3, Ntpb=1000, Nts=(1024,)), "Nb"),
    (dict(Ntpb=3000, Nts=(1024,)), "NtpbDiv"),
    (dict(Ntpb=768, Nts=(1024,)), "NtmaxDiv"),
    (dict(have_blkbufs=1), "NbHost"), (dict(Nas=(0,)), "Na0"),
    (dict(Ntpb=4096, Nts=(1024,), Nas=(3,)), "NaDiv"),
    (dict(Nc=6, Nbc=4), "Nbc"),
    (dict(incoherently_sum=1, Nant=2, Naws=3, Nc=4), "Naws"),
]


@pytest.mark.parametrize("kw,rule", BAD)
def test_validation_rejects(kw, rule):
    base = dict(No=1, Np=2, Nc=4, Ntpb=4096, Nts=(1024,), Nas=(1,), Npolout=(1,))
    base.update(kw)
    c = O.make_cfg(**base)
    assert O.ERRORS[O.derive(c)] == rule


def test_validation_derivations():
    c = O.make_cfg(No=2, Np=1, Nc=10, Ntpb=1024, Nbps=4, Nbc=1, Npolout=(4, -4),
                   Nts=(4096, 256), Nas=(2, 64))
    assert O.derive(c) == 0
    assert c.Nbps == 8 and c.expand4 == 1            # 4-bit -> expanded 8
    assert list(c.Npolout)[:2] == [1, 1]             # Np = 1 forces total power
    assert c.Nb == 4 and c.Nb_host == 4 and c.Ntmax == 4096
    assert c.Nbc == 1                                # largest factor <= 10 of 10 -> 10/10
    assert c.Nis[0] == 2 and c.Nds[1] == 1 and c.Nis[1] == 4


@pytest.mark.parametrize("nts,pols", [([1024, 8], [-4, 1]), ([1000, 100], [4, 1])])
def test_single_precision_port_tracks_the_oracle(nts, pols):
    """liboracle32.so is the same code in single precision (the FP32 CPU
    port the CPU baseline times): its dumps agree with the FP64 oracle's to
    FP32 rounding, and differ from them (it really computes in FP32)."""
    Nc, Ntpb = 2, 4000 if nts[0] == 1000 else 4096
    cfg = O.make_cfg(No=2, Np=2, Nc=Nc, Ntpb=Ntpb, Nbps=8, Npolout=pols, Nts=nts, Nas=[2, 4])
    assert O.derive(cfg) == 0
    blk = O.gen_block(Nc, Ntpb, 2, 8, 5, 8.0, 127, 0, 1, 20.0, 100.0, float(nts[0]))
    buf = O.device_layout([blk] * cfg.Nb, Nc, Ntpb, 2)
    s64, s32 = O.State(cfg), O.State(cfg, fp32=True)
    _, r64, _ = s64.process(buf, 1)
    _, r32, _ = s32.process(buf, 1)
    s64.close()
    s32.close()
    assert sorted(r64) == sorted(r32) == [0, 1]
    for i in r64:
        a, b = r64[i], r32[i]
        tot = np.abs(a).max()
        assert np.abs(a - b).max() <= 1e-5 * tot, i
        assert not np.array_equal(a, b), i
