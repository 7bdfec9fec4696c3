"""Pins the oracle's FP64 FFT (oracle/rawspec_oracle.c ro_fft), which stands in
for cuFFT C2C (reference rawspec_gpu.cu:1182-1195, 1560-1563):
  * against an O(n^2) direct DFT for every n <= 64 (odd and even),
  * against the committed numpy.fft golden vectors (tests/golden),
  * closed forms: impulse, tone, Parseval, forward/inverse mirror.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fft_vectors.npz")


@pytest.mark.parametrize("n", list(range(1, 65)))
@pytest.mark.parametrize("sign", [-1, +1])
def test_fft_matches_direct_dft(n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a = O.fft(x, sign)
    b = O.dft_naive(x, sign)
    assert np.max(np.abs(a - b)) <= 1e-12 * max(1.0, np.max(np.abs(b)))


def test_direct_dft_definition_small():
    # the direct DFT itself against the definition written out for n = 3
    x = np.array([1 + 2j, -0.5 + 0.25j, 3 - 1j])
    w = np.exp(-2j * np.pi / 3)
    want = np.array([sum(x[m] * w ** (k * m) for m in range(3)) for k in range(3)])
    assert np.allclose(O.dft_naive(x, -1), want, atol=1e-14)


def test_fft_matches_numpy_golden_vectors():
    g = np.load(GOLD, allow_pickle=False)
    sizes = sorted(int(k[2:]) for k in g.files if k.startswith("x_"))
    assert 4096 in sizes and 97 in sizes
    for n in sizes:
        x = g[f"x_{n}"]
        for sign, key in ((-1, "fwd"), (+1, "inv")):
            want = g[f"{key}_{n}"]
            got = O.fft(x, sign)
            err = np.max(np.abs(got - want)) / max(1.0, np.max(np.abs(want)))
            assert err < 1e-13, (n, key, err)


@pytest.mark.parametrize("n", [8, 1024, 1 << 16, 1 << 20])
def test_fft_impulse_and_tone(n):
    # impulse at m0 -> |X[k]| = 1 everywhere with phase exp(-2 pi i k m0/n)
    m0 = 5
    x = np.zeros(n, complex)
    x[m0] = 1.0
    X = O.fft(x, -1)
    k = np.arange(n)
    assert np.max(np.abs(X - np.exp(-2j * np.pi * k * m0 / n))) < 1e-12
    # tone exp(+2 pi i k0 m / n): forward puts n in bin k0, inverse in bin -k0
    k0 = 3
    t = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    F = O.fft(t, -1)
    B = O.fft(t, +1)
    assert abs(F[k0] - n) < 1e-9 * n and np.max(np.abs(np.delete(F, k0))) < 1e-9 * n
    assert abs(B[(-k0) % n] - n) < 1e-9 * n


@pytest.mark.parametrize("n", [1000, 4096, 12345])
def test_fft_parseval_and_mirror(n):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    F = O.fft(x, -1)
    assert abs(np.sum(np.abs(F) ** 2) - n * np.sum(np.abs(x) ** 2)) < 1e-10 * n * n
    # inverse kernel == forward of the conjugate, conjugated
    B = O.fft(x, +1)
    assert np.max(np.abs(B - np.conj(O.fft(np.conj(x), -1)))) < 1e-10 * n
