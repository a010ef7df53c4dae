"""Pins the CPU oracle (tests only) against golden fixtures, closed forms and
brute force, and shows that the checks catch plausible mistakes."""
import os

import numpy as np
import pytest

import oracle

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_fft.npz"))
SIZES = [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]


def _golden(n, tag):
    return GOLDEN[f"x_{n}"], oracle.DD(GOLDEN[f"{tag}_hi_{n}"], GOLDEN[f"{tag}_lo_{n}"])


def _dd_rel(a: oracle.DD, b: oracle.DD):
    d = a.as_longdouble() - b.as_longdouble()
    return float(np.sqrt(np.sum(np.abs(d) ** 2) / np.sum(np.abs(b.as_longdouble()) ** 2)))


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("tag", ["fwd", "inv"])
def test_dft_and_fft_match_golden(n, tag):
    x, ref = _golden(n, tag)
    inv = tag == "inv"
    if n <= 1024:
        assert _dd_rel(oracle.dft(x, inverse=inv), ref) < 2e-18
    assert _dd_rel(oracle.fft(x, inverse=inv), ref) < 2e-18


@pytest.mark.parametrize("n", [8, 64, 512, 4096])
def test_legacy8_restatement_matches_golden(n):
    x, ref = _golden(n, "fwd")
    assert _dd_rel(oracle.legacy8(x), ref) < 4e-18


@pytest.mark.parametrize("n", [16, 256, 4096])
def test_legacy16_restatement_matches_golden(n):
    x, ref = _golden(n, "fwd")
    assert _dd_rel(oracle.legacy16(x), ref) < 4e-18


@pytest.mark.parametrize("n,radices", [
    (8, [8]), (8, [2, 4]), (16, [16]), (64, [8, 8]), (64, [16, 4]), (128, [16, 8]),
    (128, [2, 64]), (256, [16, 16]), (256, [8, 4, 8]), (512, [2, 16, 16]), (512, [8, 64]),
    (4096, [16, 16, 16]),
])
def test_karatsuba_restatement_matches_golden(n, radices):
    x, ref = _golden(n, "fwd")
    assert _dd_rel(oracle.karatsuba(x, radices), ref) < 1e-17


@pytest.mark.parametrize("n", SIZES)
def test_fft_f64_within_fp64_rounding(n):
    x, ref = _golden(n, "fwd")
    got = oracle.fft_f64(x, threads=2)
    assert np.max(oracle.rel_l2(got, ref, axis=-1)) < 1e-15
    xi, refi = _golden(n, "inv")
    assert np.max(oracle.rel_l2(oracle.fft_f64(xi, inverse=True), refi, axis=-1)) < 1e-15


@pytest.mark.parametrize("n", SIZES)
def test_closed_forms(n):
    # delta -> all ones; constant -> n * delta_0
    delta = np.zeros((1, n), complex); delta[0, 0] = 1
    assert np.array_equal(oracle.fft(delta).hi, np.ones((1, n), complex))
    const = np.full((1, n), 1 + 0j)
    expect = np.zeros((1, n), complex); expect[0, 0] = n
    assert np.max(np.abs(oracle.fft(const).as_longdouble() - expect)) < 1e-15
    # pure tones exp(2 pi i k0 j / n) -> n * delta_{k0}: bin placement and sign
    rng = np.random.default_rng(n)
    for k0 in [1, n // 2 - 1, n - 1, int(rng.integers(n))]:
        j = np.arange(n, dtype=np.longdouble)
        tone = np.exp(2j * np.pi * (k0 * j % n) / n).astype(np.complex128).reshape(1, n)
        spec = oracle.fft(tone).as_longdouble()[0]
        assert int(np.argmax(np.abs(spec))) == k0
        e = np.zeros(n); e[k0] = n
        assert np.max(np.abs(spec - e)) < 1e-12
    # time shift by s multiplies bin k by exp(-2 pi i s k / n)
    x = rng.standard_normal((1, n)) + 1j * rng.standard_normal((1, n))
    s = 3 % n
    X = oracle.fft(x).as_longdouble()[0]
    Xs = oracle.fft(np.roll(x, s, axis=-1)).as_longdouble()[0]
    tw = oracle.twiddles(n).as_longdouble()
    assert np.max(np.abs(Xs - X * tw[(s * np.arange(n)) % n])) < 1e-14 * np.sqrt(n)


@pytest.mark.parametrize("n", [8, 128, 4096])
def test_invariants(n):
    rng = np.random.default_rng(7 + n)
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    X = oracle.fft(x).as_longdouble()
    # Parseval
    lhs = np.sum(np.abs(X) ** 2, axis=-1)
    rhs = n * np.sum(np.abs(x.astype(np.clongdouble)) ** 2, axis=-1)
    assert np.all(np.abs(lhs / rhs - 1) < 1e-16)
    # linearity
    a, b = 0.75 - 0.5j, -1.25 + 2j
    lin = oracle.fft(a * x[0] + b * x[1]).as_longdouble()[0]
    assert np.max(np.abs(lin - (a * X[0] + b * X[1]))) < 1e-13
    # real input -> Hermitian spectrum
    r = rng.standard_normal((1, n)) + 0j
    R = oracle.fft(r).as_longdouble()[0]
    assert np.max(np.abs(R - np.conj(R[(-np.arange(n)) % n]))) < 1e-14
    # forward then inverse returns n * x
    back = oracle.fft(X.astype(np.complex128), inverse=True).as_longdouble()
    assert np.max(np.abs(back / n - x)) < 1e-14


def test_twiddles_exact_special_values():
    t = oracle.twiddles(16).hi
    assert t[0] == 1 and t[4] == -1j and t[8] == -1 and t[12] == 1j
    assert t[2].real == np.sqrt(0.5) and t[2].imag == -np.sqrt(0.5)


@pytest.mark.parametrize("n", [64, 4096])
def test_checks_catch_planted_mistakes(n):
    """Mistakes planted in the oracle's output must fail the golden check."""
    x, ref = _golden(n, "fwd")
    good = oracle.fft(x).hi
    assert np.max(oracle.rel_l2(good, ref, axis=-1)) < 1e-15
    wrong_sign = oracle.fft(x, inverse=True).hi            # conjugated twiddles
    wrong_index = good[:, (-np.arange(n)) % n]             # k -> n-k
    transposed = good.reshape(-1, 8, n // 8).transpose(0, 2, 1).reshape(-1, n)
    dropped = good.copy(); dropped[:, 1] = 0               # a dropped term
    for bad in (wrong_sign, wrong_index, transposed, dropped):
        assert np.max(oracle.rel_l2(bad, ref, axis=-1)) > 1e-3
