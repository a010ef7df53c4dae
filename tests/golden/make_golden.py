"""Writes tests/golden/golden_fft.npz: inputs and extended-precision DFT outputs.

Expected outputs come from scipy.fft on complex256 (x87 long double, pocketfft),
independent of the project's own oracle; the N=8 and N=16 cases are also
recomputed with mpmath at 40 digits and the script refuses to write fixtures
if scipy's long-double result disagrees with mpmath beyond 1e-17 relative.
Each output is stored as a float64 pair (hi, lo) with hi + lo = value.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.fft

SIZES = [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]
HERE = os.path.dirname(os.path.abspath(__file__))


def signals(n, count, seed):
    g = np.random.Generator(np.random.Philox(seed))
    re = g.uniform(-1.0, 1.0, size=(count, n))
    im = g.uniform(-1.0, 1.0, size=(count, n))
    return re + 1j * im


def split(v):
    hi = v.astype(np.complex128)
    lo = (v - hi.astype(np.clongdouble)).astype(np.complex128)
    return hi, lo


def mp_dft(x, inverse):
    import mpmath
    mpmath.mp.dps = 40
    n = len(x)
    s = 1 if inverse else -1
    out = []
    for k in range(n):
        acc = mpmath.mpc(0)
        for j in range(n):
            acc += mpmath.mpc(float(x[j].real), float(x[j].imag)) * mpmath.expjpi(mpmath.mpf(2 * s * j * k) / n)
        out.append(acc)
    return out


def main():
    arrays = {}
    for n in SIZES:
        count = 2 if n <= 512 else 1
        x = signals(n, count, seed=1000 + n)
        xl = x.astype(np.clongdouble)
        for tag, inv in (("fwd", False), ("inv", True)):
            y = scipy.fft.ifft(xl, axis=-1) * n if inv else scipy.fft.fft(xl, axis=-1)
            if n <= 16:
                for b in range(count):
                    ref = mp_dft(x[b], inv)
                    # compare in long double precision
                    err = max(abs((y[b, k].real - np.longdouble(str(ref[k].real)))) +
                              abs((y[b, k].imag - np.longdouble(str(ref[k].imag)))) for k in range(n))
                    scale = max(abs(complex(r)) for r in ref)
                    assert float(err) / scale < 1e-17, (n, tag, float(err) / scale)
            hi, lo = split(y)
            arrays[f"x_{n}"] = x
            arrays[f"{tag}_hi_{n}"] = hi
            arrays[f"{tag}_lo_{n}"] = lo
    path = os.path.join(HERE, "golden_fft.npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
