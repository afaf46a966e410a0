"""The inflow-turbulence generator's oracle (oracle/pseudo2d.py) pinned by what the mathematics
fixes, so that the device generator can be checked against it:
  * the unnormalized inverse real transform equals a brute-force double sum with the reference's
    sign (+) and the real parts of the l = 0 and l = n1/2 columns;
  * the sample variance of the raw fields matches the closed-form expectation of the spectrum
    (amplitude sqrt(dk)*c*e(l,p), weights 2 for 0 < l < n1/2 and 1/2 for l = 0, n1/2), which a
    dropped factor in c, dk or the exponent would miss;
  * the normalization constant c of src/m_pseudo2D.F90:99-104 for a hand-checked small case;
  * after the AR(1) step every slice 1..nrturb has zero mean and unit sample deviation, slice 0
    is the raw field, and the lag-one correlation is close to 0.98.
The random phases themselves are parity unpinned (the reference draws them from random_number).
"""
import numpy as np

import pseudo2d


def test_transform_matches_brute_force():
    rng = np.random.default_rng(1)
    n1, n2 = 6, 4
    nh = n1 // 2 + 1
    X = rng.normal(size=(n2, nh)) + 1j * rng.normal(size=(n2, nh))
    y = np.fft.irfft2(X, s=(n2, n1), axes=(-2, -1)) * (n1 * n2)
    ref = np.zeros((n2, n1))
    for m in range(n2):
        for i in range(n1):
            tot = 0.0
            for l in range(nh):
                w = 1.0 if l in (0, n1 // 2) else 2.0
                z = sum(X[p, l] * np.exp(2j * np.pi * p * m / n2) for p in range(n2))
                tot += w * (z * np.exp(2j * np.pi * l * i / n1)).real
            ref[m, i] = tot
    assert np.abs(y - ref).max() < 1e-12


def test_normalization_constant_small_case():
    n1, n2 = 4, 4
    dk, c, ak, al = pseudo2d.spectrum_constants(n1, n2)
    pi2 = 2 * 3.141592653589
    r = np.sqrt(8.0) / (10.0 / np.sqrt(3.0))
    s = 0.0
    for p in (-1, 0, 1, 2):
        for l in (-1, 0, 1, 2):
            s += np.exp(-2.0 * ((pi2 / 4) ** 2 * l * l / r ** 2 + (pi2 / 4) ** 2 * p * p / r ** 2))
    s -= 1.0
    assert abs(dk - pi2 ** 2 / 16) < 1e-15
    assert abs(c - np.sqrt(1.0 / (dk * s))) < 1e-12 * c
    assert abs(ak - (pi2 / 4) ** 2 / r ** 2) < 1e-15


def test_raw_variance_matches_spectrum():
    ny, nz, nf = 24, 20, 600
    y = pseudo2d.raw_fields(ny, nz, nf, 11)
    n1, n2 = ny, nz
    dk, c, ak, al = pseudo2d.spectrum_constants(n1, n2)
    l = np.arange(n1 // 2 + 1)[:, None]
    pp = np.arange(n2)[None, :]
    p = np.where(pp < n2 // 2, pp, pp - n2)
    e2 = np.exp(-2.0 * (ak * l ** 2 + al * p ** 2))
    e2[0, 0] = 0.0
    w2 = np.full(l.shape, 2.0)
    w2[0] = w2[-1] = 0.5
    expect = dk * c * c * (w2 * e2).sum()
    var = (y ** 2).mean()
    # nf * ny * nz correlated samples; the spread of the estimate is a few percent
    assert abs(var / expect - 1.0) < 0.05, (var, expect)
    assert abs(y.mean()) < 0.05


def test_ar1_slices_normalized():
    uu, vv, ww, rr = pseudo2d.generate(24, 16, 40, 5)
    raw0 = pseudo2d.raw_fields(24, 16, 4 * 41, 5).reshape(4, 41, 16, 24)[0, 0]
    assert np.array_equal(uu[0], raw0)
    for f in (uu, vv, ww, rr):
        assert np.abs(f[1:].mean(axis=(1, 2))).max() < 1e-12
        assert np.abs(f[1:].std(axis=(1, 2), ddof=1) - 1.0).max() < 1e-12
    lag = np.mean([(uu[i] * uu[i + 1]).mean() for i in range(1, 40)])
    assert 0.95 < lag < 1.0
