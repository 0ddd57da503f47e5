"""Generate the golden fixtures of the CPU test suite.

Independent references only (no code of this repository is used):
  * GSL's published known-answer values of ranlxd1 / ranlxd2 (gsl rng/test.c:
    rng_test(gsl_rng_ranlxd1, 1, 10000, 1998227290) and
    rng_test(gsl_rng_ranlxd2, 1, 10000, 3949287736): gsl_rng_get of the
    10000th draw after gsl_rng_set(r, 1));
  * mpmath for the linear growth factor, its derivative, the COLA time
    integrals and the hypergeometric function;
  * numpy.fft for FFT vectors;
  * scipy.integrate.quad for sigma8 of the reference's input P(k) table
    (data/linear_matterpower.dat, copied next to this script as data).

Usage: python tests/golden/gen_golden.py  (writes next to this file)
"""
import json
import os

import mpmath as mp
import numpy as np
from scipy import integrate

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 30


def growth_fixtures(om):
    """D(a)/D(1) of flat LCDM: D ~ E(a) int_0^a da' / (a' E(a'))^3."""
    E = lambda a: mp.sqrt(om / a ** 3 + 1 - om)
    Du = lambda a: E(a) * mp.quad(lambda x: 1 / (x * E(x)) ** 3, [0, a])
    D1 = Du(1)
    avals = [0.05, 0.1, 0.15, 0.19, 0.21, 0.25, 0.3, 0.4, 0.5, 0.6, 0.7, 0.72, 0.7215, 0.75, 0.8, 0.9, 0.95, 1.0]
    out = []
    for a in avals:
        a = mp.mpf(a)
        d = Du(a) / D1
        qdd = a ** 3 * E(a) * mp.diff(lambda x: Du(x) / D1, a)
        out.append({"a": float(a), "D": float(d), "QdDda": float(qdd)})
    return out


def sq_fixtures(om):
    """Sq(ai, af, aRef) = int_ai^af a^-2.5 / Q(a) da / aRef^-2.5, Q = a^3 E(a)
    and Sphi = int_ai^af (a/aRef)^-3.5 da * aRef / Q(aRef) (the closed form of
    cola.c:287-292 written as an integral)."""
    Q = lambda a: a ** 3 * mp.sqrt(om / a ** 3 + 1 - om)
    cases = [(0.1, 0.2, 0.15), (0.15, 0.25, 0.2), (0.5, 0.6, 0.55), (0.9, 1.0, 0.95), (0.3, 0.35, 0.3)]
    out = []
    for ai, af, ar in cases:
        ai, af, ar = mp.mpf(ai), mp.mpf(af), mp.mpf(ar)
        sq = mp.quad(lambda a: a ** -2.5 / Q(a), [ai, af]) / ar ** -2.5
        sphi = mp.quad(lambda a: (a / ar) ** -3.5, [ai, af]) * ar / Q(ar)
        out.append({"ai": float(ai), "af": float(af), "aref": float(ar), "Sq": float(sq), "Sphi": float(sphi)})
    return out


def growth_factor_fixtures(om):
    """power.c GrowthFactor: D = H(a) int_0^a (a / (Om + OL a^3))^1.5 da."""
    ol = 1 - om
    D = lambda a: mp.sqrt(om / a ** 3 + ol) * mp.quad(lambda x: (x / (om + ol * x ** 3)) ** 1.5, [0, a])
    return [{"astart": a, "aend": 1.0, "G": float(D(1) / D(mp.mpf(a)))} for a in (0.05, 0.1, 0.2, 0.5)]


def hyp2f1_fixtures():
    pars = [(0.5, 2 / 3, 5 / 3), (-0.5, 2 / 3, 5 / 3), (-1 / 6, 0.5, 5 / 6), (-7 / 6, -0.5, -1 / 6)]
    out = []
    for a, b, c in pars:
        for z in (-0.001, -0.05, -0.3755, -0.7, -0.95, -0.9989):
            out.append({"a": a, "b": b, "c": c, "z": z, "value": float(mp.hyp2f1(a, b, c, z))})
    return out


def sigma8_of_table(path):
    """sigma8 of a (k, P) table with power.c's log-log interpolation
    (power.c:120-151) and top-hat window, by scipy.integrate.quad."""
    k, p = np.loadtxt(path, unpack=True)
    lk = np.log10(k)
    lD = np.log10(k ** 3 * p / (2 * np.pi ** 2))

    def pk(kk):
        x = np.log10(kk)
        if x < lk[0] or x > lk[-1]:
            return 0.0
        return 10 ** np.interp(x, lk, lD) / (4 * np.pi * kk ** 3)

    def integrand(kk):
        kr = 8.0 * kk
        if kr < 1e-8:
            return 0.0
        w = 3 * (np.sin(kr) / kr ** 3 - np.cos(kr) / kr ** 2)
        return 4 * np.pi * kk * kk * w * w * pk(kk)

    pts = [x for x in k if x < 500 / 8]
    val = 0.0
    edges = [0.0] + pts + [500.0 / 8]
    for lo, hi in zip(edges[:-1], edges[1:]):
        if hi > lo:
            val += integrate.quad(integrand, lo, hi, epsabs=0, epsrel=1e-10, limit=200)[0]
    return float(np.sqrt(val))


def fft_fixtures(path):
    rng = np.random.default_rng(2024)
    data = {}
    for n in (6, 10, 16):
        g = rng.standard_normal((n, n, n)).astype(np.float32)
        data[f"real_{n}"] = g
        data[f"rfftn_{n}"] = np.fft.rfftn(g.astype(np.float64)).astype(np.complex128)
        h = (rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))).astype(np.complex64)
        data[f"half_{n}"] = h
        data[f"irfftn_{n}"] = np.fft.irfftn(h.astype(np.complex128), s=(n, n, n), axes=(0, 1, 2)) * n ** 3
    for n in (7, 12, 30, 96, 105):
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        data[f"c1d_{n}"] = v
        data[f"fft1d_{n}"] = np.fft.fft(v)
    np.savez_compressed(path, **data)


def main():
    om = float(np.float32(0.273))
    cosmo = {
        "omega_m_float32": om,
        "growth": growth_fixtures(om),
        "time_integrals": sq_fixtures(om),
        "growth_factor_power_c": growth_factor_fixtures(om),
        "hyp2f1": hyp2f1_fixtures(),
    }
    with open(os.path.join(HERE, "cosmo_golden.json"), "w") as f:
        json.dump(cosmo, f, indent=1)
    kat = {
        "source": "GSL rng/test.c known-answer tests: gsl_rng_get after 10000 draws from gsl_rng_set(r, 1)",
        "ranlxd1": {"seed": 1, "n": 10000, "value": 1998227290},
        "ranlxd2": {"seed": 1, "n": 10000, "value": 3949287736},
    }
    with open(os.path.join(HERE, "ranlux_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    table = os.path.join(HERE, "linear_matterpower.dat")
    with open(os.path.join(HERE, "sigma8_golden.json"), "w") as f:
        json.dump({"table": "linear_matterpower.dat", "sigma8_scipy": sigma8_of_table(table)}, f, indent=1)
    fft_fixtures(os.path.join(HERE, "fft_golden.npz"))


if __name__ == "__main__":
    main()
