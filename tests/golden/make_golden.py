"""Generate tests/golden/witt_poly_roots.npz: exact critical-curve roots of
tiny star fields, independent of any Aberth iteration.

For a circular field with the approximate smooth sheet and at most 32 stars
(no tree, no FMM), Witt's function is
    f0(z) = gamma + theta^2 sum_i m_i / (z - s_i)^2 - (1 - kappa + kappa_star) e^{-i phi}
(weisluke/ccf include/ccf_functions.cuh:69-86 with alpha_smooth.cuh:141-236).
Multiplying by prod_i (z - s_i)^2 gives the degree-2N polynomial
    P(z) = (gamma - A e^{-i phi}) prod_i (z - s_i)^2 + theta^2 sum_i m_i prod_{j != i} (z - s_j)^2,
whose roots are found with mpmath.polyroots at 50 digits.

Run from the repository root: python tests/golden/make_golden.py
"""
import os

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def poly_mul(a, b):
    out = [mp.mpc(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def witt_roots(stars, masses, kappa, kappa_star, gamma, phi, theta=1.0):
    mp.mp.dps = 50
    A = 1 - mp.mpf(kappa) + mp.mpf(kappa_star)
    c = mp.mpf(gamma) - A * mp.expj(-mp.mpf(phi))
    n = len(stars)
    lin = [[mp.mpc(1), -mp.mpc(s.real, s.imag)] for s in stars]  # z - s_i, highest power first
    sq = [poly_mul(l, l) for l in lin]
    Q = [mp.mpc(1)]
    for q in sq:
        Q = poly_mul(Q, q)
    P = [c * x for x in Q]
    for i in range(n):
        term = [mp.mpc(theta) ** 2 * mp.mpf(masses[i])]
        for j in range(n):
            if j != i:
                term = poly_mul(term, sq[j])
        off = len(P) - len(term)
        for k, x in enumerate(term):
            P[off + k] += x
    r = mp.polyroots(P, maxsteps=500, extraprec=200)
    return np.array([complex(x) for x in r])


def main():
    rng = np.random.default_rng(20261202)
    cases = []
    # (num_stars, kappa, kappa_star, gamma)
    for n, kappa, ks, gamma in [(1, 0.5, 0.3, 0.15), (2, 0.4, 0.4, 0.3), (4, 0.6, 0.35, 0.2), (6, 0.45, 0.45, 0.1)]:
        r = 3.0 * np.sqrt(rng.random(n))
        a = 2 * np.pi * rng.random(n)
        s = r * np.exp(1j * a)
        if n == 1:
            s = np.array([0.0 + 0.0j])
        m = rng.uniform(0.5, 1.5, n) if n > 1 else np.array([1.0])
        rad = np.sqrt(m.sum() / ks)  # circular field whose kappa_star is ks (star.cuh:159)
        num_phi = 20
        roots = np.stack([witt_roots(s, m, kappa, ks, gamma, 2 * np.pi * k / num_phi) for k in range(num_phi + 1)])
        cases.append(dict(stars=np.stack([s.real, s.imag, m], axis=1), rad=rad, kappa=kappa, kappa_star=ks, gamma=gamma,
                          num_phi=num_phi, roots=roots))
    out = {}
    for i, c in enumerate(cases):
        for k, v in c.items():
            out[f"c{i}_{k}"] = np.asarray(v)
    out["ncases"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "witt_poly_roots.npz"), **out)
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    main()
