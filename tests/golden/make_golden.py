"""Writes the golden fixtures in tests/golden/ from numpy/scipy alone (never from the oracle or the reference).

* epg_cases.npz    -- echo trains from a full-state EPG simulation: complex (F+, F-, Z) state vectors over all
                      2*ETL+1 dephasing orders, relaxation over TE/2, the refocusing rotation matrix, a gradient shift,
                      relaxation over TE/2 (Prasloski 2012 appendix; Hennig 1988). No state truncation.
* nnls_cases.npz   -- scipy.optimize.nnls solutions of seeded random problems (the reference's rand_NNLS_data shape:
                      A > 0, b = A x with x of alternating sign), plain and Tikhonov-augmented [A; mu I].
* svd_cases.npz    -- numpy singular values of EPG basis matrices.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import scipy.optimize

HERE = Path(__file__).resolve().parent


def epg_full(etl, alpha_deg, TE, T2, T1, beta_deg=180.0):
    """Decay curve of the A*90, A*180, A*beta, ... train (A = alpha/180) by full-state EPG simulation.

    State k >= 0 holds (F+_k, F-_k, Z_k). Per echo: relax TE/2, rotate every state with
        [[c^2, s^2, -i sin], [s^2, c^2, i sin], [-i sin/2, i sin/2, cos]]   (c = cos(a/2), s = sin(a/2)),
    shift dephasing orders (F+_k -> F+_{k+1}, F-_k -> F-_{k-1}, F-_0 -> F+_0 ... expressed through the
    reference's state-1-based bookkeeping: state j's (F, Fbar, Z) after the shift is (F_{j-1}, Fbar_{j+1}, Z_j),
    state 1's F takes state 1's Fbar), relax TE/2, read the echo as |F of state 1|.
    """
    A = alpha_deg / 180.0
    E1, E2 = np.exp(-(TE / 2) / T1), np.exp(-(TE / 2) / T2)
    E = np.array([E2, E2, E1])

    def rot(a):
        a = np.deg2rad(a)
        c2, s2, s = np.cos(a / 2) ** 2, np.sin(a / 2) ** 2, np.sin(a)
        return np.array([[c2, s2, -1j * s], [s2, c2, 1j * s], [-0.5j * s, 0.5j * s, np.cos(a)]])

    nst = 2 * etl + 2
    M = np.zeros((nst, 3), dtype=complex)
    M[0, 0] = np.sin(np.deg2rad(A * 90))
    dc = np.empty(etl)
    for i in range(etl):
        R = rot(A * 180 if i == 0 else A * beta_deg)
        M = (M * E) @ R.T
        new = np.zeros_like(M)
        new[0, 0] = M[0, 1]
        new[1:, 0] = M[:-1, 0]
        new[:-1, 1] = M[1:, 1]
        new[:, 2] = M[:, 2]
        M = new * E
        dc[i] = abs(M[0, 0])
    return dc


def rand_nnls(rng, m, n):
    x = rng.random(n)
    x[rng.integers(0, 2)::2] *= -1
    A = rng.random((m, n))
    return A, A @ x


def main():
    rng = np.random.default_rng(20240601)
    # EPG cases: (ETL, alpha, TE, T2, T1, beta)
    params = []
    for etl in (4, 5, 7, 16, 32, 48, 56):
        for alpha in (180.0, 165.0, 137.3, 120.0, 60.0):
            for beta in (180.0, 172.0, 150.0):
                params.append((etl, alpha, 10e-3, 0.05, 1.0, beta))
    params += [(10, 169.0, 9e-3, 10.1e-3, 0.98, 176.0), (64, 150.0, 39e-3, 1.1, 151.0, 180.0)]
    curves = np.zeros((len(params), 64))
    for r, p in enumerate(params):
        curves[r, : p[0]] = epg_full(*p)
    np.savez_compressed(HERE / "epg_cases.npz", params=np.array(params), curves=curves)

    # NNLS cases
    sizes = [(1, 1), (2, 2), (5, 3), (3, 5), (8, 8), (13, 5), (16, 16), (25, 32), (32, 25), (48, 40)]
    As, bs, xs, mus, Ats, bts, xts = [], [], [], [], [], [], []
    for (m, n) in sizes:
        for t in range(3):
            A, b = rand_nnls(rng, m, n)
            x, _ = scipy.optimize.nnls(A, b, maxiter=50 * n)
            mu = [1e-6, 1e-2, 1.0][t]
            Aa = np.vstack([A, mu * np.eye(n)])
            ba = np.concatenate([b, np.zeros(n)])
            xt, _ = scipy.optimize.nnls(Aa, ba, maxiter=50 * n)
            As.append(A.ravel()); bs.append(b); xs.append(x); mus.append(mu); xts.append(xt)
    def pack(arrs):  # ragged arrays flattened with offsets (no pickle)
        off = np.cumsum([0] + [a.size for a in arrs])
        return np.concatenate([a.ravel() for a in arrs]), off
    A_flat, A_off = pack([np.asarray(a) for a in As])
    b_flat, b_off = pack(bs)
    x_flat, x_off = pack(xs)
    xt_flat, xt_off = pack(xts)
    np.savez_compressed(HERE / "nnls_cases.npz", sizes=np.repeat(np.array(sizes), 3, axis=0), A=A_flat, A_off=A_off,
                        b=b_flat, b_off=b_off, x=x_flat, x_off=x_off, mu=np.array(mus), xt=xt_flat, xt_off=xt_off)

    # SVD cases: EPG bases (columns = curves over a T2 grid), numpy singular values
    T2 = np.exp(np.linspace(np.log(10e-3), np.log(2.0), 40))
    mats, svals = [], []
    for (m, alpha) in ((48, 150.0), (32, 170.0), (24, 130.0)):
        B = np.stack([epg_full(m, alpha, 10e-3, t, 1.0) for t in T2], axis=1)
        mats.append(B)
        svals.append(np.linalg.svd(B, compute_uv=False))
    M_flat, M_off = pack(mats)
    s_flat, s_off = pack(svals)
    np.savez_compressed(HERE / "svd_cases.npz", shapes=np.array([m.shape for m in mats]), M=M_flat, M_off=M_off,
                        s=s_flat, s_off=s_off)
    print("wrote", [p.name for p in HERE.glob("*.npz")])


if __name__ == "__main__":
    main()
