"""Writes tests/golden/forced_256_lundgren.npz: the float64 oracle's record of
the 256^3 constant-energy Lundgren-forced DNS (nu = 1/1000, dt = 1e-3, seeded
C k^4 exp(-2 (k/4)^2) initial field, seed 0) over its first 10 RK4 steps.

Stored per step n = 0..10: the full shell spectrum (bins 0..smax), E, Omega
and eps = 2 nu Omega; at steps 0, 1 and 10 the values of the state at 2048
fixed retained modes (indices drawn once, seed 2024, from the modes inside
the 2/3 box). The GPU parity test compares its own run of the same
configuration with these numbers, so the box does not have to run a 256^3
oracle (about 10-20 s per RK4 step on a CPU).

Run from the repository root: python tests/golden/make_forced_256.py
(about 4 minutes on 8 cores)."""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

import hit3d_oracle as O  # noqa: E402

N, NU, DT, NSTEPS, SEED, NSAMPLE = 256, 1e-3, 1e-3, 10, 0, 2048
OUT = os.path.join(HERE, "forced_256_lundgren.npz")


def sums(u, g):
    a2 = np.sum(np.abs(u) ** 2, axis=0)
    full = np.bincount(O.shells(g)[0].ravel(), weights=((g.weights * a2) / (2 * O._ntot2(g))).ravel(),
                       minlength=O.shells(g)[1] + 1)
    return full, O.energy(u, g), O.enstrophy(u, g)


def main():
    g = O.make_grid(N)
    rhs = O.make_ns_rhs(g)
    params = O.NSParams(nu=NU, forcing=O.LundgrenForcing(nu=NU))
    u = O.random_field(SEED, g, O.config_spectrum)
    mask = np.broadcast_to(g.dealias, u.shape[1:])
    retained = np.flatnonzero(np.broadcast_to(mask, u.shape))
    idx = np.sort(np.random.default_rng(2024).choice(retained, NSAMPLE, replace=False))
    spectra, E, Om, samples = [], [], [], {}
    t0 = time.time()
    for n in range(NSTEPS + 1):
        if n > 0:
            u = O.rk4_step(rhs, u, params, DT, (n - 1) * DT)
        s, e, om = sums(u, g)
        spectra.append(s)
        E.append(e)
        Om.append(om)
        if n in (0, 1, NSTEPS):
            samples[n] = u.ravel()[idx]
        print(f"step {n}: E = {e:.15f}, Omega = {om:.6f} ({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(OUT, N=N, nu=NU, dt=DT, seed=SEED, index=idx.astype(np.int64),
                        spectra=np.asarray(spectra), E=np.asarray(E), Omega=np.asarray(Om),
                        eps=2 * NU * np.asarray(Om), sample0=samples[0], sample1=samples[1],
                        sample10=samples[NSTEPS])
    print("wrote", OUT)


if __name__ == "__main__":
    main()
