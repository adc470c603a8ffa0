"""Write the golden fixtures of the SPH derivative pass with the CPU oracle.

    python tests/golden/make_golden.py

Each fixture holds the inputs (internal nodes and, where boundaries apply,
the ghost rows the oracle made), the neighbour CSR, the summed density and
every evaluateDerivatives field of the oracle's reference-order (scatter)
loop.  Workloads:
  converging3d_8  3-D converging flow, 8^3 jittered lattice, reflecting box
  tube1d_200      1-D two-state tube (left rho=1, P=1; right rho=0.125,
                  P=0.1; equal-mass spacing), reflecting walls
  blast2d_30      2-D 30x30 jittered lattice, rho=1, P=1e-6, with the
                  energy of the central nodes raised (point blast), periodic
The fixtures are data written by this repository's oracle; they pin the
oracle against regressions and the device path against stored vectors."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "llnl-spheral_amd")):
    sys.path.insert(0, p)

import oracle  # noqa: E402
from sphgpu import problems  # noqa: E402


def tube1d(n=200, gamma=1.4):
    """Equal-mass two-state tube on [0, 1], split at 0.5 by mass."""
    rhoL, rhoR, PL, PR = 1.0, 0.125, 1.0, 0.1
    mtot = 0.5 * rhoL + 0.5 * rhoR
    m = mtot / n
    nL = int(round(n * 0.5 * rhoL / mtot))
    nR = n - nL
    xL = (np.arange(nL) + 0.5) * (0.5 / nL)
    xR = 0.5 + (np.arange(nR) + 0.5) * (0.5 / nR)
    pos = np.concatenate([xL, xR])[:, None]
    rho = np.concatenate([np.full(nL, rhoL), np.full(nR, rhoR)])
    P = np.concatenate([np.full(nL, PL), np.full(nR, PR)])
    eps = P / ((gamma - 1.0) * rho)
    dx = np.concatenate([np.full(nL, 0.5 / nL), np.full(nR, 0.5 / nR)])
    H = (1.0 / (1.51 * dx))[:, None]
    P2, cs = problems.gamma_law(rho, eps, gamma)
    return dict(pos=pos, vel=np.zeros((n, 1)), mass=np.full(n, m), rho=rho, eps=eps, P=P2, cs=cs, H=H,
                omega=np.ones(n))


def blast2d(n=30, gamma=5.0 / 3.0):
    pos, mass, H, dx = problems.lattice(2, n, jitter=0.01, seed=7)
    N = pos.shape[0]
    rho = np.ones(N)
    eps = np.full(N, 1.0e-6 / ((gamma - 1.0) * 1.0))
    r = np.sqrt(((pos - 0.5) ** 2).sum(1))
    near = r < 1.5 * dx
    eps[near] = 1.0 / (mass[near].sum())
    P, cs = problems.gamma_law(rho, eps, gamma)
    return dict(pos=pos, vel=np.zeros((N, 2)), mass=mass, rho=rho, eps=eps, P=P, cs=cs, H=H, omega=np.ones(N))


WORKLOADS = {
    "converging3d_8": (3, lambda: problems.converging_flow(8)[0], "reflecting"),
    "tube1d_200": (1, tube1d, "reflecting"),
    "blast2d_30": (2, blast2d, "periodic"),
}


def compute(dim, st, kind):
    N = st["pos"].shape[0]
    bo = oracle.box_boundaries(dim, kind=kind)
    full, off, nbr, info = oracle.ghost_pipeline(dim, st, bo)
    T = oracle.Table(dim, 1000)
    rho = oracle.sum_density(T, N, full["pos"], full["mass"], full["H"], off, nbr)
    rho_all = np.array(full["rho"])
    rho_all[:N] = rho
    rho_all = oracle.apply_ghosts(dim, N, bo, info["control"], info["sub_of"], rho_all, 0)
    full["rho"] = rho_all
    out = oracle.eval_derivs(T, oracle.make_params(), N, full, off, nbr)
    return full, off, nbr, rho, out


def main():
    for name, (dim, make, kind) in WORKLOADS.items():
        st = make()
        full, off, nbr, rho, out = compute(dim, st, kind)
        arrays = {f"in_{k}": np.asarray(v) for k, v in st.items()}
        arrays.update({f"full_{k}": np.asarray(v) for k, v in full.items()})
        arrays.update(off=off, nbr=nbr, rhoSum_pass=rho, ncalc=np.array(out.pop("ncalc")))
        arrays.update({f"out_{k}": np.asarray(v) for k, v in out.items()})
        arrays.update(dim=np.array(dim), kind=np.array(kind))
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **arrays)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
