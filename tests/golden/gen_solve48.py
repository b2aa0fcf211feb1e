"""Record the oracle's C2 / C3 solves on the 48^3 Kershaw mesh as fixtures.

The full-size pressure-type Poisson solves (48^3 = 110 592 elements, N = 7,
Kershaw eps = 0.3, right-hand side 3 pi^2 sin sin sin + 1e-3 Gaussian noise,
seed 0, weighted by the mass matrix; p-MG 7-3-1, coarse level by two FP64 PCG
iterations on the assembled order-1 matrix preconditioned by one AMG V-cycle;
flexible CG to 1e-8 relative) run through the oracle (oracle/mg.py) in this
container, which takes a long time on the CPU; the GPU test
(tests/test_gpu_solve48.py) then checks the library's iteration counts (+-1)
and solutions (1e-8 relative) against what this script stores:

  solve48_<case>.npz: iterations, res0, res, sample indices and the solution
  at those indices, max |x|, the weighted 2-norm of x, and the residual
  history.

Cases: fp64_chebyshev_jacobi (FP64 multigrid, 1st-kind Chebyshev-Jacobi degree
2), mixed_opt4chebyshev_ras (FP32 multigrid, optimal 4th-kind Chebyshev
degree 4 around RAS, FDM contractions rounded as BF16 with FP32 accumulation).

Usage: python tests/golden/gen_solve48.py [n] [case ...]   (n = 48 by default)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

CASES = {
    "fp64_chebyshev_jacobi": dict(T=np.float64, degree=2, cheby="cheby1", S="jacobi", fmt=None),
    "mixed_opt4chebyshev_ras": dict(T=np.float32, degree=4, cheby="opt4", S="ras", fmt="bf16"),
}
NSAMPLE = 20000


def rhs_of(mesh, seed=0):
    """3 pi^2 sin(pi x) sin(pi y) sin(pi z) + 1e-3 N(0,1), times the mass matrix GwJ."""
    u = np.sin(np.pi * mesh.x) * np.sin(np.pi * mesh.y) * np.sin(np.pi * mesh.z)
    f = 3 * np.pi ** 2 * u + 1e-3 * np.random.default_rng(seed).standard_normal(u.size)
    gw = mesh.ggeo.reshape(mesh.E, 7, mesh.Np)[:, 6].ravel()
    return gw * f


def sample_indices(L, seed=1):
    return np.sort(np.random.default_rng(seed).choice(L, size=min(NSAMPLE, L), replace=False))


def main():
    import oracle
    import oracle.mg as omg
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 48
    names = sys.argv[2:] or list(CASES)
    oracle.build()
    t0 = time.time()
    mesh = oracle.BoxMesh(n, n, n, 7, "kershaw", 0.3)
    eo = oracle.Elliptic(mesh)
    rhs = rhs_of(mesh)
    idx = sample_indices(mesh.Nlocal)
    print("mesh %d^3 ready in %.1f s" % (n, time.time() - t0), flush=True)
    for name in names:
        kw = CASES[name]
        t1 = time.time()
        om = omg.MG(mesh, [7, 3, 1], coarse="amg", coarse_tol=0.0, coarse_maxit=2, seed=0, **kw)
        print("%s: setup %.1f s, rho %s" % (name, time.time() - t1,
                                             [getattr(L, "rho", None) for L in om.levels]), flush=True)
        t1 = time.time()
        hist = []

        def precon(r, _hist=hist):
            _hist.append(time.time())
            if len(_hist) % 5 == 0:
                print("  %s: %d V-cycles, %.0f s" % (name, len(_hist), time.time() - t1), flush=True)
            return om.apply(r)
        x, info = omg.fcg_solve(eo, rhs, precon, flexible=True, tol=1e-8, maxit=1000)
        w = eo.invDegree
        print("%s: %d iterations, res %.3e / %.3e, %.1f s" % (name, info["iterations"], info["res"],
                                                              info["res0"], time.time() - t1), flush=True)
        out = os.path.join(ROOT, "tests", "golden", "solve48_%s%s.npz" % (name, "" if n == 48 else "_n%d" % n))
        np.savez_compressed(out, n=n, iterations=info["iterations"], res0=info["res0"], res=info["res"],
                            history=np.asarray(info.get("history", []), dtype=np.float64),
                            idx=idx.astype(np.int64), x=x[idx], xmax=np.max(np.abs(x)),
                            xnorm=np.sqrt(np.sum(w * x * x)))
        print("wrote", out, flush=True)
        del om


if __name__ == "__main__":
    main()
