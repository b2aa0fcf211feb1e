"""Generate golden vectors from the reference's own compiled serial kernels.

Runs only where the reference tree is present (the build container): it builds
oracle/_ref/libnrsref.so from the reference sources (oracle/build_ref.sh) and
records inputs and outputs of
  * ellipticPartialAxCoeffHex3D (Poisson, p_lambda=0; Helmholtz, p_lambda=1),
    FP64 and FP32, orders N = 1,2,3,4,5,7,9;
  * ellipticBlockPartialAxCoeffHex3D (3 fields), N = 3, 7;
  * ellipticBlockUpdatePCG;
  * ellipticPreconCoarsenHex3D / ProlongateHex3D for (Nf,Nc) pairs;
  * gslib gs() (double/float/int; add, mul, min, max) on box-mesh ids and on
    random id lists with repeats and zeros;
  * preFDM / fusedFDM (ASM and RAS) / postFDM in FP32 for Nq = 2, 4, 8.
Inputs are seeded (numpy PCG64); element counts are tiny so the fixtures stay
small.  Output: tests/golden/*.npz (data only).

usage: python tests/golden/gen_golden.py [REFERENCE_ROOT] [fixture ...]
(fixtures: ax block_ax updatepcg transfer gs fdm; default all)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

ci = lambda v: C.byref(C.c_int(v))  # noqa: E731
dp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731


def gen_ax(R, rng):
    out = {}
    for N in [1, 2, 3, 4, 5, 7, 9]:
        Nq = N + 1
        Np = Nq ** 3
        E = 3 if N <= 5 else 2
        D = oracle.dmatrix(N)
        S = D.T.copy()
        for T, tn in [(np.float64, "double"), (np.float32, "float")]:
            Dt, St = D.astype(T), S.astype(T)
            ggeo = rng.uniform(-1, 1, E * 7 * Np).astype(T)
            q = rng.uniform(-1, 1, E * Np).astype(T)
            elist = np.array(rng.permutation(E), dtype=np.int32)
            # Poisson, constant lambda0
            l0 = np.array([rng.uniform(0.5, 2.0)], dtype=T)
            l1 = np.zeros(1, dtype=T)
            Aq = np.zeros_like(q)
            getattr(R, "ref_ax_poisson_%s_nq%d" % (tn, Nq))(
                ci(E), ci(E * Np), ci(E * Np), dp(elist), dp(ggeo), dp(Dt), dp(St), dp(l0), dp(l1),
                dp(q), dp(Aq))
            key = "poisson_%s_N%d" % (tn, N)
            out.update({key + "_ggeo": ggeo, key + "_q": q, key + "_D": Dt, key + "_elist": elist,
                        key + "_lambda0": l0, key + "_Aq": Aq})
            # Helmholtz with coefficient fields
            L0 = rng.uniform(0.5, 2.0, E * Np).astype(T)
            L1 = rng.uniform(0.0, 10.0, E * Np).astype(T)
            Aq = np.zeros_like(q)
            getattr(R, "ref_ax_helm_%s_nq%d" % (tn, Nq))(
                ci(E), ci(E * Np), ci(E * Np), dp(elist), dp(ggeo), dp(Dt), dp(St), dp(L0), dp(L1),
                dp(q), dp(Aq))
            key = "helm_%s_N%d" % (tn, N)
            out.update({key + "_ggeo": ggeo, key + "_q": q, key + "_D": Dt, key + "_elist": elist,
                        key + "_lambda0": L0, key + "_lambda1": L1, key + "_Aq": Aq})
    return out


def gen_block_ax(R, rng):
    out = {}
    for N in [3, 7]:
        Nq = N + 1
        Np = Nq ** 3
        E = 2
        L = E * Np
        D = oracle.dmatrix(N)
        S = D.T.copy()
        ggeo = rng.uniform(-1, 1, E * 7 * Np)
        q = rng.uniform(-1, 1, 3 * L)
        L0 = rng.uniform(0.5, 2.0, 3 * L)
        L1 = rng.uniform(0.0, 10.0, 3 * L)
        elist = np.arange(E, dtype=np.int32)
        Aq = np.zeros(3 * L)
        getattr(R, "ref_blockax_nq%d" % Nq)(ci(E), ci(L), ci(L), dp(elist), dp(ggeo), dp(D), dp(S),
                                             dp(L0), dp(L1), dp(q), dp(Aq))
        key = "block_N%d" % N
        out.update({key + "_ggeo": ggeo, key + "_q": q, key + "_D": D, key + "_lambda0": L0,
                    key + "_lambda1": L1, key + "_Aq": Aq})
    return out


def gen_updatepcg(R, rng):
    n = 1000
    w = rng.uniform(0.1, 1.0, n)
    Ap = rng.uniform(-1, 1, n)
    r = rng.uniform(-1, 1, n)
    alpha = 0.37
    r_out = r.copy()
    rdotr = np.zeros(1)
    R.ref_updatepcg(ci(n), ci(n), dp(w), dp(Ap), C.byref(C.c_double(alpha)), dp(r_out), dp(rdotr))
    return {"updatepcg_w": w, "updatepcg_Ap": Ap, "updatepcg_r": r, "updatepcg_alpha": np.array([alpha]),
            "updatepcg_r_out": r_out, "updatepcg_rdotr": rdotr}


def gen_transfer(R, rng):
    out = {}
    for nf, nc in [(8, 4), (8, 2), (4, 2), (8, 6), (6, 4), (10, 4), (10, 6), (6, 2)]:
        Nf, Nc = nf - 1, nc - 1
        E = 2
        Rm = oracle.transfer_R(Nf, Nc)       # [NqC][NqF]
        qf = rng.uniform(-1, 1, E * nf ** 3)
        qc = np.zeros(E * nc ** 3)
        getattr(R, "ref_coarsen_%d_%d" % (nf, nc))(ci(E), dp(Rm), dp(qf), dp(qc))
        qc_in = rng.uniform(-1, 1, E * nc ** 3)
        qf_acc = rng.uniform(-1, 1, E * nf ** 3)
        qf_out = qf_acc.copy()
        getattr(R, "ref_prolongate_%d_%d" % (nf, nc))(ci(E), dp(Rm), dp(qc_in), dp(qf_out))
        key = "transfer_%d_%d" % (Nf, Nc)
        out.update({key + "_R": Rm, key + "_qf": qf, key + "_qc": qc, key + "_qc_in": qc_in,
                    key + "_qf_acc": qf_acc, key + "_qf_out": qf_out})
    return out


def gen_gs(R, rng):
    out = {}
    cases = {}
    m = oracle.BoxMesh(3, 2, 2, 3, "kershaw", 0.3)
    cases["box"] = m.gid.copy()
    ids = rng.integers(0, 40, 300).astype(np.int64)    # zeros = excluded, many repeats
    cases["random"] = ids
    cases["singletons"] = np.arange(1, 51, dtype=np.int64)
    for name, ids in cases.items():
        n = len(ids)
        out["gs_%s_ids" % name] = ids
        for dom, T in [(0, np.float64), (1, np.float32), (2, np.int32)]:
            if dom == 2:
                v = rng.integers(-1000, 1000, n).astype(np.int32)
            else:
                v = rng.uniform(-1, 1, n).astype(T)
            out["gs_%s_v_%s" % (name, T.__name__)] = v
            for op, opn in enumerate(["add", "mul", "min", "max"]):
                if dom == 2 and opn == "mul":
                    continue
                r = v.copy()
                R.ref_gslib_gs(dp(ids), C.c_uint(n), dp(r), dom, op)
                out["gs_%s_%s_%s" % (name, opn, T.__name__)] = r
    return out


def gen_fdm(R, rng):
    out = {}
    f32 = lambda n: rng.uniform(-1, 1, n).astype(np.float32)   # noqa: E731
    for nq in [2, 4, 8]:
        nl = nq + 2
        E = 2
        key = "fdm_nq%d" % nq
        u = f32(E * nq ** 3)
        work1 = np.zeros(E * nl ** 3, dtype=np.float32)
        getattr(R, "ref_prefdm_nq%d" % nq)(ci(E), dp(u), dp(work1))
        out.update({key + "_pre_u": u, key + "_pre_work1": work1})
        Sx, Sy, Sz = f32(E * nl * nl), f32(E * nl * nl), f32(E * nl * nl)
        invL = rng.uniform(0.1, 1.0, E * nl ** 3).astype(np.float32)
        w1 = f32(E * nl ** 3)
        elist = np.arange(E, dtype=np.int32)
        out.update({key + "_Sx": Sx, key + "_Sy": Sy, key + "_Sz": Sz, key + "_invL": invL,
                    key + "_w1": w1})
        Su = np.zeros(E * nl ** 3, dtype=np.float32)
        uw = w1.copy()
        getattr(R, "ref_fusedfdm_asm_nq%d" % nq)(ci(E), dp(elist), dp(Su), dp(Sx), dp(Sy), dp(Sz),
                                                  dp(invL), dp(uw))
        out.update({key + "_asm_Su": Su, key + "_asm_u": uw})
        wts = rng.uniform(0.1, 1.0, E * nq ** 3).astype(np.float32)
        Su = np.zeros(E * nq ** 3, dtype=np.float32)
        uw = w1.copy()
        getattr(R, "ref_fusedfdm_ras_nq%d" % nq)(ci(E), dp(elist), dp(Su), dp(Sx), dp(Sy), dp(Sz),
                                                  dp(invL), dp(wts), dp(uw))
        out.update({key + "_ras_wts": wts, key + "_ras_Su": Su})
        a, b = f32(E * nl ** 3), f32(E * nl ** 3)
        Su = np.zeros(E * nq ** 3, dtype=np.float32)
        getattr(R, "ref_postfdm_nq%d" % nq)(ci(E), dp(a.copy()), dp(b.copy()), dp(Su), dp(wts))
        out.update({key + "_post_work1": a, key + "_post_work2": b, key + "_post_Su": Su})
    return out


GENERATORS = {"ax": gen_ax, "block_ax": gen_block_ax, "updatepcg": gen_updatepcg,
              "transfer": gen_transfer, "gs": gen_gs}


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    only = set(sys.argv[2:])
    oracle.build()
    if oracle.build_ref(ref_root) is None:
        print("reference tree not found; nothing generated")
        return 1
    R = oracle.ref_lib()
    rng = np.random.default_rng(20261202)      # one stream through the original five, in order
    for name, fn in GENERATORS.items():
        data = fn(R, rng)
        if not only or name in only:
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
    if not only or "fdm" in only:
        np.savez_compressed(os.path.join(HERE, "fdm.npz"), **gen_fdm(R, np.random.default_rng(20261203)))
    print("golden fixtures written to", HERE)
    return 0


if __name__ == "__main__":
    sys.exit(main())
