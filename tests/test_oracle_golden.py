"""The oracle against golden vectors recorded from the reference's own compiled
serial kernels and its vendored gslib (tests/golden/gen_golden.py)."""
import numpy as np
import pytest

NS = [1, 2, 3, 4, 5, 7, 9]


def rel(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))) /
                 max(np.max(np.abs(b.astype(np.float64))), 1e-300))


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("tn", ["double", "float"])
def test_ax_poisson_golden(orc, golden, N, tn):
    g = golden("ax")
    k = "poisson_%s_N%d" % (tn, N)
    T = np.float64 if tn == "double" else np.float32
    Aq = orc.ax(N, g[k + "_ggeo"], g[k + "_D"], g[k + "_q"], lambda0=float(g[k + "_lambda0"][0]),
                elist=g[k + "_elist"], dtype=T)
    tol = 1e-14 if tn == "double" else 2e-6
    assert rel(Aq, g[k + "_Aq"]) < tol


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("tn", ["double", "float"])
def test_ax_helmholtz_golden(orc, golden, N, tn):
    g = golden("ax")
    k = "helm_%s_N%d" % (tn, N)
    T = np.float64 if tn == "double" else np.float32
    Aq = orc.ax(N, g[k + "_ggeo"], g[k + "_D"], g[k + "_q"], lambda0=g[k + "_lambda0"],
                lambda1=g[k + "_lambda1"], poisson=False, elist=g[k + "_elist"], dtype=T)
    tol = 1e-14 if tn == "double" else 2e-6
    assert rel(Aq, g[k + "_Aq"]) < tol


@pytest.mark.parametrize("N", [3, 7])
def test_block_ax_golden(orc, golden, N):
    g = golden("block_ax")
    k = "block_N%d" % N
    Np = (N + 1) ** 3
    L = g[k + "_q"].size // 3
    E = L // Np
    Aq = np.zeros(3 * L)
    orc.lib().orc_block_ax_f64(N, E, 3, L, L, g[k + "_ggeo"], g[k + "_D"].ravel().copy(),
                               g[k + "_lambda0"], g[k + "_lambda1"], g[k + "_q"], Aq)
    assert rel(Aq, g[k + "_Aq"]) < 1e-14


def test_update_pcg_golden(orc, golden):
    g = golden("updatepcg")
    r, rr = orc.update_pcg(g["updatepcg_w"], g["updatepcg_Ap"], float(g["updatepcg_alpha"][0]),
                           g["updatepcg_r"])
    assert rel(r, g["updatepcg_r_out"]) < 1e-15
    assert abs(rr - g["updatepcg_rdotr"][0]) <= 1e-13 * abs(g["updatepcg_rdotr"][0])


@pytest.mark.parametrize("Nf,Nc", [(7, 3), (7, 1), (3, 1), (7, 5), (5, 3), (9, 3), (9, 5), (5, 1)])
def test_transfers_golden(orc, golden, Nf, Nc):
    g = golden("transfer")
    k = "transfer_%d_%d" % (Nf, Nc)
    R = g[k + "_R"]
    assert rel(orc.transfer_R(Nf, Nc), R) < 1e-14   # the oracle's R equals the recorded one
    qc = orc.coarsen(Nf, Nc, R, g[k + "_qf"])
    assert rel(qc, g[k + "_qc"]) < 1e-14
    qf = orc.prolongate(Nf, Nc, R, g[k + "_qc_in"], g[k + "_qf_acc"])
    assert rel(qf, g[k + "_qf_out"]) < 1e-14


@pytest.mark.parametrize("case", ["box", "random", "singletons"])
@pytest.mark.parametrize("T", [np.float64, np.float32, np.int32])
@pytest.mark.parametrize("op", ["add", "mul", "min", "max"])
def test_gs_against_gslib(orc, golden, case, T, op):
    if T == np.int32 and op == "mul":
        pytest.skip("int mul not recorded")
    g = golden("gs")
    ids = g["gs_%s_ids" % case]
    v = g["gs_%s_v_%s" % (case, T.__name__)]
    ref = g["gs_%s_%s_%s" % (case, op, T.__name__)]
    out = orc.Ogs(ids).gs(v, op)
    if T == np.int32 or op in ("min", "max"):
        assert np.array_equal(out, ref)
    else:
        tol = 1e-14 if T == np.float64 else 1e-6
        assert rel(out, ref) < tol
