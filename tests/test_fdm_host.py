"""Host-only pieces of the Schwarz/FDM setup, called through the C ABI (no
GPU work): the 1D generalized eigenproblem (own Jacobi eigensolver in place
of LAPACK dsygv) and the element lengths, against the oracle (oracle/fdm.py,
itself pinned in test_oracle_fdm.py)."""
import numpy as np
import pytest

import oracle.fdm as F


@pytest.fixture(scope="module")
def nk():
    import nekrs_amd
    return nekrs_amd


def aligned(S, ref):
    """S with each column's sign chosen to agree with the matching column of ref."""
    S = S.copy()
    for m in range(S.shape[1]):
        if np.dot(S[:, m], ref[:, m]) < 0:
            S[:, m] = -S[:, m]
    return S


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7, 9])
@pytest.mark.parametrize("lbc,rbc", [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 1), (2, 2)])
def test_fdm_1d_matches_oracle(nk, orc, N, lbc, rbc):
    _, w = orc.gll(N)
    D = orc.dmatrix(N)
    ll, lm, lr = 0.31, 0.77, 1.9
    S, lam = nk.fdm_1d(N, lbc, rbc, ll, lm, lr)
    So, lamo = F.matrices_1d(D, w, lbc, rbc, ll, lm, lr)
    assert np.allclose(lam, lamo, rtol=1e-11, atol=1e-12 * np.max(lamo))
    nonzero = np.any(So != 0, axis=0)
    assert np.array_equal(nonzero, np.any(S != 0, axis=0))
    assert np.allclose(aligned(S, So)[:, nonzero], So[:, nonzero], rtol=1e-9, atol=1e-10)
    # the pair itself: S^T A S = diag(lam), S^T B S = I on the kept columns
    a = F.stiffness_1d(D, w, lbc, rbc, ll, lm, lr)
    b = F.mass_1d(w, lbc, rbc, ll, lm, lr)
    keep = np.ones(N + 3, bool)
    keep[[r for r in range(N + 3) if np.all(S[r] == 0)]] = False
    Sk = S[:, nonzero]
    Ak = a[np.ix_(keep, keep)]
    Bk = b[np.ix_(keep, keep)]
    assert np.allclose(Sk[keep].T @ Ak @ Sk[keep], np.diag(lam[nonzero]), atol=1e-9 * np.max(lam))
    assert np.allclose(Sk[keep].T @ Bk @ Sk[keep], np.eye(Sk.shape[1]), atol=1e-10)


def test_fdm_1d_rejects_bad_arguments(nk):
    with pytest.raises(nk.NrsError):
        nk.fdm_1d(3, 0, 0, -1.0, 1.0, 1.0)
    with pytest.raises(nk.NrsError):
        nk.fdm_1d(3, 3, 0, 1.0, 1.0, 1.0)


@pytest.mark.parametrize("N", [1, 2, 3, 7])
@pytest.mark.parametrize("mp", ["kershaw", "sine", "none"])
def test_element_lengths_match_oracle(nk, orc, N, mp):
    mo = orc.BoxMesh(4, 3, 2, N, mp, 0.3)
    mg = nk.BoxMesh(4, 3, 2, N, mp, 0.3)
    assert np.array_equal(mg.gid, mo.gid)
    left, mid, right = nk.fdm_element_lengths(mg)
    lo, mo_, ro = F.element_lengths(N, mo.E, mo.x, mo.y, mo.z, mo.w, orc.Ogs(mo.gid))
    for a, b in ((left, lo), (mid, mo_), (right, ro)):
        assert np.allclose(a, b, rtol=1e-12, atol=0)


def test_element_lengths_rotated_elements(nk, orc):
    """Neighbour lengths through face links on a mesh whose elements are
    stored in rotated local orientations (ids and coordinates permuted)."""
    N = 3
    m = nk.BoxMesh(3, 2, 2, N, "kershaw", 0.3)
    rng = np.random.default_rng(5)
    Nq = N + 1
    x, y, z, g = (a.reshape(m.E, Nq, Nq, Nq).copy() for a in (m.x, m.y, m.z, m.gid))
    for e in range(m.E):
        t = rng.integers(0, 4)
        for a in (x, y, z, g):
            a[e] = np.rot90(a[e], k=t, axes=(1, 2))      # rotate about the k axis
    class M:
        pass
    r = M()
    r.N, r.E = N, m.E
    r.x, r.y, r.z, r.gid = (np.ascontiguousarray(a.ravel()) for a in (x, y, z, g))
    left, mid, right = nk.fdm_element_lengths(r)
    lo, mo_, ro = F.element_lengths(N, m.E, r.x, r.y, r.z, orc.gll(N)[1], orc.Ogs(r.gid))
    for a, b in ((left, lo), (mid, mo_), (right, ro)):
        assert np.allclose(a, b, rtol=1e-12, atol=0)
