"""Pins for the Schwarz / FDM oracle (oracle/fdm.py):
  * preFDM / fusedFDM (ASM, RAS) / postFDM against golden vectors recorded
    from the reference's compiled serial kernels (tests/golden/fdm.npz);
  * the 1D stiffness / mass pair equals the principal sub-matrix of a
    three-element 1D SEM assembly (an independent construction);
  * the generalized eigenpairs satisfy S^T A S = Lambda, S^T B S = I;
  * the FDM box solve inverts the separable 3D operator on its active set;
  * on one undeformed element with Dirichlet faces the smoother is the exact
    inverse of the assembled operator (checked through the oracle's Ax);
  * element lengths on a box with x -> x^2 spacing, by hand;
  * extended-mesh mask counts and ASM weights by direct counting."""
import numpy as np
import pytest

import oracle.fdm as F


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def ref_S(g, key, E, nl):
    # stored per element as S[component][eigenvector], row-major
    return np.stack([g[key + s].reshape(E, nl, nl) for s in ("_Sx", "_Sy", "_Sz")])


@pytest.mark.parametrize("nq", [2, 4, 8])
def test_prefdm_golden(golden, nq):
    g = golden("fdm")
    k = "fdm_nq%d" % nq
    w = F.pre_fdm(g[k + "_pre_u"], nq - 1)
    assert np.array_equal(w.ravel(), g[k + "_pre_work1"])


@pytest.mark.parametrize("nq", [2, 4, 8])
def test_fusedfdm_golden(golden, nq):
    g = golden("fdm")
    k = "fdm_nq%d" % nq
    nl, E = nq + 2, 2
    S = ref_S(g, k, E, nl)
    invL = g[k + "_invL"].reshape(E, nl, nl, nl)
    out = F.fused_fdm(g[k + "_w1"].reshape(E, nl, nl, nl), S, invL)
    # ASM writes the whole box; RAS the restricted interior times wts
    assert rel(out.ravel(), g[k + "_asm_Su"]) < 2e-6
    f = slice(1, nl - 1)
    ras = out[:, f, f, f].ravel() * g[k + "_ras_wts"]
    assert rel(ras, g[k + "_ras_Su"]) < 2e-6
    # ASM also leaves its ghost-face results in u for postFDM
    u = g[k + "_asm_u"].reshape(E, nl, nl, nl)
    for gh, _ in F.ghost_faces(nl):
        assert rel(u[gh], out[gh]) < 2e-6


@pytest.mark.parametrize("nq", [2, 4, 8])
def test_postfdm_golden(golden, nq):
    g = golden("fdm")
    k = "fdm_nq%d" % nq
    su = F.post_fdm(g[k + "_post_work2"], g[k + "_post_work1"], g[k + "_ras_wts"], nq - 1)
    assert rel(su, g[k + "_post_Su"]) < 1e-6


def assembled_three(D, w, ll, lm, lr):
    """1D SEM stiffness and (diagonal) mass of three elements of lengths
    ll, lm, lr, assembled on 3N+1 nodes."""
    n = len(w) - 1
    K = np.zeros((3 * n + 1, 3 * n + 1))
    M = np.zeros(3 * n + 1)
    for e, h in enumerate((ll, lm, lr)):
        s = e * n
        K[s:s + n + 1, s:s + n + 1] += (2.0 / h) * (D.T @ np.diag(w) @ D)
        M[s:s + n + 1] += 0.5 * h * w
    return K, M


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7])
@pytest.mark.parametrize("lbc,rbc", [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 1)])
def test_1d_matrices_against_assembly(orc, N, lbc, rbc):
    _, w = orc.gll(N)
    D = orc.dmatrix(N)
    ll, lm, lr = 0.7, 1.3, 0.45
    a = F.stiffness_1d(D, w, lbc, rbc, ll, lm, lr)
    b = F.mass_1d(w, lbc, rbc, ll, lm, lr)
    nl = N + 3
    assert np.allclose(a, a.T, rtol=1e-14, atol=1e-14)
    assert np.count_nonzero(b - np.diag(np.diag(b))) == 0
    K, M = assembled_three(D, w, ll, lm, lr)
    # patch: left element's second-to-last node, own element, right element's second node
    patch = [N - 1] + list(range(N, 2 * N + 1)) + [2 * N + 1]
    # active rows: ghost ends present only for bc 0; Dirichlet ends drop the boundary node
    act = list(range(nl))
    drop = set()
    if lbc != 0:
        drop.add(0)
    if lbc == 1:
        drop.add(1)
    if rbc != 0:
        drop.add(nl - 1)
    if rbc == 1:
        drop.add(nl - 2)
    act = [i for i in act if i not in drop]
    Kp = K[np.ix_(patch, patch)]
    Mp = M[patch]
    if lbc == 2:       # Neumann: no left neighbour, the boundary node keeps only its own element
        Kp[1, 1] -= (2.0 / ll) * (D.T @ np.diag(w) @ D)[N, N]
        Mp[1] -= 0.5 * ll * w[N]
    if rbc == 2:
        Kp[nl - 2, nl - 2] -= (2.0 / lr) * (D.T @ np.diag(w) @ D)[0, 0]
        Mp[nl - 2] -= 0.5 * lr * w[0]
    assert np.allclose(a[np.ix_(act, act)], Kp[np.ix_(act, act)], rtol=1e-13, atol=1e-13)
    assert np.allclose(np.diag(b)[act], Mp[act], rtol=1e-14, atol=0)
    for i in drop:     # identity rows, decoupled
        assert a[i, i] == 1.0 and b[i, i] == 1.0
        assert np.count_nonzero(a[i]) == 1 and np.count_nonzero(a[:, i]) == 1


@pytest.mark.parametrize("N", [1, 3, 7, 9])
@pytest.mark.parametrize("lbc,rbc", [(0, 0), (1, 0), (1, 1), (2, 1)])
def test_generalized_eigenpairs(orc, N, lbc, rbc):
    import scipy.linalg
    _, w = orc.gll(N)
    D = orc.dmatrix(N)
    a = F.stiffness_1d(D, w, lbc, rbc, 0.9, 1.1, 1.4)
    b = F.mass_1d(w, lbc, rbc, 0.9, 1.1, 1.4)
    lam, V = scipy.linalg.eigh(a, b, type=1, lower=False, driver="gv")
    assert np.allclose(V.T @ a @ V, np.diag(lam), atol=1e-10 * np.max(lam))
    assert np.allclose(V.T @ b @ V, np.eye(len(lam)), atol=1e-12)
    assert np.all(np.diff(lam) >= 0)
    S, lam2 = F.matrices_1d(D, w, lbc, rbc, 0.9, 1.1, 1.4)
    assert np.array_equal(lam, lam2)
    nl = N + 3
    zero = ([0] if lbc > 0 else []) + ([1] if lbc == 1 else []) + \
           ([nl - 1] if rbc > 0 else []) + ([nl - 2] if rbc == 1 else [])
    for r in zero:
        assert np.all(S[r] == 0)


@pytest.mark.parametrize("N", [2, 3, 7])
@pytest.mark.parametrize("bcs", [(0, 0, 0, 0, 0, 0), (1, 0, 0, 1, 1, 1), (1, 1, 1, 1, 1, 1),
                                 (2, 0, 1, 0, 0, 2)])
def test_fdm_box_inverts_separable_operator(orc, N, bcs):
    _, w = orc.gll(N)
    D = orc.dmatrix(N)
    lens = [(0.5, 0.8, 1.1), (1.0, 0.6, 0.6), (0.9, 1.2, 0.7)]
    A, B, S, lam, act = [], [], [], [], []
    nl = N + 3
    for d in range(3):
        lbc, rbc = bcs[2 * d], bcs[2 * d + 1]
        A.append(F.stiffness_1d(D, w, lbc, rbc, *lens[d]))
        B.append(F.mass_1d(w, lbc, rbc, *lens[d]))
        s, l = F.matrices_1d(D, w, lbc, rbc, *lens[d])
        S.append(s[None])
        lam.append(l)
        a = np.ones(nl, bool)
        for r in ([0] if lbc > 0 else []) + ([1] if lbc == 1 else []) + \
                 ([nl - 1] if rbc > 0 else []) + ([nl - 2] if rbc == 1 else []):
            a[r] = False
        act.append(a)
    K3 = (np.kron(B[2], np.kron(B[1], A[0])) + np.kron(B[2], np.kron(A[1], B[0])) +
          np.kron(A[2], np.kron(B[1], B[0])))
    diag = lam[2][:, None, None] + lam[1][None, :, None] + lam[0][None, None, :]
    invL = np.where(diag > F.EPS_DIAG, 1.0 / diag, 0.0)[None]
    r = np.random.default_rng(N).uniform(-1, 1, (1, nl, nl, nl))
    u = F.fdm_box(r, np.stack(S), invL, T=np.float64)
    mask = (act[2][:, None, None] & act[1][None, :, None] & act[0][None, None, :]).ravel()
    Ka = K3[np.ix_(mask, mask)]
    ref = np.linalg.solve(Ka, r.ravel()[mask])
    assert rel(u.ravel()[mask], ref) < 1e-11
    assert np.all(u.ravel()[~mask] == 0)


@pytest.mark.parametrize("N", [2, 3, 5, 7])
@pytest.mark.parametrize("mode", ["ras", "asm"])
def test_single_element_exact_inverse(orc, N, mode):
    m = orc.BoxMesh(1, 1, 1, N)
    el = orc.Elliptic(m)
    fine = (m.N, m.x, m.y, m.z, m.w, el.ogs_full)
    S = F.Schwarz(m, fine, el.maskIds, el.ogs, el.invDegree, mode=mode, T=np.float64)
    r = np.random.default_rng(N).uniform(-1, 1, m.Nlocal)
    r[el.maskIds] = 0
    z = S.apply(r)
    act = np.ones(m.Nlocal, bool)
    act[el.maskIds] = False
    assert rel(el.operator(z)[act], r[act]) < 1e-12
    assert np.all(z[el.maskIds] == 0)


def test_element_lengths_by_hand(orc):
    nx, ny, nz, N = 4, 3, 2, 4
    m = orc.BoxMesh(nx, ny, nz, N)
    m.x = m.x ** 2                               # planes stay planes: width = difference of squares
    m.z = 0.5 * m.z
    ogs = orc.Ogs(m.gid)
    left, mid, right = F.element_lengths(N, m.E, m.x, m.y, m.z, m.w, ogs)
    for e in range(m.E):
        ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
        wx = lambda i: ((i + 1) ** 2 - i ** 2) / nx ** 2        # noqa: E731
        assert abs(mid[0, e] - wx(ex)) < 1e-13
        assert abs(mid[1, e] - 1.0 / ny) < 1e-13
        assert abs(mid[2, e] - 0.5 / nz) < 1e-13
        assert abs(left[0, e] - (wx(ex - 1) if ex > 0 else wx(ex))) < 1e-13
        assert abs(right[0, e] - (wx(ex + 1) if ex < nx - 1 else wx(ex))) < 1e-13
        assert abs(left[1, e] - 1.0 / ny) < 1e-13 and abs(right[2, e] - 0.5 / nz) < 1e-13


def test_element_lengths_order_one(orc):
    m = orc.BoxMesh(3, 2, 2, 1)
    ogs = orc.Ogs(m.gid)
    left, mid, right = F.element_lengths(1, m.E, m.x, m.y, m.z, m.w, ogs)
    assert np.allclose(mid[0], 1 / 3) and np.allclose(mid[1], 0.5)
    assert np.array_equal(left, mid) and np.array_equal(right, mid)


@pytest.mark.parametrize("N", [1, 3, 4])
def test_extended_mask_counts(orc, N):
    m = orc.BoxMesh(3, 2, 2, N)
    me = orc.BoxMesh(3, 2, 2, N + 2)
    masked = np.zeros(me.Nlocal, bool)
    masked[F.extended_mask(me)] = True
    Nq = N + 1
    per = masked.reshape(m.E, -1)
    for e in range(m.E):
        ndir = int(np.sum(m.EToB[6 * e:6 * e + 6] == 1))
        # unmasked: the Nq^3 interior plus the Nq^2 face interiors of non-Dirichlet faces
        assert np.sum(~per[e]) == Nq ** 3 + (6 - ndir) * Nq ** 2


@pytest.mark.parametrize("N", [2, 3, 5])
def test_asm_weights_by_counting(orc, N):
    m = orc.BoxMesh(3, 2, 2, N)
    el = orc.Elliptic(m)
    fine = (m.N, m.x, m.y, m.z, m.w, el.ogs_full)
    S = F.Schwarz(m, fine, el.maskIds, el.ogs, el.invDegree, mode="asm", T=np.float64)
    Nq = N + 1
    c = np.ones((m.E, Nq, Nq, Nq))
    k, j, i = np.meshgrid(np.arange(Nq), np.arange(Nq), np.arange(Nq), indexing="ij")
    # face f has a neighbour when it is not a boundary face; the overlap lands on
    # the element's own layer 1 (left faces) or N-1 (right faces)
    layer = {4: (i, 1), 2: (i, N - 1), 1: (j, 1), 3: (j, N - 1), 0: (k, 1), 5: (k, N - 1)}
    for e in range(m.E):
        for f, (idx, l) in layer.items():
            if m.EToB[6 * e + f] == 0:
                c[e] += (idx == l)
    ref = 1.0 / el.ogs.gs(c.ravel(), "add")
    assert rel(S.wts, ref) < 1e-15
