"""Oracle for the algebraic coarse solve of the order-1 level.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  The reference solves the
order-1 problem with one BoomerAMG V-cycle on its assembled FEM matrix
(src/solvers/elliptic/MG/coarseLevel.cpp:47-222, settings :80-93: 1 cycle, 1
smoothing sweep; matrix from MG/ellipticBuildFEM.cpp).  hypre is not buildable
here, so the library replaces it by smoothed-aggregation AMG, restated here
step by step as the library builds it:

  assembly      element matrices = the level operator applied to unit vectors
                (one per element node), summed into rows numbered by first
                appearance of each unmasked global id
  strength      j strong for i if |a_ij| >= theta sqrt(a_ii a_jj)   (theta = 0.08)
  aggregation   three passes in row order: (1) a node whose strong
                neighbours are all free forms an aggregate with them; (2) a
                free node joins the aggregate of its first aggregated strong
                neighbour; (3) a free node forms an aggregate with its free
                strong neighbours
  prolongation  P = (I - omega D^-1 A) T, T = aggregate indicator with unit
                columns, omega = 4 / (3 rho(D^-1 A)), rho from 20 power steps
                from a fixed splitmix64 vector; for a singular (all-Neumann)
                matrix the columns are the null vector B restricted to the
                aggregate and normalised, B_coarse = the aggregates' norms
                (standard smoothed aggregation), and the last level is the
                pseudo-inverse
  coarse matrix P^T A P; levels until n <= 256 or 12 levels; the last is
                inverted densely (Cholesky)
  V-cycle       pre: 2 sweeps damped Jacobi (omega_s = 4 / (3 rho)), first
                from x = 0; residual; restrict R = P^T; recurse; prolongate and
                add; post: 2 sweeps, so the cycle is symmetric
"""
import numpy as np
import scipy.sparse as sp

from .mg import splitmix_uniform

THETA = 0.08
COARSEST = 256
MAX_LEVELS = 12
SWEEPS = 2
POWER_STEPS = 20


def assemble(L, op_local, rows):
    """Assembled matrix of the level operator on the unmasked unique nodes.
    op_local(v): element-local operator (no gather-scatter, no mask) on an
    E-vector; rows[node] = row number or -1 (masked)."""
    E, Np = L.E, L.Np
    n = int(rows.max()) + 1
    I, J, V = [], [], []
    rr = rows.reshape(E, Np)
    for b in range(Np):
        u = np.zeros(E * Np)
        u.reshape(E, Np)[:, b] = 1.0
        col = op_local(u).reshape(E, Np)          # column b of every element matrix
        for a in range(Np):
            keep = (rr[:, a] >= 0) & (rr[:, b] >= 0)
            I.append(rr[keep, a])
            J.append(rr[keep, b])
            V.append(col[keep, a])
    A = sp.coo_matrix((np.concatenate(V), (np.concatenate(I), np.concatenate(J))), shape=(n, n))
    A = A.tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A


def row_numbers(gids_masked):
    """Row of each node: unique nonzero ids numbered by first appearance, -1 for 0."""
    rows = np.full(gids_masked.size, -1, dtype=np.int64)
    first = {}
    for i, g in enumerate(gids_masked.tolist()):
        if g == 0:
            continue
        r = first.setdefault(g, len(first))
        rows[i] = r
    return rows


def strength(A, theta=THETA):
    A = A.tocsr()
    d = np.abs(A.diagonal())
    rowi = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    keep = (A.indices != rowi) & (np.abs(A.data) >= theta * np.sqrt(d[rowi] * d[A.indices]))
    S = sp.csr_matrix((np.ones(keep.sum()), (rowi[keep], A.indices[keep])), shape=A.shape)
    S.sort_indices()
    return S


def aggregate(S):
    n = S.shape[0]
    agg = np.full(n, -1, dtype=np.int64)
    ip, ix = S.indptr, S.indices
    na = 0
    for i in range(n):                       # pass 1
        nb = ix[ip[i]:ip[i + 1]]
        if agg[i] < 0 and np.all(agg[nb] < 0):
            agg[i] = na
            agg[nb] = na
            na += 1
    first_pass = agg.copy()
    for i in range(n):                       # pass 2
        if agg[i] >= 0:
            continue
        for j in ix[ip[i]:ip[i + 1]]:
            if first_pass[j] >= 0:
                agg[i] = first_pass[j]
                break
    for i in range(n):                       # pass 3
        if agg[i] >= 0:
            continue
        agg[i] = na
        for j in ix[ip[i]:ip[i + 1]]:
            if agg[j] < 0:
                agg[j] = na
        na += 1
    return agg, na


def tentative(agg, na, B=None):
    """T[i, agg[i]] = B_i / ||B on agg[i]|| (B = 1: unit aggregate indicators) and
    the coarse null vector T^T B (the aggregates' norms), so that T B_c = B."""
    n = agg.size
    B = np.ones(n) if B is None else B
    norm = np.sqrt(np.bincount(agg, weights=B * B, minlength=na))
    return sp.csr_matrix((B / norm[agg], (np.arange(n), agg)), shape=(n, na)), norm


def rho_dinv_a(A, seed=7):
    Dinv = 1.0 / A.diagonal()
    v = splitmix_uniform(seed, A.shape[0]) - 0.5
    v /= np.linalg.norm(v)
    lam = 0.0
    for _ in range(POWER_STEPS):
        w = Dinv * (A @ v)
        lam = np.linalg.norm(w)
        if not lam > 0:
            return 1.0
        v = w / lam
    return lam


class AMG:
    """Smoothed-aggregation hierarchy of an SPD CSR matrix."""

    def __init__(self, A, coarsest=COARSEST, null_space=False):
        """null_space: A is singular with the constant vector as null space (an
        all-Neumann problem): the last level is the Moore-Penrose pseudo-inverse."""
        self.levels = []
        # with a null space the constant fine vector is carried down as B
        # (T B_c = B, so every Galerkin matrix keeps B as its null vector) and the
        # dense level is the pseudo-inverse (the null direction cut off);
        # without one B stays 1 on every level (unit aggregate indicators)
        dense_inv = ((lambda M: np.linalg.pinv(M, rcond=1e-10, hermitian=True)) if null_space
                     else np.linalg.inv)
        B = np.ones(A.shape[0])
        A = sp.csr_matrix(A)
        while True:
            lv = {"A": A, "Dinv": 1.0 / A.diagonal(), "B": B}
            self.levels.append(lv)
            n = A.shape[0]
            if n <= coarsest or len(self.levels) == MAX_LEVELS:
                lv["inv"] = dense_inv(A.toarray())
                break
            rho = rho_dinv_a(A)
            lv["omega"] = 4.0 / (3.0 * rho)
            agg, na = aggregate(strength(A))
            if na >= n:                       # no coarsening possible: solve here
                lv["inv"] = dense_inv(A.toarray())
                break
            T, Bc = tentative(agg, na, B if null_space else None)
            B = Bc if null_space else np.ones(na)
            P = (T - lv["omega"] * sp.diags(lv["Dinv"]) @ (A @ T)).tocsr()
            lv["P"] = P
            lv["R"] = P.T.tocsr()
            A = (lv["R"] @ A @ P).tocsr()

    def vcycle(self, b, k=0):
        lv = self.levels[k]
        if "inv" in lv:
            return lv["inv"] @ b
        A, Dinv, om = lv["A"], lv["Dinv"], lv["omega"]
        x = om * Dinv * b
        for _ in range(SWEEPS - 1):
            x = x + om * Dinv * (b - A @ x)
        r = b - A @ x
        x = x + lv["P"] @ self.vcycle(lv["R"] @ r, k + 1)
        for _ in range(SWEEPS):
            x = x + om * Dinv * (b - A @ x)
        return x
