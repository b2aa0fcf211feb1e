"""Oracle for the projection initial guess (SolutionProjection).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Restates, in numpy,
src/solvers/elliptic/ellipticSolutionProjection.cpp:

  pre(r)    :269-280 + computePreProjection :141-192: from the projStart-th
            solve on, alpha_k = <xx_k, r>_W, xbar = sum alpha_k xx_k, and
            r -= A xbar (ACONJ) or r -= sum alpha_k bb_k (CLASSIC, bb_k = A xx_k)
  post(x)   :282-288 + computePostProjection :194-235: append the correction
            x as a new basis vector (or restart the basis from x + xbar when
            it is full), x += xbar, bb = A xx_new, then updateProjectionSpace
            (:37-139): A-orthogonalise the new vector against the basis with
            one pass of coefficients alpha_k = <xx_k, A xx_new>_W, drop it when
            sqrt(<xx_new, A xx_new> - sum alpha_k^2) / <xx_new, A xx_new> <= 1e-7
            (the reference's test, kept as written), else scale it to unit
            A-norm; a dropped vector resets the basis to x.

W is invDegree (repeated per field for block vectors), so <a, b>_W is the
inner product of the assembled vectors.  The driver `solve` restates the part
of ellipticSolve around it (ellipticSolve.cpp:91-186).
"""
import numpy as np

PREVIOUS, CLASSIC, ACONJ = 0, 1, 2
DROP_TOL = 1e-7


class SolutionProjection:
    def __init__(self, operator, weight, kind=ACONJ, max_vecs=8, start=5):
        self.A = operator            # masked, assembled operator on (block) vectors
        self.w = weight
        self.kind = kind
        self.max_vecs = max_vecs
        self.start = start
        self.xx = []                 # basis (A-orthonormal)
        self.bb = []                 # CLASSIC: A xx_k; ACONJ: [A xbar or A xx_new]
        self.xbar = None
        self.step = 0
        self.prev = 0
        self.dropped = 0             # vectors discarded as dependent (diagnostic)

    def dot(self, a, b):
        return float(np.sum(self.w * a * b))

    def pre(self, r):
        """Returns the projected residual (a new array) and whether it ran."""
        self.step += 1
        if self.step < self.start or not self.xx:
            return r, False
        self.prev = len(self.xx)
        alpha = [self.dot(v, r) for v in self.xx]
        self.xbar = sum(a * v for a, v in zip(alpha, self.xx))
        if self.kind == CLASSIC:
            return r - sum(a * b for a, b in zip(alpha, self.bb)), True
        return r - self.A(self.xbar), True

    def _update_space(self):
        m = len(self.xx)
        new = self.xx[-1]
        bnew = self.bb[-1] if self.kind == CLASSIC else self.bb[0]
        alpha = [self.dot(v, bnew) for v in self.xx]
        norm_orig = alpha[-1]
        for k in range(m - 1):
            new = new - alpha[k] * self.xx[k]
            if self.kind == CLASSIC:
                bnew = bnew - alpha[k] * self.bb[k]
        sum_alpha = sum(a * a for a in alpha[:-1])
        with np.errstate(invalid="ignore", divide="ignore"):
            norm_new = np.sqrt(np.float64(norm_orig - sum_alpha))
            test = norm_new / np.float64(norm_orig)
        if test > DROP_TOL:
            self.xx[-1] = new / norm_new
            if self.kind == CLASSIC:
                self.bb[-1] = bnew / norm_new
        else:
            self.xx.pop()
            if self.kind == CLASSIC:
                self.bb.pop()
            self.dropped += 1

    def post(self, x):
        """x: the solver's correction; returns the full correction x + xbar."""
        if self.step < self.start:
            return x
        if not self.xx:
            self.xx = [x.copy()]
        elif len(self.xx) == self.max_vecs:
            x = x + self.xbar
            self.xx = [x.copy()]
            if self.kind == CLASSIC:
                self.bb = []
        else:
            self.xx.append(x.copy())
            x = x + self.xbar
        before = len(self.xx)
        self._apply_new()
        self._update_space()
        if len(self.xx) < before:
            self.xx = [x.copy()]
            if self.kind == CLASSIC:
                self.bb = []
            self._apply_new()
            self._update_space()
        return x

    def _apply_new(self):
        Ax = self.A(self.xx[-1])
        if self.kind == CLASSIC:
            self.bb = self.bb[:len(self.xx) - 1] + [Ax]
        else:
            self.bb = [Ax]


def residual(el, rhs, x0, nfields=1):
    """r = gs(mask(rhs - A_local x0)) per field (ellipticSolve.cpp:91-103)."""
    from . import ax
    m = el.mesh
    L = m.Nlocal
    out = []
    for f in range(nfields):
        sl = slice(f * L, (f + 1) * L)
        r = rhs[sl] - ax(m.N, m.ggeo, m.D, x0[sl], el.lambda0, el.lambda1, el.poisson)
        r[el.maskIds] = 0
        out.append(el.ogs.gs(r, "add"))
    return np.concatenate(out)


def solve(el, proj, rhs, x0, inner, nfields=1):
    """One ellipticSolve with the projection: inner(r) -> (correction, stats)
    solves A dx = r from dx = 0 with a tolerance relative to the projected
    residual.  Returns (x, stats) with stats["res00"], ["proj_vectors"]."""
    L = el.mesh.Nlocal * nfields
    w = np.tile(el.invDegree, nfields)
    r = residual(el, rhs, x0, nfields)
    res00 = np.sqrt(np.sum(w * r * r) / el.volume)
    ran = False
    if proj is not None:
        r, ran = proj.pre(r)
    dx, st = inner(r)
    if proj is not None:
        dx = proj.post(dx)
    st = dict(st)
    st["res00"] = res00
    st["proj_vectors"] = proj.prev if ran else 0
    assert dx.size == L
    return x0 + dx, st
