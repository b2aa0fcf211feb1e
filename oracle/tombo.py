"""Oracle for the pressure right-hand side of the TOMBO splitting.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  A numpy restatement of
tombo::pressureSolve's right-hand side (src/navierStokes/tombo.cpp:11-121)
and of the kernels it calls:

  vgeo        kernels/mesh/geometricFactorsHex3D.okl:27-143 (rx..tz, J, JW)
  sgeo        kernels/mesh/surfaceGeometricFactorsHex3D.okl (n, sJ, WSJ)
  invLMM      src/mesh/meshComputeInvLMM.cpp (1 / gs(JW))
  curl        kernels/nrs/curlHex3D.okl (scaleJW = 1)
  gradient    kernels/nrs/gradientVolumeHex3D.okl
  stress      kernels/nrs/pressureStressHex3D.okl (VELOCITY STRESSFORMULATION)
  rhsU        kernels/nrs/pressureRhsHex3D.okl
  wdiv        kernels/nrs/wDivergenceVolumeHex3D.okl
  addQtl      kernels/nrs/pressureAddQtl.okl
  surface     kernels/nrs/divergenceSurfaceHex3D.okl (bcType ids of src/bdry/bcType.h)

E-vectors are [e][k][j][i]; fields [3][Nlocal].  The reference's OKL kernels
need OCCA to build, so parity rests on the closed forms and identities of
tests/test_oracle_tombo.py.
"""
import numpy as np

from . import Ogs, dmatrix, gll

BC_W, BC_INT, BC_V, BC_SYMX, BC_SYMY, BC_SYMZ, BC_SYM = 1, 2, 3, 4, 5, 6, 7
SYM = (BC_SYMX, BC_SYMY, BC_SYMZ, BC_SYM)


def _d(D, u):
    """(d/dr, d/ds, d/dt) of element fields u[E][k][j][i]."""
    return (np.einsum("im,ekjm->ekji", D, u), np.einsum("jm,ekmi->ekji", D, u),
            np.einsum("km,emji->ekji", D, u))


def _dT(D, a, b, c):
    """D_r^T a + D_s^T b + D_t^T c (the transposed derivatives of wdiv)."""
    return (np.einsum("mi,ekjm->ekji", D, a) + np.einsum("mj,ekmi->ekji", D, b) +
            np.einsum("mk,emji->ekji", D, c))


class Geometry:
    def __init__(self, N, x, y, z, gids):
        self.N, self.Nq = N, N + 1
        Nq = self.Nq
        self.E = x.size // Nq ** 3
        E = self.E
        r, w = gll(N)
        self.w = w
        self.D = dmatrix(N)
        sh = (E, Nq, Nq, Nq)
        X, Y, Z = x.reshape(sh), y.reshape(sh), z.reshape(sh)
        xr, xs, xt = _d(self.D, X)
        yr, ys, yt = _d(self.D, Y)
        zr, zs, zt = _d(self.D, Z)
        J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt)
        Ji = 1.0 / J
        W3 = w[None, :, None, None] * w[None, None, :, None] * w[None, None, None, :]
        self.J = J
        self.JW = J * W3
        self.rx, self.ry, self.rz = (ys * zt - zs * yt) * Ji, -(xs * zt - zs * xt) * Ji, (xs * yt - ys * xt) * Ji
        self.sx, self.sy, self.sz = -(yr * zt - zr * yt) * Ji, (xr * zt - zr * xt) * Ji, -(xr * yt - yr * xt) * Ji
        self.tx, self.ty, self.tz = (yr * zs - zr * ys) * Ji, -(xr * zs - zr * xs) * Ji, (xr * ys - yr * xs) * Ji
        self.ogs = Ogs(gids)
        lmm = self.ogs.gs(self.JW.ravel().copy(), "add")
        self.invLMM = 1.0 / lmm

    def vgeo(self):
        """[E][10][Np] as the library stores it."""
        f = [self.rx, self.ry, self.rz, self.sx, self.sy, self.sz, self.tx, self.ty, self.tz, self.JW]
        return np.stack([a.reshape(self.E, -1) for a in f], axis=1)

    def face(self, f):
        """Index tuple selecting face f's nodes (arrays [E][b][a], face node
        n = a + b*Nq), and the unit normal and WSJ there
        (surfaceGeometricFactorsHex3D)."""
        Nq = self.Nq
        sel = {0: (slice(None), 0, slice(None), slice(None)),        # k = 0: [e][j][i]
               1: (slice(None), slice(None), 0, slice(None)),        # j = 0: [e][k][i]
               2: (slice(None), slice(None), slice(None), Nq - 1),   # i = N: [e][k][j]
               3: (slice(None), slice(None), Nq - 1, slice(None)),   # j = N: [e][k][i]
               4: (slice(None), slice(None), slice(None), 0),        # i = 0: [e][k][j]
               5: (slice(None), Nq - 1, slice(None), slice(None))}[f]   # k = N: [e][j][i]
        g = {0: (-self.tx, -self.ty, -self.tz), 1: (-self.sx, -self.sy, -self.sz),
             2: (self.rx, self.ry, self.rz), 3: (self.sx, self.sy, self.sz),
             4: (-self.rx, -self.ry, -self.rz), 5: (self.tx, self.ty, self.tz)}[f]
        nx, ny, nz = (a[sel] for a in g)
        sJ = np.sqrt(nx * nx + ny * ny + nz * nz)
        nx, ny, nz = nx / sJ, ny / sJ, nz / sJ
        sJ = sJ * self.J[sel]
        # n = a + b Nq with a the faster face index: [e][b][a] after the slicing above
        WSJ = sJ * self.w[None, None, :] * self.w[None, :, None]
        return sel, nx, ny, nz, WSJ

    # ---- kernels
    def _fields(self, U, off):
        L = self.E * self.Nq ** 3
        sh = (self.E, self.Nq, self.Nq, self.Nq)
        return [U[c * off:c * off + L].reshape(sh) for c in range(3)]

    def curl(self, U, off=None):
        """JW curl(U) (curlHex3D, scaleJW = 1); returns [3][Nlocal]."""
        L = self.E * self.Nq ** 3
        u, v, w = self._fields(U, off or L)
        ur, us, ut = _d(self.D, u)
        vr, vs, vt = _d(self.D, v)
        wr, ws, wt = _d(self.D, w)
        uy = self.ry * ur + self.sy * us + self.ty * ut
        uz = self.rz * ur + self.sz * us + self.tz * ut
        vx = self.rx * vr + self.sx * vs + self.tx * vt
        vz = self.rz * vr + self.sz * vs + self.tz * vt
        wx = self.rx * wr + self.sx * ws + self.tx * wt
        wy = self.ry * wr + self.sy * ws + self.ty * wt
        return np.concatenate([((wy - vz) * self.JW).ravel(), ((uz - wx) * self.JW).ravel(),
                               ((vx - uy) * self.JW).ravel()])

    def gradient(self, p):
        """JW grad(p) (gradientVolumeHex3D); returns [3][Nlocal]."""
        pr, ps, pt = _d(self.D, p.reshape(self.E, self.Nq, self.Nq, self.Nq))
        gx = self.JW * (self.rx * pr + self.sx * ps + self.tx * pt)
        gy = self.JW * (self.ry * pr + self.sy * ps + self.ty * pt)
        gz = self.JW * (self.rz * pr + self.sz * ps + self.tz * pt)
        return np.concatenate([gx.ravel(), gy.ravel(), gz.ravel()])

    def stress(self, U, mue, div, off=None):
        """The stress-formulation term JW (w + (-2/3) div grad(mue) / mue) that
        pressureStressHex3D.okl subtracts from the second curl, with
        w = 2 sym(grad U) . grad(mue) / mue; returns [3][Nlocal]."""
        L = self.E * self.Nq ** 3
        sh = (self.E, self.Nq, self.Nq, self.Nq)
        u, v, w = self._fields(U, off or L)
        m = mue.reshape(sh)

        def grad(f):
            fr, fs, ft = _d(self.D, f)
            return (self.rx * fr + self.sx * fs + self.tx * ft,
                    self.ry * fr + self.sy * fs + self.ty * ft,
                    self.rz * fr + self.sz * fs + self.tz * ft)
        dudx, dudy, dudz = grad(u)
        dvdx, dvdy, dvdz = grad(v)
        dwdx, dwdy, dwdz = grad(w)
        mx, my, mz = grad(m)
        inv = 1.0 / m
        mx, my, mz = mx * inv, my * inv, mz * inv
        w1 = mx * 2 * dudx + my * (dudy + dvdx) + mz * (dudz + dwdx)
        w2 = mx * (dudy + dvdx) + my * 2 * dvdy + mz * (dvdz + dwdy)
        w3 = mx * (dudz + dwdx) + my * (dvdz + dwdy) + mz * 2 * dwdz
        d = div.reshape(sh)
        sc = -2.0 / 3.0
        return np.concatenate([(self.JW * (w1 + sc * d * mx)).ravel(), (self.JW * (w2 + sc * d * my)).ravel(),
                               (self.JW * (w3 + sc * d * mz)).ravel()])

    def wdiv(self, U, off=None):
        """Weak divergence sum_d D_d^T (JW (d_d . U)) (wDivergenceVolumeHex3D)."""
        L = self.E * self.Nq ** 3
        u, v, w = self._fields(U, off or L)
        a = self.JW * (self.rx * u + self.ry * v + self.rz * w)
        b = self.JW * (self.sx * u + self.sy * v + self.sz * w)
        c = self.JW * (self.tx * u + self.ty * v + self.tz * w)
        return _dT(self.D, a, b, c).ravel()

    def surface(self, bc, coef, F, U, offU):
        """divergenceSurfaceHex3D's contribution (added to nothing): per face
        node -WSJ coef n.U on V/INT faces, -WSJ n.F on symmetry faces."""
        L = self.E * self.Nq ** 3
        out = np.zeros((self.E, self.Nq, self.Nq, self.Nq))
        Uf = self._fields(U, offU)
        Ff = self._fields(F, L)
        bc = np.asarray(bc).reshape(self.E, 6)
        for f in (0, 5, 1, 3, 2, 4):             # the reference's face order
            sel, nx, ny, nz, WSJ = self.face(f)
            t = bc[:, f][:, None, None]
            fluxV = coef * (nx * Uf[0][sel] + ny * Uf[1][sel] + nz * Uf[2][sel])
            fluxS = nx * Ff[0][sel] + ny * Ff[1][sel] + nz * Ff[2][sel]
            flux = np.where((t == BC_V) | (t == BC_INT), fluxV,
                            np.where(np.isin(t, SYM), fluxS, 0.0))
            out[sel] += -WSJ * flux
        return out.ravel()

    def pressure_rhs(self, bc, Ue, U, BF, mue, irho, div, g0idt, visc=True, off=None, stress=False):
        """tombo::pressureSolve's rhs (tombo.cpp:11-121); stress: VELOCITY
        STRESSFORMULATION = TRUE (tombo.cpp:57-66), the second curl less the
        stress term."""
        L = self.E * self.Nq ** 3
        off = off or L
        rep = lambda a: np.tile(a, 3)
        if visc:
            W = self.curl(Ue, off)
            W = self.ogs.gs(W, "add", nvec=3, stride=L) * rep(self.invLMM)
            NC = self.curl(W)
            if stress:
                NC = NC - self.stress(Ue, mue, div, off)
            gD = self.gradient(div)
            nu = rep(mue * irho)
            B = np.concatenate([BF[c * off:c * off + L] for c in range(3)])
            rhsU = B - nu * (NC - (4.0 / 3.0) * gD)
        else:
            rhsU = np.concatenate([BF[c * off:c * off + L] for c in range(3)])
        rhsU = self.ogs.gs(rhsU.copy(), "add", nvec=3, stride=L) * rep(self.invLMM)
        rhs = self.wdiv(rhsU) + g0idt * self.JW.ravel() * div
        return rhs + self.surface(bc, g0idt, rhsU, U, off)
