""".re2 mesh files for tests.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  A numpy restatement of
the binary layout read by 3rd_party/nek5000/core/reader_re2.f (readp_re2_mesh
:68-157, buf_to_xyz :430-471, readp_re2_curve, readp_re2_bc / buf_to_bc
:503-541) and by re2::nelg (src/io/re2Reader.cpp:3-41): an 80-byte header,
the float 6.54321, element records [group, x[8], y[8], z[8]] with the corners
in the preprocessor's cyclic order, the curve count and records, then the
boundary count and records [element, face, bc[5], code] in 8-byte words.
`write` makes test meshes (any corner order, any codes); `read` reads them
back independently of the library.
"""
import struct

import numpy as np

CYCLIC_TO_LEX = [0, 1, 3, 2, 4, 5, 7, 6]      # meshNekReader.cpp:30
FILE_FACE_TO_SOLVER = [1, 2, 3, 4, 0, 5]      # meshNekReader.cpp:31
# corners (lexicographic i + 2j + 4k) of each file face 1..6 (s-, r+, s+, r-, t-, t+)
FILE_FACE_CORNERS = [(0, 1, 4, 5), (1, 3, 5, 7), (2, 3, 6, 7), (0, 2, 4, 6), (0, 1, 2, 3), (4, 5, 6, 7)]


def write(path, corners, codes=None, group=None, version="#v002", byteorder="<", params=None,
          curves=None):
    """corners: [E][8][3] lexicographic; codes: {(e, file_face 0..5): "W  "};
    params: {(e, file_face): (bc1, bc2)} - for "P  " the connected element
    (1-based) and its file face (1..6), as buf_to_bc reads them; curves: list
    of (e, side 1..12, (c1..c5), type char) curved-side records."""
    corners = np.asarray(corners, dtype=np.float64)
    E = corners.shape[0]
    codes = codes or {}
    hdr = "%5s %9d %1d %9d hdr" % (version, E, 3, E)
    buf = bytearray(hdr.ljust(80).encode())
    buf += struct.pack(byteorder + "f", 6.54321)
    for e in range(E):
        g = 0 if group is None else group[e]
        cyc = corners[e][CYCLIC_TO_LEX]                 # the involution maps both ways
        rec = [float(g)] + list(cyc[:, 0]) + list(cyc[:, 1]) + list(cyc[:, 2])
        buf += struct.pack(byteorder + "25d", *rec)
    curves = curves or []
    buf += struct.pack(byteorder + "d", float(len(curves)))
    for e, sd, cp, t in curves:
        buf += struct.pack(byteorder + "7d", float(e + 1), float(sd), *[float(v) for v in cp])
        buf += t.ljust(8).encode()[:8]
    buf += struct.pack(byteorder + "d", float(len(codes)))
    for (e, f), c in sorted(codes.items()):
        p1, p2 = (params or {}).get((e, f), (0.0, 0.0))
        buf += struct.pack(byteorder + "7d", float(e + 1), float(f + 1), float(p1), float(p2), 0, 0, 0)
        buf += c.ljust(8).encode()[:8]
    with open(path, "wb") as fh:
        fh.write(bytes(buf))


def read(path):
    """(corners [E][8][3] lexicographic, codes [E][6] in the solver's face order)."""
    data = open(path, "rb").read()
    hdr = data[:80].decode()
    E = int(hdr[5:15])
    bo = "<" if abs(struct.unpack("<f", data[80:84])[0] - 6.54321) < 1e-5 else ">"
    off = 84
    corners = np.zeros((E, 8, 3))
    for e in range(E):
        rec = struct.unpack(bo + "25d", data[off:off + 200])
        off += 200
        cyc = np.array(rec[1:]).reshape(3, 8).T
        corners[e] = cyc[CYCLIC_TO_LEX]
    ncurve = int(struct.unpack(bo + "d", data[off:off + 8])[0])
    off += 8 + 64 * ncurve
    nbc = int(struct.unpack(bo + "d", data[off:off + 8])[0])
    off += 8
    codes = np.full((E, 6), "   ", dtype="<U3")
    for _ in range(nbc):
        eg, f = struct.unpack(bo + "2d", data[off:off + 16])
        c = data[off + 56:off + 59].decode()
        codes[int(eg) - 1, FILE_FACE_TO_SOLVER[int(f) - 1]] = c
        off += 64
    return corners, codes


def read_curves(path):
    """Curved-side records [(e 0-based, side, (c1..c5), type)] (buf_to_curve,
    3rd_party/nek5000/core/reader_re2.f:473-498), read independently of the library."""
    data = open(path, "rb").read()
    E = int(data[:80].decode()[5:15])
    bo = "<" if abs(struct.unpack("<f", data[80:84])[0] - 6.54321) < 1e-5 else ">"
    off = 84 + 200 * E
    n = int(struct.unpack(bo + "d", data[off:off + 8])[0])
    off += 8
    out = []
    for _ in range(n):
        rec = struct.unpack(bo + "7d", data[off:off + 56])
        out.append((int(rec[0]) - 1, int(rec[1]), tuple(rec[2:7]), data[off + 56:off + 57].decode()))
        off += 64
    return out


# the midpoint of Nek's edges 1..12 on the 3 x 3 x 3 points (i, j, k) in {0, 1, 2}
# (xyzquad's eindx, 3rd_party/nek5000/core/genxyz.f:1346-1349, 1-based i + 3j + 9k + 1)
EDGE_MID = [(1, 0, 0), (2, 1, 0), (1, 2, 0), (0, 1, 0), (1, 0, 2), (2, 1, 2), (1, 2, 2), (0, 1, 2),
            (0, 0, 1), (2, 0, 1), (2, 2, 1), (0, 2, 1)]


def midside_map(N, corners, curves, x, y, z):
    """Restates Nek5000's xyzquad (genxyz.f:1334-1410, from linquad :1235-1267)
    for elements with midside-node ('m') edges: the trilinear map on the
    3 x 3 x 3 points {-1, 0, 1}^3, the 'm' midpoints put in, vertex + edge
    Gordon-Hall extension (gh_face_extend_3d with gh_type = 2, navier5.f:2501-2590),
    then quadratic interpolation to the GLL nodes.  x, y, z [E][(N+1)^3] are
    overwritten in place for those elements; order 1 keeps the trilinear map."""
    from oracle import gll
    corners = np.asarray(corners, dtype=np.float64)
    byelem = {}
    for e, sd, cp, t in curves:
        if t == "m":
            byelem.setdefault(e, []).append((sd, cp))
        elif t.strip():
            raise ValueError("curve type %r not restated" % t)
    if N == 1:
        return
    zg = np.array([-1.0, 0.0, 1.0])
    lin = np.stack([0.5 * (1 - zg), 0.5 * (1 + zg)])          # [end][point]
    r, _ = gll(N)
    Lq = np.stack([0.5 * r * (r - 1), 1 - r * r, 0.5 * r * (r + 1)], 1)   # [gll][3]
    Nq = N + 1
    for e, lst in byelem.items():
        cx = corners[e].reshape(2, 2, 2, 3)                   # [k][j][i][xyz] (lexicographic)
        q = np.einsum("ai,bj,ck,cbad->kjid", lin, lin, lin, cx)   # [k][j][i][xyz] on {-1,0,1}^3
        for sd, cp in lst:
            i, j, k = EDGE_MID[sd - 1]
            q[k, j, i] = cp[:3]
        # vertex interpolant, then the edge corrections along x, y and z edges
        v = np.einsum("ai,bj,ck,cbad->kjid", lin, lin, lin, q[::2, ::2, ::2])
        dq = q - v
        g = v.copy()
        g += np.einsum("bj,ck,cbid->kjid", lin, lin, dq[::2, ::2, :])
        g += np.einsum("ai,ck,cjad->kjid", lin, lin, dq[::2, :, ::2])
        g += np.einsum("ai,bj,kbad->kjid", lin, lin, dq[:, ::2, ::2])
        out = np.einsum("ia,jb,kc,cbad->kjid", Lq, Lq, Lq, g)   # [k][j][i][xyz] at GLL
        for d, arr in enumerate((x, y, z)):
            arr[e * Nq ** 3:(e + 1) * Nq ** 3] = out[..., d].ravel()


def box_corners(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """Corners [E][8][3] of an nx x ny x nz box, elements x fastest."""
    xs = [np.linspace(lo[a], hi[a], n + 1) for a, n in enumerate((nx, ny, nz))]
    out = np.zeros((nx * ny * nz, 8, 3))
    e = 0
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for c in range(8):
                    a, b, d = c & 1, (c >> 1) & 1, (c >> 2) & 1
                    out[e, c] = (xs[0][i + a], xs[1][j + b], xs[2][k + d])
                e += 1
    return out


def rotations():
    """The 24 proper rotations of the cube as permutations of the lexicographic
    corners: new corner c takes old corner perm[c]."""
    pts = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)]) * 2 - 1
    perms = []
    for P in _rotation_matrices():
        img = pts @ P.T
        perm = [int(np.where((pts == img[c]).all(1))[0][0]) for c in range(8)]
        perms.append(perm)
    return perms


def _rotation_matrices():
    import itertools
    mats = []
    for p in itertools.permutations(range(3)):
        for s in itertools.product((1, -1), repeat=3):
            M = np.zeros((3, 3))
            for r in range(3):
                M[r, p[r]] = s[r]
            if np.linalg.det(M) > 0:
                mats.append(M)
    return mats
