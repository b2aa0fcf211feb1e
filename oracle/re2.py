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


def write(path, corners, codes=None, group=None, version="#v002", byteorder="<", params=None):
    """corners: [E][8][3] lexicographic; codes: {(e, file_face 0..5): "W  "};
    params: {(e, file_face): (bc1, bc2)} - for "P  " the connected element
    (1-based) and its file face (1..6), as buf_to_bc reads them."""
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
    buf += struct.pack(byteorder + "d", 0.0)            # no curved sides
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
