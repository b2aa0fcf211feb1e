"""Curved sides of .re2 meshes (host code of the library, no GPU): the
curve records read back as written (either byte order); elements with
midside-node ('m') edges get the triquadratic Gordon-Hall map of Nek5000's
xyzquad, checked against the oracle's restatement (oracle/re2.py) on rotated
boxes with some edges curved, and both pinned by what the mathematics fixes:
midpoints at the trilinear midpoints change nothing, a deformation in
Q211 + Q121 + Q112 (quadratic along one axis, linear along the others) is
reproduced exactly at every GLL node, and nodes that share a global id share
their coordinates (the curved mesh stays conforming).  Other curve types are
refused."""
import numpy as np
import pytest

import oracle.re2 as r2

# lexicographic end corners of Nek's edges 1..12 (from the 3 x 3 x 3 midpoints)
EDGE_ENDS = []
for (i, j, k) in r2.EDGE_MID:
    ends = []
    for t in (0, 2):
        p = [t if v == 1 else v for v in (i, j, k)]
        ends.append(p[0] // 2 + 2 * (p[1] // 2) + 4 * (p[2] // 2))
    EDGE_ENDS.append(tuple(ends))


@pytest.fixture(scope="module")
def lib():
    import nekrs_amd
    return nekrs_amd


def phi(p, a=0.3, b=-0.2, c=0.25):
    """Q211 + Q121 + Q112 deformation of [0,1]^3 (fixes the faces' planes' corners)."""
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    return np.stack([x + a * x * (1 - x) * (1 + y), y + b * y * (1 - y) * (1 + z),
                     z + c * z * (1 - z) * (1 + x)], -1)


def rotated_box(nx, ny, nz, seed):
    c = r2.box_corners(nx, ny, nz)
    rots = r2.rotations()
    rng = np.random.default_rng(seed)
    for e in range(c.shape[0]):
        c[e] = c[e][rots[rng.integers(0, 24)]]
    return c


def curved_records(corners, fn, edges_of=None):
    recs = []
    for e in range(corners.shape[0]):
        for k, (a, b) in enumerate(EDGE_ENDS):
            if edges_of is not None and k not in edges_of(e):
                continue
            mid = fn(0.5 * (corners[e, a] + corners[e, b]))
            recs.append((e, k + 1, (mid[0], mid[1], mid[2], 0.0, 0.0), "m"))
    return recs


@pytest.mark.parametrize("byteorder", ["<", ">"])
def test_curve_records_round_trip(lib, tmp_path, byteorder):
    c = rotated_box(2, 2, 2, 1)
    rng = np.random.default_rng(2)
    recs = [(int(rng.integers(0, 8)), int(rng.integers(1, 13)), tuple(rng.uniform(-1, 1, 5)), t)
            for t in "mmCsm"]
    path = tmp_path / "c.re2"
    r2.write(path, c, curves=recs, byteorder=byteorder)
    got = lib.re2_curves(path)
    assert lib.read_re2(path)["ncurve"] == len(recs)
    assert list(got["elem"]) == [r[0] for r in recs]
    assert list(got["side"]) == [r[1] for r in recs]
    assert np.array_equal(got["params"], np.array([r[2] for r in recs]))
    assert got["type"] == "mmCsm"
    assert r2.read_curves(path) == [(e, s, tuple(p), t) for e, s, p, t in recs]


@pytest.mark.parametrize("N", [2, 5, 7])
def test_midside_matches_oracle(lib, N):
    """Random midpoints on a random subset of edges of rotated elements."""
    c = rotated_box(3, 2, 2, N)
    rng = np.random.default_rng(10 + N)
    sub = {e: set(rng.choice(12, size=int(rng.integers(1, 13)), replace=False)) for e in range(c.shape[0])}
    recs = curved_records(c, lambda p: p + rng.uniform(-0.04, 0.04, 3), lambda e: sub[e])
    vx, vy, vz = (np.ascontiguousarray(c[:, :, d].ravel()) for d in range(3))
    x, y, z, gid = lib.mesh_from_vertices(N, vx, vy, vz)
    xo, yo, zo = x.copy(), y.copy(), z.copy()
    curves = {"elem": [r[0] for r in recs], "side": [r[1] for r in recs],
              "params": np.array([r[2] for r in recs]), "type": "".join(r[3] for r in recs)}
    lib.mesh_midside(N, vx, vy, vz, curves, x, y, z)
    r2.midside_map(N, c, recs, xo, yo, zo)
    for a, b in ((x, xo), (y, yo), (z, zo)):
        assert np.max(np.abs(a - b)) < 1e-14
    assert np.max(np.abs(x - lib.mesh_from_vertices(N, vx, vy, vz)[0])) > 1e-3   # actually curved


@pytest.mark.parametrize("impl", ["library", "oracle"])
def test_straight_midpoints_change_nothing(lib, impl):
    N = 6
    c = rotated_box(2, 3, 2, 4)
    recs = curved_records(c, lambda p: p)
    vx, vy, vz = (np.ascontiguousarray(c[:, :, d].ravel()) for d in range(3))
    x0, y0, z0, _ = lib.mesh_from_vertices(N, vx, vy, vz)
    x, y, z = x0.copy(), y0.copy(), z0.copy()
    if impl == "library":
        lib.mesh_midside(N, vx, vy, vz, {"elem": [r[0] for r in recs], "side": [r[1] for r in recs],
                                         "params": np.array([r[2] for r in recs]),
                                         "type": "m" * len(recs)}, x, y, z)
    else:
        r2.midside_map(N, c, recs, x, y, z)
    for a, b in ((x, x0), (y, y0), (z, z0)):
        assert np.max(np.abs(a - b)) < 1e-14


@pytest.mark.parametrize("impl", ["library", "oracle"])
@pytest.mark.parametrize("N", [3, 7])
def test_quadratic_deformation_reproduced(lib, impl, N):
    """Corners and edge midpoints taken from phi: every GLL node lands on phi of
    its trilinear position, and shared nodes agree between elements."""
    c = rotated_box(3, 2, 2, 20 + N)
    recs = curved_records(c, phi)
    cp = phi(c)
    vx, vy, vz = (np.ascontiguousarray(cp[:, :, d].ravel()) for d in range(3))
    bx, by, bz = (np.ascontiguousarray(c[:, :, d].ravel()) for d in range(3))
    xl, yl, zl, _ = lib.mesh_from_vertices(N, bx, by, bz)      # trilinear positions in the box
    want = phi(np.stack([xl, yl, zl], -1))
    x, y, z, gid = lib.mesh_from_vertices(N, vx, vy, vz)
    if impl == "library":
        lib.mesh_midside(N, vx, vy, vz, {"elem": [r[0] for r in recs], "side": [r[1] for r in recs],
                                         "params": np.array([r[2] for r in recs]),
                                         "type": "m" * len(recs)}, x, y, z)
    else:
        r2.midside_map(N, cp, recs, x, y, z)
    got = np.stack([x, y, z], -1)
    assert np.max(np.abs(got - want)) < 1e-13
    # conforming: one position per global id
    order = np.argsort(gid, kind="stable")
    g, p = gid[order], got[order]
    same = g[1:] == g[:-1]
    assert np.max(np.abs(p[1:][same] - p[:-1][same])) < 1e-13


def test_other_curve_types_refused(lib):
    c = rotated_box(2, 1, 1, 3)
    vx, vy, vz = (np.ascontiguousarray(c[:, :, d].ravel()) for d in range(3))
    x, y, z, _ = lib.mesh_from_vertices(4, vx, vy, vz)
    cv = {"elem": [0, 1], "side": [1, 2], "params": np.zeros((2, 5)), "type": "mC"}
    with pytest.raises(lib.NrsError) as ex:
        lib.mesh_midside(4, vx, vy, vz, cv, x, y, z)
    assert ex.value.code == -4 and "'C'" in str(ex.value)
    with pytest.raises(ValueError):
        r2.midside_map(4, c, [(0, 1, (0, 0, 0, 0, 0), "C")], x, y, z)


def test_re2mesh_applies_midside_nodes(lib, tmp_path):
    """Re2Mesh reads the records and applies them; the ids are those of the
    straight mesh."""
    N = 5
    c = rotated_box(2, 2, 2, 9)
    recs = curved_records(c, phi)
    cp = phi(c)
    path = tmp_path / "curved.re2"
    r2.write(path, cp, curves=recs)
    r2.write(tmp_path / "straight.re2", cp)
    m = lib.Re2Mesh(path, N)
    s = lib.Re2Mesh(tmp_path / "straight.re2", N)
    assert np.array_equal(m.gid, s.gid)
    xo, yo, zo = s.x.copy(), s.y.copy(), s.z.copy()
    r2.midside_map(N, cp, recs, xo, yo, zo)
    assert np.max(np.abs(m.x - xo)) < 1e-14 and np.max(np.abs(m.z - zo)) < 1e-14
    assert np.max(np.abs(m.x - s.x)) > 1e-3
