"""MULTIGRID SCHEDULE parsing (nrs_parse_mg_schedule, host code of the library,
no GPU) against the oracle's restatement of parseMultigridSchedule
(MG/parseMultigridSchedule.cpp:7-131) and the level setup's reading of it
(MG/ellipticMultiGridLevelSetup.cpp:160-167): hand-worked known answers from
those rules (the degree before restriction comes from the schedule's up leg),
the reference's own error cases, and 400 random schedules on which the library
and the oracle must agree exactly (levels, degrees, valid or not)."""
import numpy as np
import pytest

import oracle.mg as omg


@pytest.fixture(scope="module")
def lib():
    import nekrs_amd
    return nekrs_amd


def parse(lib, s, N, default=3):
    try:
        o = lib.parse_mg_schedule(s, N, default)
    except lib.NrsError as e:
        return None, str(e)
    n = o.nLevels
    return (list(o.levels[:n]), list(o.preDegree[:n]), list(o.postDegree[:n])), ""


KNOWN = [
    # (schedule, N, default) -> (levels, pre, post)
    ("p=7+degree=2, p=3+degree=2, p=1, p=3+degree=2, p=7+degree=2", 7, 3,
     ([7, 3, 1], [2, 2, 0], [2, 2, 0])),
    # kershaw.par:20 (commented there): degree 3 throughout
    ("p=7+degree=3, p=3+degree=3, p=1, p=3+degree=3, p=7+degree=3", 7, 2,
     ([7, 3, 1], [3, 3, 0], [3, 3, 0])),
    # asymmetric: before restriction the up leg's degree, after prolongation the down leg's
    ("p=7+degree=2, p=3+degree=1, p=1, p=3+degree=4, p=7+degree=5", 7, 3,
     ([7, 3, 1], [5, 4, 0], [2, 1, 0])),
    # entries without a degree take the default
    ("p=7, p=5, p=3+degree=4, p=1, p=3, p=5+degree=1, p=7", 7, 3,
     ([7, 5, 3, 1], [3, 1, 3, 0], [3, 3, 4, 0])),
    # white space anywhere, empty entries dropped
    ("  p = 7 + degree = 2 ,, p=1 ,p=7+degree=2  ", 7, 3, ([7, 1], [2, 0], [2, 0])),
    # a coarsest order above 1
    ("p=9+degree=2, p=5+degree=3, p=3, p=5+degree=3, p=9+degree=2", 9, 3,
     ([9, 5, 3], [2, 3, 0], [2, 3, 0])),
]


@pytest.mark.parametrize("case", KNOWN, ids=[str(i) for i in range(len(KNOWN))])
def test_known_answers(lib, case):
    s, N, d, want = case
    got, err = parse(lib, s, N, d)
    assert err == "" and got == want, (got, err)
    lv, pre, post, errs = omg.parse_mg_schedule(s, N, d)
    assert errs == [] and (lv, pre, post) == want


ERRORS = [
    ("p=7+degree=2, p=3, p=1, p=7+degree=2", 7, "down leg and up leg orders must be identical"),
    ("p=5+degree=2, p=1, p=5+degree=2", 7, "does not match polynomial degree 7"),
    ("p=7+smoother=2, p=1, p=7", 7, "unknown token"),
    ("degree=2, p=1", 1, "order not specified"),
    ("p=7+degree=0, p=1, p=7", 7, "outside 1..16"),
    ("p=7+degree=2, p=-1, p=7", 7, "< 1"),
    # as in the reference, 'sweeps' holds a 'p' and so reads as an order
    ("p=7+sweeps=2, p=1, p=7", 7, "down leg and up leg orders must be identical"),
    ("", 7, "no levels"),
]


def coarse_degree(s):
    """True when an entry of the coarsest order carries a degree token."""
    ent = [[t for t in e.replace(" ", "").split("+") if t] for e in s.replace(" ", "").split(",") if e]
    orders = [omg._stoi(t) for e in ent for t in e if "p" in t]
    if not orders:
        return False
    m = min(orders)
    return any(any("p" in t and omg._stoi(t) == m for t in e) and any("degree" in t for t in e)
               for e in ent)


def test_coarsest_degree_refused(lib):
    """A degree on the coarsest order is refused by the library.  The reference's
    parser accepts it and its level setup (ellipticMultiGridLevelSetup.cpp:162-167)
    would make it the coarsest level's degree after prolongation while std::map's
    operator[] sets the one before restriction to 0; the library keeps the coarsest
    level's own smoothing and says so instead of silently ignoring the entry."""
    s = "p=7+degree=2, p=3+degree=2, p=1+degree=5, p=3+degree=2, p=7+degree=2"
    got, err = parse(lib, s, 7)
    assert got is None and "coarsest order 1" in err, err
    lv, pre, post, errs = omg.parse_mg_schedule(s, 7, 3)     # the reference's parser accepts it
    assert errs == [] and lv == [7, 3, 1]
    assert coarse_degree(s) and not coarse_degree(s.replace("p=1+degree=5", "p=1"))
    assert parse(lib, s.replace("p=1+degree=5", "p=1"), 7)[0] == ([7, 3, 1], [2, 2, 0], [2, 2, 0])


@pytest.mark.parametrize("case", ERRORS, ids=[str(i) for i in range(len(ERRORS))])
def test_errors(lib, case):
    s, N, msg = case
    got, err = parse(lib, s, N)
    assert got is None and msg in err, err
    _, _, _, errs = omg.parse_mg_schedule(s, N, 3)
    assert any(msg in e for e in errs), errs


def random_schedule(rng):
    N = int(rng.integers(2, 10))
    k = int(rng.integers(1, 5))
    orders = sorted(set(int(v) for v in rng.integers(1, N, k)) | {N}, reverse=True)
    coarse = orders[-1]
    deg = lambda: "+degree=%d" % rng.integers(0, 18) if rng.random() < 0.7 else ""
    down = ["p=%d%s" % (o, deg()) for o in orders[:-1]]
    up = ["p=%d%s" % (o, deg()) for o in orders[-2::-1]]
    mid = ["p=%d%s" % (coarse, deg() if rng.random() < 0.3 else "")]
    ent = down + mid + up
    r = rng.random()
    if r < 0.1 and up:
        ent.pop()                                     # up leg shorter
    elif r < 0.15:
        ent.insert(int(rng.integers(0, len(ent) + 1)), "q=3")   # unknown token
    elif r < 0.2:
        N = N + 1                                     # finest != N
    return ", ".join(ent), N


def test_random_schedules_agree_with_oracle(lib):
    rng = np.random.default_rng(11)
    valid = 0
    for _ in range(400):
        s, N = random_schedule(rng)
        d = int(rng.integers(1, 9))
        got, err = parse(lib, s, N, d)
        lv, pre, post, errs = omg.parse_mg_schedule(s, N, d)
        if not errs and coarse_degree(s):
            assert got is None and "coarsest order" in err, (s, err)
            continue
        assert (got is None) == bool(errs), (s, N, err, errs)
        if got is not None:
            valid += 1
            assert got == (lv, pre, post), (s, got, (lv, pre, post))
    assert valid > 100
