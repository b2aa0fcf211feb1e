"""GPU parity of the kernel-level entry points (Ax, block Ax, diagonal, mask,
transfers, geometric factors, gather-scatter) against the oracle and the
golden vectors.  Bars: FP64 Ax <= 1e-12 relative, gather-scatter FP64 bit-exact
against the oracle (same reduction order) and <= 1e-13 relative against gslib,
integer work bit-exact, FP32 <= 2e-5 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NS = [1, 2, 3, 4, 5, 6, 7, 8, 9]


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("N", NS)
@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("kind", ["poisson", "helm_field", "poisson_field", "helm_const"])
def test_ax_vs_oracle(gpu, torch_mod, orc, N, T, kind):
    torch = torch_mod
    rng = np.random.default_rng(100 + N)
    E = 37
    m = orc.BoxMesh(E, 1, 1, N, "kershaw", 0.3)   # real geometry, E not a multiple of the packing
    Np = m.Np
    q = rng.uniform(-1, 1, E * Np)
    poisson = kind.startswith("poisson")
    field = kind.endswith("field")
    l0 = rng.uniform(0.5, 2, E * Np) if field else np.array([1.7])
    l1 = rng.uniform(0, 10, E * Np) if field else np.array([3.1])
    ref = orc.ax(N, m.ggeo.astype(T), m.D.astype(T), q.astype(T),
                 lambda0=l0.astype(T) if field else float(l0[0]),
                 lambda1=l1.astype(T) if field else float(l1[0]), poisson=poisson, dtype=T)
    Aq = torch.zeros(E * Np, dtype=torch.float64 if T == np.float64 else torch.float32, device="cuda")
    gpu.ax(N, dev(torch, m.ggeo.astype(T)), dev(torch, m.D.astype(T).ravel()), dev(torch, q.astype(T)),
           Aq, lambda0=dev(torch, l0.astype(T)), lambda1=dev(torch, l1.astype(T)), poisson=poisson,
           lambda_field=field)
    torch.cuda.synchronize()
    tol = 1e-12 if T == np.float64 else 2e-5
    assert rel(host(Aq), ref) < tol


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 7, 9])
def test_ax_golden_with_element_list(gpu, torch_mod, golden, N):
    torch = torch_mod
    g = golden("ax")
    for tn, T in [("double", np.float64), ("float", np.float32)]:
        k = "poisson_%s_N%d" % (tn, N)
        q = g[k + "_q"]
        Aq = torch.full((q.size,), 7.0, dtype=torch.float64 if T == np.float64 else torch.float32,
                        device="cuda")
        gpu.ax(N, dev(torch, g[k + "_ggeo"]), dev(torch, g[k + "_D"].ravel()), dev(torch, q), Aq,
               lambda0=dev(torch, g[k + "_lambda0"]), element_list=dev(torch, g[k + "_elist"]))
        k2 = "helm_%s_N%d" % (tn, N)
        Ah = torch.zeros_like(Aq)
        gpu.ax(N, dev(torch, g[k2 + "_ggeo"]), dev(torch, g[k2 + "_D"].ravel()), dev(torch, g[k2 + "_q"]),
               Ah, lambda0=dev(torch, g[k2 + "_lambda0"]), lambda1=dev(torch, g[k2 + "_lambda1"]),
               poisson=False, lambda_field=True, element_list=dev(torch, g[k2 + "_elist"]))
        torch.cuda.synchronize()
        tol = 1e-12 if T == np.float64 else 2e-5
        assert rel(host(Aq), g[k + "_Aq"]) < tol
        assert rel(host(Ah), g[k2 + "_Aq"]) < tol


def test_ax_partial_element_list_leaves_others(gpu, torch_mod, orc):
    torch = torch_mod
    N, E = 7, 10
    m = orc.BoxMesh(E, 1, 1, N, "kershaw", 0.3)
    q = np.random.default_rng(0).uniform(-1, 1, m.Nlocal)
    ref = orc.ax(N, m.ggeo, m.D, q)
    elist = np.array([7, 2, 5], dtype=np.int32)
    Aq = torch.full((m.Nlocal,), 3.0, dtype=torch.float64, device="cuda")
    gpu.ax(N, dev(torch, m.ggeo), dev(torch, m.D.ravel()), dev(torch, q), Aq,
           element_list=dev(torch, elist))
    out = host(Aq).reshape(E, -1)
    for e in range(E):
        if e in elist:
            assert rel(out[e], ref.reshape(E, -1)[e]) < 1e-12
        else:
            assert np.all(out[e] == 3.0)


@pytest.mark.parametrize("N", [3, 7])
def test_block_ax_golden(gpu, torch_mod, golden, N):
    torch = torch_mod
    g = golden("block_ax")
    k = "block_N%d" % N
    L = g[k + "_q"].size // 3
    E = L // (N + 1) ** 3
    Aq = torch.zeros(3 * L, dtype=torch.float64, device="cuda")
    gpu.block_ax(N, dev(torch, g[k + "_ggeo"]), dev(torch, g[k + "_D"].ravel()),
                 dev(torch, g[k + "_lambda0"]), dev(torch, g[k + "_lambda1"]), dev(torch, g[k + "_q"]),
                 Aq, L, L, E)
    torch.cuda.synchronize()
    assert rel(host(Aq), g[k + "_Aq"]) < 1e-12


@pytest.mark.parametrize("N", [1, 3, 7, 9])
def test_build_diagonal_vs_oracle(gpu, torch_mod, orc, N):
    torch = torch_mod
    m = orc.BoxMesh(3, 2, 2, N, "kershaw", 0.3)
    l0, l1 = np.array([1.3]), np.array([0.7])
    for poisson in (True, False):
        ref = orc.build_diag(N, m.ggeo, m.D, 1.3, 0.7, poisson)
        d = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
        gpu.build_diagonal(N, dev(torch, m.ggeo), dev(torch, m.D.ravel()), d, dev(torch, l0),
                           dev(torch, l1), poisson=poisson)
        torch.cuda.synchronize()
        assert rel(host(d), ref) < 1e-12


@pytest.mark.parametrize("Nf,Nc", [(7, 3), (7, 1), (3, 1), (7, 5), (5, 3), (9, 3), (9, 5), (5, 1)])
def test_transfers_golden(gpu, torch_mod, golden, Nf, Nc):
    torch = torch_mod
    g = golden("transfer")
    k = "transfer_%d_%d" % (Nf, Nc)
    R = dev(torch, g[k + "_R"].ravel())
    qc = torch.zeros(g[k + "_qc"].size, dtype=torch.float64, device="cuda")
    gpu.coarsen(Nf, Nc, R, dev(torch, g[k + "_qf"]), qc)
    qf = dev(torch, g[k + "_qf_acc"])
    gpu.prolongate(Nf, Nc, R, dev(torch, g[k + "_qc_in"]), qf)
    torch.cuda.synchronize()
    assert rel(host(qc), g[k + "_qc"]) < 1e-13
    assert rel(host(qf), g[k + "_qf_out"]) < 1e-13
    # FP32 path
    qc32 = torch.zeros(qc.numel(), dtype=torch.float32, device="cuda")
    gpu.coarsen(Nf, Nc, R.float(), dev(torch, g[k + "_qf"]).float(), qc32)
    torch.cuda.synchronize()
    assert rel(host(qc32), g[k + "_qc"]) < 2e-5


@pytest.mark.parametrize("N", [1, 3, 7, 9])
@pytest.mark.parametrize("mp", ["kershaw", "sine"])
def test_geometric_factors_vs_oracle(gpu, torch_mod, orc, N, mp):
    torch = torch_mod
    m = orc.BoxMesh(3, 2, 2, N, mp, 0.3)
    ggeo = torch.zeros(m.E * 7 * m.Np, dtype=torch.float64, device="cuda")
    gpu.geometric_factors(N, dev(torch, m.x), dev(torch, m.y), dev(torch, m.z), ggeo)
    out = host(ggeo).reshape(m.E, 7, m.Np)
    ref = m.ggeo.reshape(m.E, 7, m.Np)
    for c in range(7):
        assert np.max(np.abs(out[:, c] - ref[:, c])) <= 1e-12 * np.max(np.abs(ref[:, [0, 2, 5]]))


def _gs_cases(orc):
    rng = np.random.default_rng(11)
    cases = {
        "box_N7": orc.BoxMesh(4, 3, 2, 7, "kershaw", 0.3).gid,
        "box_N1": orc.BoxMesh(5, 4, 3, 1).gid,
        "box_N4": orc.BoxMesh(2, 2, 2, 4).gid,
        "random": rng.integers(0, 500, 5000).astype(np.int64),
        "sparse_ids": (rng.integers(0, 300, 3000) * 1_000_003).astype(np.int64),
        "singletons": np.arange(1, 1001, dtype=np.int64),
        "all_excluded": np.zeros(64, dtype=np.int64),
    }
    # a masked handle: zero out boundary ids of a box mesh
    m = orc.BoxMesh(3, 3, 3, 3)
    mids = orc.mask_ids(m, orc.Ogs(m.gid))
    cases["masked_box"] = orc.masked_gids(m.gid, mids)
    return cases


@pytest.mark.parametrize("op", ["add", "min", "max", "mul"])
def test_gather_scatter_vs_oracle(gpu, torch_mod, orc, op):
    torch = torch_mod
    rng = np.random.default_rng(12)
    for name, ids in _gs_cases(orc).items():
        og_o, og_g = orc.Ogs(ids), gpu.Ogs(ids)
        n = ids.size
        for T, tt in [(np.float64, torch.float64), (np.float32, torch.float32), (np.int32, torch.int32)]:
            if T == np.int32:
                if op == "mul":
                    continue
                v = rng.integers(-1000, 1000, n).astype(np.int32)
            else:
                v = rng.uniform(0.5, 1.5, n).astype(T) if op == "mul" else rng.uniform(-1, 1, n).astype(T)
            ref = og_o.gs(v, op)
            d = dev(torch, v)
            og_g.gather_scatter(d, op)
            torch.cuda.synchronize()
            out = host(d)
            assert np.array_equal(out, ref), (name, T, op)   # same order of reduction: bit exact


def test_gather_scatter_vs_gslib_golden(gpu, torch_mod, golden):
    torch = torch_mod
    g = golden("gs")
    for case in ["box", "random", "singletons"]:
        ids = g["gs_%s_ids" % case]
        og = gpu.Ogs(ids)
        for T in [np.float64, np.float32, np.int32]:
            for op in ["add", "mul", "min", "max"]:
                if T == np.int32 and op == "mul":
                    continue
                v = dev(torch, g["gs_%s_v_%s" % (case, T.__name__)])
                og.gather_scatter(v, op)
                torch.cuda.synchronize()
                ref = g["gs_%s_%s_%s" % (case, op, T.__name__)]
                if T == np.int32 or op in ("min", "max"):
                    assert np.array_equal(host(v), ref)
                else:
                    assert rel(host(v), ref) < (1e-13 if T == np.float64 else 1e-6)


@pytest.mark.parametrize("nvec,op", [(3, "add"), (3, "min"), (3, "mul"), (2, "add"), (4, "max")])
def test_gather_scatter_many_vectors(gpu, torch_mod, orc, nvec, op):
    """Several vectors stride apart in one call (gatherScatterMany): three take
    the kernel form that loads every vector's members before storing, other
    counts the loop over vectors; both bit for bit the oracle, padding untouched."""
    torch = torch_mod
    ids = orc.BoxMesh(3, 2, 2, 5, "kershaw", 0.3).gid
    n = ids.size
    stride = n + 100
    rng = np.random.default_rng(2)
    v = rng.uniform(0.5, 1.5, nvec * stride) if op == "mul" else rng.uniform(-1, 1, nvec * stride)
    og_o, og_g = orc.Ogs(ids), gpu.Ogs(ids)
    ref = og_o.gs(v, op, nvec=nvec, stride=stride)
    d = dev(torch, v)
    og_g.gather_scatter(d, op, nvec=nvec, stride=stride)
    torch.cuda.synchronize()
    assert np.array_equal(host(d), ref)
    for f in range(nvec):                                    # padding untouched
        assert np.array_equal(host(d)[f * stride + n:(f + 1) * stride], v[f * stride + n:(f + 1) * stride])


def test_gather_scatter_start_finish(gpu, torch_mod, orc):
    """oogs::start / finish as ellipticOperator uses them: start, then the
    caller keeps writing v (the second half of the elements' values, as the Ax of
    the local elements does), then finish: equal bit for bit to the one-call
    gather-scatter of the final v, for one and for three vectors; a finish
    without a start, a second start and a finish with other arguments are refused."""
    torch = torch_mod
    ids = orc.BoxMesh(4, 3, 2, 5, "kershaw", 0.3).gid
    n = ids.size
    og = gpu.Ogs(ids)
    rng = np.random.default_rng(21)
    for nvec, stride in ((1, n), (3, n + 64)):
        a = rng.uniform(-1, 1, nvec * stride)
        b = rng.uniform(-1, 1, nvec * stride)
        half = np.zeros(nvec * stride, dtype=bool)
        for f in range(nvec):
            half[f * stride + n // 2: f * stride + n] = True
        final = np.where(half, b, a)
        ref = dev(torch, final)
        og.gather_scatter(ref, "add", nvec=nvec, stride=stride)
        v = dev(torch, a)
        og.start(v, "add", nvec=nvec, stride=stride)
        v[torch.from_numpy(half).cuda()] = dev(torch, b[half])    # the caller's work between the halves
        og.finish(v, "add", nvec=nvec, stride=stride)
        torch.cuda.synchronize()
        assert torch.equal(v, ref), nvec
    v = dev(torch, rng.uniform(-1, 1, n))
    with pytest.raises(gpu.NrsError):
        og.finish(v)                                   # no start
    og.start(v)
    with pytest.raises(gpu.NrsError):
        og.start(v)                                    # start twice
    with pytest.raises(gpu.NrsError):
        og.finish(v, "min")                            # other arguments
    og.finish(v)


def test_inv_degree_and_info(gpu, torch_mod, orc):
    ids = orc.BoxMesh(4, 3, 2, 7, "kershaw", 0.3).gid
    og_o, og_g = orc.Ogs(ids), gpu.Ogs(ids)
    assert np.array_equal(host(og_g.inv_degree()), og_o.invDegree)
    info = og_g.info()
    sizes = np.diff(og_o.offsets)
    assert info["rows"] == int(np.sum(sizes > 1))
    assert info["members"] == int(np.sum(sizes[sizes > 1]))
    assert info["singletons"] == int(np.sum(sizes == 1))


def test_mask_kernel(gpu, torch_mod):
    torch = torch_mod
    v = torch.arange(1000, dtype=torch.float64, device="cuda") + 1
    ids = torch.tensor([0, 5, 999, 500], dtype=torch.int32, device="cuda")
    gpu.mask(ids, v)
    torch.cuda.synchronize()
    out = host(v)
    assert np.all(out[[0, 5, 999, 500]] == 0)
    assert np.count_nonzero(out == 0) == 4
