"""GPU parity of the drop-in kernels against the oracle on identical inputs.

Bar: bit-exact for indices, counts, rho and the push (whose arithmetic follows
the reference NumPy path), 1e-9 relative for Poisson against SciPy SuperLU.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import kernels as K
from oracle.poisson import SuperLUPoisson, efield

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from thruster_cx_code_amd import _lib
    return _lib.load()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _random_state(seed, n, nx, ny, dx, dy, vscale=3e4, pad=0.1):
    r = np.random.default_rng(seed)
    x = r.uniform(-pad, (nx - 1) * dx + pad, n).astype(np.float32)
    y = r.uniform(-pad, (ny - 1) * dy + pad, n).astype(np.float32)
    v = [r.normal(0, vscale, n).astype(np.float32) for _ in range(3)]
    # exact half-cell ties and edges
    x[:4] = np.float32([0.0, dx * 0.5, dx * 1.5, (nx - 1) * dx])
    return [x, y] + v


@pytest.mark.parametrize("n_rep,with_b", [(1, False), (7, False), (3, True)])
def test_push_bit_exact(lib, n_rep, with_b):
    print(f"[gpu] push n_rep={n_rep} with_b={with_b}", flush=True)
    nx, ny, dx, dy = 97, 41, 0.0605, 0.0579
    r = np.random.default_rng(10 + n_rep)
    Ex = r.normal(0, 2e5, (ny, nx)).astype(np.float32)
    Ey = r.normal(0, 2e5, (ny, nx)).astype(np.float32)
    if with_b:
        Bx, By, Bz = [r.normal(0, 0.3, (ny, nx)).astype(np.float32) for _ in range(3)]
    else:
        Bx = By = Bz = np.zeros((ny, nx), np.float32)
    n = 50001
    st = _random_state(3, n, nx, ny, dx, dy)
    dt, qm = 5e-10 / 3, -1.602e-19 / (131.293 * 1.6605e-27 / 200)
    ref = [a.copy() for a in st]
    for _ in range(n_rep):
        K.push_boris(*ref, Ex, Ey, Bx, By, Bz, n, dx, dy, nx, ny, dt, qm)
    d = [_dev(a) for a in st]
    bt = [_dev(b) for b in (Bx, By, Bz)] if with_b else [None, None, None]
    ex_d, ey_d = _dev(Ex), _dev(Ey)  # keep every device tensor alive across the call
    rc = lib.bemcs_push_boris(*[_p(t) for t in d], n, _p(ex_d), _p(ey_d), *[_p(t) for t in bt],
                              nx, ny, dx, dy, dt, qm, n_rep, _stream())
    assert rc == 0
    for a, t in zip(ref, d):
        g = t.cpu().numpy()
        assert np.array_equal(g.view(np.uint32), a.view(np.uint32)), np.max(np.abs(g - a))


@pytest.mark.parametrize("b", [0.0605, 0.0579, 6.1925 / 1023, 2.3184 / 383, 0.05, 1.0, 3.0])
def test_fastdiv_certified_divisors(lib, b):
    """The reciprocal division equals IEEE division for all 2^32 dividends."""
    print(f"[gpu] fastdiv b={b}", flush=True)
    b32 = np.float32(b)
    r32 = np.float32(1.0) / b32
    bad = C.c_uint64(99)
    assert lib.bemcs_fastdiv_check(float(b32), float(r32), C.byref(bad), _stream()) == 0
    assert bad.value == 0
    assert lib.bemcs_fastdiv_enabled(float(b32), _stream()) == 1


def test_fastdiv_check_detects_wrong_reciprocal(lib):
    """The exhaustive check is live: a reciprocal 2^-10 off fails it."""
    print("[gpu] fastdiv planted error", flush=True)
    b32 = np.float32(0.0605)
    r32 = np.float32(1.0) / b32 * np.float32(1.0 + 2.0 ** -10)
    bad = C.c_uint64(0)
    assert lib.bemcs_fastdiv_check(float(b32), float(r32), C.byref(bad), _stream()) == 0
    assert bad.value > 1_000_000
    assert lib.bemcs_fastdiv_enabled(0.0, _stream()) == 0


def test_push_ieee_division_path_equal(lib, monkeypatch):
    """BEMCS_FASTDIV=0 (IEEE division) and the default give the same bits."""
    print("[gpu] push fastdiv vs ieee", flush=True)
    nx, ny, dx, dy = 97, 41, 0.0605, 0.0579
    r = np.random.default_rng(5)
    Ex = _dev(r.normal(0, 2e5, (ny, nx)).astype(np.float32))
    Ey = _dev(r.normal(0, 2e5, (ny, nx)).astype(np.float32))
    st = _random_state(4, 30001, nx, ny, dx, dy, pad=2.0)
    out = []
    for env in (None, "0"):
        if env is None:
            monkeypatch.delenv("BEMCS_FASTDIV", raising=False)
        else:
            monkeypatch.setenv("BEMCS_FASTDIV", env)
        d = [_dev(a) for a in st]
        rc = lib.bemcs_push_boris(*[_p(t) for t in d], len(st[0]), _p(Ex), _p(Ey), None, None, None,
                                  nx, ny, dx, dy, 1e-10, -9.6e9, 5, _stream())
        assert rc == 0
        out.append([t.cpu().numpy().view(np.uint32) for t in d])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_drop_in_push_numpy_arrays(lib):
    from thruster_cx_code_amd.kernels import push_particles_boris_hip
    nx, ny, dx, dy = 33, 17, 0.05, 0.05
    r = np.random.default_rng(1)
    Ex = r.normal(0, 1e5, (ny, nx)).astype(np.float32); Ey = r.normal(0, 1e5, (ny, nx)).astype(np.float32)
    Z = np.zeros((ny, nx), np.float32)
    st = _random_state(2, 1000, nx, ny, dx, dy)
    cap = [np.concatenate([a, np.full(10, 7.0, np.float32)]) for a in st]  # buffer longer than num_p
    ref = [a.copy() for a in cap]
    K.push_boris(*ref, Ex, Ey, Z, Z, Z, 1000, dx, dy, nx, ny, 5e-10, 7.3e5)
    push_particles_boris_hip(*cap, Ex, Ey, Z, Z, Z, 1000, dx, dy, nx, ny, 5e-10, 7.3e5)
    for a, b in zip(cap, ref):
        assert np.array_equal(a, b)


def test_deposit_and_rho_bit_exact(lib):
    from thruster_cx_code_amd.kernels import accumulate_rho_hip
    print("[gpu] deposit", flush=True)
    nx, ny, dx, dy = 128, 48, 6.1925 / 127, 2.3184 / 47
    r = np.random.default_rng(5)
    xi = r.uniform(-0.2, 6.4, 300000).astype(np.float32); yi = r.uniform(-0.1, 2.4, 300000).astype(np.float32)
    xe = r.uniform(0, 6.2, 120000).astype(np.float32); ye = r.uniform(0, 2.3, 120000).astype(np.float32)
    xi[:3] = np.float32([np.nan, dx * 2.5, -1e20])
    cd = 1.602e-19 * 1000.0 / ((dx * 1e-3) * (dy * 1e-3) * 1e-3)
    ref = np.zeros((ny, nx), np.float32)
    K.accumulate_rho(xi, yi, ref, len(xi), dx, dy, nx, ny, cd)
    K.accumulate_rho(xe, ye, ref, len(xe), dx, dy, nx, ny, -cd)
    got = np.zeros((ny, nx), np.float32)
    accumulate_rho_hip(xi, yi, got, len(xi), dx, dy, nx, ny, cd)
    accumulate_rho_hip(xe, ye, got, len(xe), dx, dy, nx, ny, -cd)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    cnt = torch.zeros(nx * ny, dtype=torch.int32, device="cuda")
    xd, yd = _dev(xi), _dev(yi)
    assert lib.bemcs_deposit_ngp_counts(_p(xd), _p(yd), len(xi), dx, dy, nx, ny, _p(cnt), _stream()) == 0
    assert np.array_equal(cnt.cpu().numpy().astype(np.uint32), K.deposit_counts(xi, yi, dx, dy, nx, ny))


def test_substeps_bit_exact(lib):
    print("[gpu] substeps", flush=True)
    nx, ny, dx, dy = 128, 48, 6.1925 / 127, 2.3184 / 47
    r = np.random.default_rng(6)
    Ex = r.normal(0, 5e5, (ny, nx)).astype(np.float32); Ey = r.normal(0, 5e5, (ny, nx)).astype(np.float32)
    Exy = np.stack([Ex, Ey], -1)
    for qm, n in [(7.3e5, 100000), (-1.47e8, 77777), (-1.47e8, 0), (-1.47e8, 1500000)]:
        st = _random_state(int(n) + 1, n, nx, ny, dx, dy, vscale=4e4) if n else None
        dt = 5e-10
        if n:
            want_n, want_ds, _ = K.substeps(*st, Ex, Ey, dx, dy, nx, ny, qm, dt)
        else:
            want_n, want_ds = 1, 0.0
        d = [_dev(a) for a in st] if n else [torch.zeros(1, device="cuda") for _ in range(5)]
        cnt = torch.tensor([n], dtype=torch.int32, device="cuda")
        bits = torch.zeros(1, dtype=torch.int64, device="cuda")
        nsub = torch.zeros(1, dtype=torch.int32, device="cuda")
        exy_d = _dev(Exy)
        assert lib.bemcs_substeps(*[_p(t) for t in d], _p(cnt), n, _p(exy_d), nx, ny, dx, dy, qm, dt,
                                  dt ** 2, 0.25, _p(bits), _p(nsub), _stream()) == 0
        got_ds = bits.cpu().numpy().view(np.float64)[0]
        assert got_ds == want_ds
        assert int(nsub.item()) == want_n


def test_percentile_bit_exact(lib):
    for n in [6, 7, 50, 1001, 65537, 300000]:
        r = np.random.default_rng(n)
        a = (np.abs(r.normal(0, 5, n)) ).astype(np.float32)
        a[: min(n, 10)] = a[0]  # ties
        nd = torch.tensor([n], dtype=torch.int32, device="cuda")
        out = torch.zeros(1, dtype=torch.float64, device="cuda")
        ad = _dev(a)
        assert lib.bemcs_percentile(_p(ad), _p(nd), n, 95.0, 5, None, _p(out), _stream()) == 0
        assert np.float32(out.item()) == np.percentile(a, 95), n
    # heavy ties: the ranks either side of the 95th percentile fall on the
    # same value or on neighbouring values
    for n, levels in [(100, 3), (1000, 4), (4097, 2), (65536, 7), (20, 20)]:
        r = np.random.default_rng(n + levels)
        a = r.integers(0, levels, n).astype(np.float32) * np.float32(0.37)
        nd = torch.tensor([n], dtype=torch.int32, device="cuda")
        out = torch.zeros(1, dtype=torch.float64, device="cuda")
        ad = _dev(a)
        assert lib.bemcs_percentile(_p(ad), _p(nd), n, 95.0, 5, None, _p(out), _stream()) == 0
        assert np.float32(out.item()) == np.percentile(a, 95), (n, levels)
    nd = torch.tensor([5], dtype=torch.int32, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    ad = _dev(a)
    lib.bemcs_percentile(_p(ad), _p(nd), 5, 95.0, 5, None, _p(out), _stream())
    assert np.isnan(out.item())


def _geometry(nx, ny, periodic):
    """Electrode masks of the two-grid test geometry on an nx x ny mesh."""
    from oracle.sim import OracleSimulator
    from conftest import c1_params
    o = OracleSimulator()
    p = c1_params(nx=nx, ny=ny, ions=0, electrons=0, geometry="one_hole" if periodic else "half_hole")
    o.build_domain(p)
    return o


@pytest.mark.parametrize("nx,ny,periodic", [(128, 48, False), (256, 96, False), (130, 60, True),
                                            (257, 97, False), (1024, 384, False)])
def test_poisson_solve_vs_superlu(lib, nx, ny, periodic, monkeypatch):
    """Repeated solves on one handle (with eager work and host synchronisation
    in between) against SuperLU.  The first handle replays captured HIP graphs,
    the second launches every kernel directly (BEMCS_POISSON_GRAPHS=0): both
    must give identical bits on every replay."""
    print(f"[gpu] poisson {nx}x{ny} periodic={periodic}", flush=True)
    o = _geometry(nx, ny, periodic)
    ib = np.ascontiguousarray(o.isBound.astype(np.uint8))
    handles = []
    for graphs in ("1", "0"):
        monkeypatch.setenv("BEMCS_POISSON_GRAPHS", graphs)
        h = C.c_void_p()
        assert lib.bemcs_poisson_create(ib.ctypes.data_as(C.c_void_p), nx, ny, int(periodic), C.byref(h)) == 0
        handles.append(h)
    monkeypatch.delenv("BEMCS_POISSON_GRAPHS")
    try:
        r = np.random.default_rng(nx)
        for rep in range(3):
            b = o.poisson.rhs_vector(r.normal(0, 1.0, nx * ny) * 1e-3, o.V_fixed.ravel() * (1 + rep), 1.0)
            b[o.poisson.kind.ravel() == 2] = r.normal(0, 0.1, np.count_nonzero(o.poisson.kind == 2))
            want = o.poisson.solve(b)
            bd = _dev(b)
            outs = []
            for h in handles:
                V = torch.zeros(nx * ny, dtype=torch.float64, device="cuda")
                its, rel = C.c_int32(0), C.c_double(0)
                rc = lib.bemcs_poisson_solve(h, _p(bd), _p(V), 1e-13, 300, C.byref(its), C.byref(rel), _stream())
                assert rc == 0, lib.bemcs_last_error()
                got = V.cpu().numpy()
                err = np.max(np.abs(got - want)) / np.max(np.abs(want))
                print(f"  rep {rep} iters={its.value} relres={rel.value:.2e} err_inf={err:.2e}")
                assert err < 1e-9
                assert its.value < 60
                outs.append(got)
                _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
            assert np.array_equal(outs[0], outs[1])
    finally:
        for h in handles:
            lib.bemcs_poisson_free(h)


def test_efield_bit_exact(lib):
    for periodic in (False, True):
        r = np.random.default_rng(3)
        V = r.normal(0, 300, (40, 70))
        Ex, Ey = efield(V, 0.05, 0.04, periodic)
        E = torch.zeros(2 * V.size, dtype=torch.float32, device="cuda")
        vd = _dev(V)
        assert lib.bemcs_efield(_p(vd), 70, 40, 0.05 * 1e-3, 0.04 * 1e-3, int(periodic), _p(E), _stream()) == 0
        g = E.cpu().numpy().reshape(40, 70, 2)
        assert np.array_equal(g[..., 0], Ex) and np.array_equal(g[..., 1], Ey)


def test_poisson_error_vs_tolerance(lib):
    """Error against SuperLU as a function of the PCG tolerance on the
    1024 x 384 two-grid mesh (documents the margin of the simulator default
    poisson_rtol = 5e-10)."""
    nx, ny = 1024, 384
    o = _geometry(nx, ny, False)
    r = np.random.default_rng(7)
    rho = np.abs(r.normal(0, 1.0, nx * ny)) * 0.05
    coeff = (o.dx * 1e-3) ** 2 / 8.854e-12
    b = o.poisson.rhs_vector(rho, o.V_fixed.ravel(), coeff)
    want = o.poisson.solve(b)
    ib = np.ascontiguousarray(o.isBound.astype(np.uint8))
    h = C.c_void_p()
    assert lib.bemcs_poisson_create(ib.ctypes.data_as(C.c_void_p), nx, ny, 0, C.byref(h)) == 0
    try:
        bd = _dev(b)
        errs = {}
        for rtol in (1e-9, 1e-10, 1e-11, 1e-12, 1e-13):
            V = torch.zeros(nx * ny, dtype=torch.float64, device="cuda")
            its, rel = C.c_int32(0), C.c_double(0)
            assert lib.bemcs_poisson_solve(h, _p(bd), _p(V), rtol, 400, C.byref(its), C.byref(rel), _stream()) == 0
            err = np.max(np.abs(V.cpu().numpy() - want)) / np.max(np.abs(want))
            errs[rtol] = err
            print(f"  rtol {rtol:.0e}: iters {its.value} relres {rel.value:.2e} err_inf {err:.2e}")
        # the difference levels off near 1e-11: SuperLU's own rounding error at
        # this condition number, not the tolerance
        assert errs[1e-9] < 2e-10
        for rt in (1e-10, 1e-11, 1e-12, 1e-13):
            assert errs[rt] < 5e-11
    finally:
        lib.bemcs_poisson_free(h)


def _picard_case(lib, nx, ny, periodic, seed):
    o = _geometry(nx, ny, periodic)
    r = np.random.default_rng(seed)
    n = nx * ny
    bufs = {
        "ci": torch.from_numpy(r.poisson(3.0, n).astype(np.int32)).cuda(),
        "ce": torch.from_numpy(r.poisson(2.0, n).astype(np.int32)).cuda(),
        "Vf": _dev(o.V_fixed.ravel().astype(np.float64)),
        "V": _dev(o.V_fixed.ravel().astype(np.float64)),
        "rho": torch.zeros(n, dtype=torch.float32, device="cuda"),
        "E": torch.zeros(2 * n, dtype=torch.float32, device="cuda"),
    }
    cell_vol = (o.dx * 1e-3) * (o.dy * 1e-3) * 1e-3
    cd = 1.602e-19 * 1000.0 / cell_vol
    return o, bufs, cd


def _picard_params(o, bufs, cd):
    from thruster_cx_code_amd import _lib as L
    p = L.Picard()
    p.cnt_ion, p.cnt_e = _p(bufs["ci"]), _p(bufs["ce"])
    p.V_fixed, p.V, p.rho_out, p.Exy = _p(bufs["Vf"]), _p(bufs["V"]), _p(bufs["rho"]), _p(bufs["E"])
    p.cd_ion, p.cd_e = cd, -cd
    p.coeff = (o.dx * 1e-3) ** 2 / 8.854187817e-12
    p.V_plasma, p.Te_up, p.n0, p.q, p.omega = 1020.0, 3.0, 1e17, 1.602e-19, 0.2
    p.dx_m, p.dy_m = o.dx * 1e-3, o.dy * 1e-3
    p.rtol, p.iterations, p.max_iter = 1e-9, 5, 200
    return p


def test_poisson_batch_picard_matches_single(lib):
    """Three uneven meshes (one periodic) solved as one batch give the same
    bits as each handle's own Picard call, over two calls (warm start)."""
    from thruster_cx_code_amd import _lib as L
    print("[gpu] poisson batch picard", flush=True)
    specs = [(128, 48, False, 1), (257, 97, False, 2), (130, 60, True, 3)]
    singles, batched, hs_single, hs_batch = [], [], [], []
    try:
        for nx, ny, per, seed in specs:
            o, b1, cd = _picard_case(lib, nx, ny, per, seed)
            _, b2, _ = _picard_case(lib, nx, ny, per, seed)
            ib = np.ascontiguousarray(o.isBound.astype(np.uint8))
            for lst in (hs_single, hs_batch):
                h = C.c_void_p()
                assert lib.bemcs_poisson_create(ib.ctypes.data_as(C.c_void_p), nx, ny, int(per), C.byref(h)) == 0
                lst.append(h)
            singles.append((o, b1, cd))
            batched.append((o, b2, cd))
        arr = (C.c_void_p * len(specs))(*[h.value for h in hs_batch])
        batch = C.c_void_p()
        assert lib.bemcs_poisson_batch_create(arr, len(specs), C.byref(batch)) == 0
        assert lib.bemcs_poisson_batch_size(batch) == len(specs)
        try:
            for call in range(2):
                its_single = []
                for h, (o, b, cd) in zip(hs_single, singles):
                    it, rel = C.c_int32(0), C.c_double(0)
                    p = _picard_params(o, b, cd)
                    assert lib.bemcs_poisson_picard(h, C.byref(p), C.byref(it), C.byref(rel), _stream()) == 0, \
                        lib.bemcs_last_error()
                    its_single.append(it.value)
                ps = (L.Picard * len(specs))(*[_picard_params(o, b, cd) for o, b, cd in batched])
                its = (C.c_int32 * len(specs))()
                rels = (C.c_double * len(specs))()
                assert lib.bemcs_poisson_batch_picard(batch, ps, its, rels, _stream()) == 0, lib.bemcs_last_error()
                print(f"  call {call}: single iters {its_single} batch iters {list(its)}")
                assert list(its) == its_single
                for (_, b1, _), (_, b2, _) in zip(singles, batched):
                    for key in ("V", "E", "rho"):
                        assert torch.equal(b1[key].view(torch.int32) if key != "V" else b1[key].view(torch.int64),
                                           b2[key].view(torch.int32) if key != "V" else b2[key].view(torch.int64)), key
        finally:
            lib.bemcs_poisson_batch_free(batch)
    finally:
        for h in hs_single + hs_batch:
            lib.bemcs_poisson_free(h)


def test_poisson_update_equals_create(lib):
    """An operator updated in place to a new electrode map (eroded cells,
    moved grid faces) solves bit-identically to a handle created from that
    map, and both agree with SuperLU."""
    from oracle.poisson import SuperLUPoisson
    nx, ny = 257, 97
    o = _geometry(nx, ny, False)
    A = o.isBound.copy()
    B = A.copy()
    r = np.random.default_rng(5)
    cand = np.argwhere(A[1:-1, 1:-1]) + 1
    for iy, ix in cand[r.choice(len(cand), 40, replace=False)]:
        B[iy, ix] = False                          # eroded cells
    cols = np.where(A.any(axis=0))[0]
    B[: ny // 2, cols[cols > 2].min() - 3] = True  # a bowed face
    ha, hb = C.c_void_p(), C.c_void_p()
    ia = np.ascontiguousarray(A.astype(np.uint8))
    ibm = np.ascontiguousarray(B.astype(np.uint8))
    assert lib.bemcs_poisson_create(ia.ctypes.data_as(C.c_void_p), nx, ny, 0, C.byref(ha)) == 0
    assert lib.bemcs_poisson_create(ibm.ctypes.data_as(C.c_void_p), nx, ny, 0, C.byref(hb)) == 0
    try:
        bd_map = _dev(ibm.ravel())
        assert lib.bemcs_poisson_update(ha, _p(bd_map), _stream()) == 0, lib.bemcs_last_error()
        sp = SuperLUPoisson(B, False)
        b = sp.rhs_vector(r.normal(0, 1.0, nx * ny) * 1e-3, o.V_fixed.ravel(), 1.0)
        want = sp.solve(b)
        bd = _dev(b)
        outs = []
        for h in (ha, hb):
            V = torch.zeros(nx * ny, dtype=torch.float64, device="cuda")
            its, rel = C.c_int32(0), C.c_double(0)
            assert lib.bemcs_poisson_solve(h, _p(bd), _p(V), 1e-12, 300, C.byref(its), C.byref(rel), _stream()) == 0
            outs.append(V.cpu().numpy())
            err = np.max(np.abs(outs[-1] - want)) / np.max(np.abs(want))
            print(f"[gpu] poisson update: iters {its.value} err {err:.2e}", flush=True)
            assert err < 1e-10
        assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))
    finally:
        lib.bemcs_poisson_free(ha)
        lib.bemcs_poisson_free(hb)


def test_poisson_shared_operator_and_update(lib):
    """Handles created from one electrode map share their operator arrays;
    updating one of them to a new map gives it arrays of its own: it then
    solves like a handle created from the new map, the other keeps solving
    the old map's system bit for bit, and a solve set made before the update
    refuses to run."""
    from oracle.poisson import SuperLUPoisson
    from thruster_cx_code_amd import _lib as L
    nx, ny = 257, 97
    o = _geometry(nx, ny, False)
    A = o.isBound.copy()
    B = A.copy()
    r = np.random.default_rng(7)
    cand = np.argwhere(A[1:-1, 1:-1]) + 1
    for iy, ix in cand[r.choice(len(cand), 30, replace=False)]:
        B[iy, ix] = False
    ia = np.ascontiguousarray(A.astype(np.uint8))
    ibm = np.ascontiguousarray(B.astype(np.uint8))
    h1, h2, h3 = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for h, m in ((h1, ia), (h2, ia), (h3, ibm)):
        assert lib.bemcs_poisson_create(m.ctypes.data_as(C.c_void_p), nx, ny, 0, C.byref(h)) == 0

    def solve(h, b):
        V = torch.zeros(nx * ny, dtype=torch.float64, device="cuda")
        bd = _dev(b)
        its, rel = C.c_int32(0), C.c_double(0)
        assert lib.bemcs_poisson_solve(h, _p(bd), _p(V), 1e-12, 300, C.byref(its), C.byref(rel),
                                       _stream()) == 0, lib.bemcs_last_error()
        return V.cpu().numpy()

    batch = C.c_void_p()
    try:
        spA, spB = SuperLUPoisson(A, False), SuperLUPoisson(B, False)
        src = r.normal(0, 1.0, nx * ny) * 1e-3
        bA = spA.rhs_vector(src, o.V_fixed.ravel(), 1.0)
        bB = spB.rhs_vector(src, o.V_fixed.ravel(), 1.0)
        v1, v2 = solve(h1, bA), solve(h2, bA)
        assert np.array_equal(v1.view(np.uint64), v2.view(np.uint64))
        assert np.max(np.abs(v1 - spA.solve(bA))) / np.max(np.abs(v1)) < 1e-10
        arr = (C.c_void_p * 2)(h1.value, h2.value)
        assert lib.bemcs_poisson_batch_create(arr, 2, C.byref(batch)) == 0
        print("[gpu] shared operator: update one handle", flush=True)
        bd_map = _dev(ibm.ravel())
        assert lib.bemcs_poisson_update(h1, _p(bd_map), _stream()) == 0, lib.bemcs_last_error()
        w1, w3 = solve(h1, bB), solve(h3, bB)
        assert np.array_equal(w1.view(np.uint64), w3.view(np.uint64))
        assert np.max(np.abs(w1 - spB.solve(bB))) / np.max(np.abs(w1)) < 1e-10
        assert np.array_equal(solve(h2, bA).view(np.uint64), v2.view(np.uint64))
        ps = (L.Picard * 2)()
        its, rels = (C.c_int32 * 2)(), (C.c_double * 2)()
        assert lib.bemcs_poisson_batch_picard(batch, ps, its, rels, _stream()) == L.EINVAL
        assert b"create the set again" in lib.bemcs_last_error()
    finally:
        if batch:
            lib.bemcs_poisson_batch_free(batch)
        for h in (h1, h2, h3):
            lib.bemcs_poisson_free(h)


def test_poisson_update_rejects_invalid_map(lib):
    nx, ny = 128, 48
    o = _geometry(nx, ny, False)
    ia = np.ascontiguousarray(o.isBound.astype(np.uint8))
    h = C.c_void_p()
    assert lib.bemcs_poisson_create(ia.ctypes.data_as(C.c_void_p), nx, ny, 0, C.byref(h)) == 0
    try:
        bad = ia.copy()
        bad[:, 0] = 0                               # no plasma boundary: unknowns on the x edge
        assert lib.bemcs_poisson_update(h, _p(_dev(bad.ravel())), _stream()) != 0
        assert b"edge" in lib.bemcs_last_error()
    finally:
        lib.bemcs_poisson_free(h)


@pytest.mark.parametrize("nx,ny", [(1024, 384), (2048, 768)])
def test_two_level_down_launch_same_bits(lib, nx, ny, monkeypatch):
    """The fused down launch of the last two coarse levels computes what two
    single-level launches compute, in the same order: solves give identical
    bits and iteration counts with it (BEMCS_MG_FUSE=1) and without it (the
    default)."""
    from conftest import c1_params
    from thruster_cx_code_amd import DigitalTwinSimulator
    print(f"[gpu] two-level down {nx}x{ny}", flush=True)
    g = DigitalTwinSimulator()
    g.build_domain(c1_params(nx=nx, ny=ny, ions=0, electrons=0))
    ib = np.ascontiguousarray(g.isBound.astype(np.uint8))
    handles = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("BEMCS_MG_FUSE", fuse)
        h = C.c_void_p()
        assert lib.bemcs_poisson_create(ib.ctypes.data_as(C.c_void_p), nx, ny, 0, C.byref(h)) == 0
        handles.append(h)
    monkeypatch.delenv("BEMCS_MG_FUSE")
    try:
        assert lib.bemcs_poisson_levels(handles[0]) >= 5
        r = np.random.default_rng(nx)
        for rep in range(2):
            bd = _dev(r.normal(0.0, 1.0, nx * ny) * np.where(ib.ravel() != 0, 100.0, 1e-3))
            outs = []
            for h in handles:
                V = torch.zeros(nx * ny, dtype=torch.float64, device="cuda")
                its, rel = C.c_int32(0), C.c_double(0)
                rc = lib.bemcs_poisson_solve(h, _p(bd), _p(V), 1e-10, 300, C.byref(its), C.byref(rel), _stream())
                assert rc == 0, lib.bemcs_last_error()
                outs.append((V.cpu().numpy(), its.value))
            print(f"  rep {rep}: iterations {outs[0][1]} / {outs[1][1]}")
            assert outs[0][1] == outs[1][1]
            assert np.array_equal(outs[0][0], outs[1][0])
    finally:
        for h in handles:
            lib.bemcs_poisson_free(h)


def test_picard_stage_ahead_same_bits(lib, monkeypatch):
    """The launches that follow a Picard solve go out before the host reads
    whether the solve converged, guarded by the stop flags.  A first-graph
    prediction of 1 iteration (every solve of the first call needs further
    rounds, each enqueuing the guarded launches again; in the batch the cases
    converge in different rounds) and one of 40 (launches past convergence)
    give the same bits and iteration counts, singly and batched."""
    from thruster_cx_code_amd import _lib as L
    print("[gpu] picard stage ahead", flush=True)
    specs = [(128, 48, False, 1), (257, 97, False, 2), (130, 60, True, 3)]
    runs = {}
    for pred in ("40", "1"):
        monkeypatch.setenv("BEMCS_PCG_PREDICT", pred)
        for mode in ("single", "batch"):
            cases, hs = [], []
            batch = C.c_void_p()
            try:
                for nx, ny, per, seed in specs:
                    o, b, cd = _picard_case(lib, nx, ny, per, seed)
                    ib = np.ascontiguousarray(o.isBound.astype(np.uint8))
                    h = C.c_void_p()
                    assert lib.bemcs_poisson_create(ib.ctypes.data_as(C.c_void_p), nx, ny, int(per),
                                                    C.byref(h)) == 0
                    hs.append(h)
                    cases.append((o, b, cd))
                if mode == "batch":
                    arr = (C.c_void_p * len(specs))(*[h.value for h in hs])
                    assert lib.bemcs_poisson_batch_create(arr, len(specs), C.byref(batch)) == 0
                out = []
                for call in range(2):
                    if mode == "single":
                        its = []
                        for h, (o, b, cd) in zip(hs, cases):
                            it, rel = C.c_int32(0), C.c_double(0)
                            p = _picard_params(o, b, cd)
                            assert lib.bemcs_poisson_picard(h, C.byref(p), C.byref(it), C.byref(rel),
                                                            _stream()) == 0, lib.bemcs_last_error()
                            its.append(it.value)
                    else:
                        ps = (L.Picard * len(specs))(*[_picard_params(o, b, cd) for o, b, cd in cases])
                        itc = (C.c_int32 * len(specs))()
                        rels = (C.c_double * len(specs))()
                        assert lib.bemcs_poisson_batch_picard(batch, ps, itc, rels, _stream()) == 0, \
                            lib.bemcs_last_error()
                        its = list(itc)
                    out.append((its, [(b["V"].cpu().numpy().copy(), b["E"].cpu().numpy().copy())
                                      for _, b, _ in cases]))
                runs[(pred, mode)] = out
                print(f"  predict {pred} {mode}: iterations {[o[0] for o in out]}")
            finally:
                if batch.value:
                    lib.bemcs_poisson_batch_free(batch)
                for h in hs:
                    lib.bemcs_poisson_free(h)
    monkeypatch.delenv("BEMCS_PCG_PREDICT")
    ref = runs[("40", "single")]
    for key, out in runs.items():
        for (its_r, fields_r), (its, fields) in zip(ref, out):
            assert its == its_r, key
            for (Vr, Er), (V, E) in zip(fields_r, fields):
                assert np.array_equal(V.view(np.uint64), Vr.view(np.uint64)), key
                assert np.array_equal(E.view(np.uint32), Er.view(np.uint32)), key


def test_stream_pointer_follows_current_stream(lib):
    """stream_ptr() returns the raw handle of torch's current stream, also
    inside a `torch.cuda.stream` context and for an explicit stream."""
    import torch
    from thruster_cx_code_amd.device import stream_ptr

    def raw(p):
        return p.value or 0          # the default stream's handle is 0 (NULL)
    assert raw(stream_ptr()) == torch.cuda.current_stream().cuda_stream
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        assert raw(stream_ptr()) == s.cuda_stream != 0
    assert raw(stream_ptr(s)) == s.cuda_stream
    assert raw(stream_ptr()) == torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("sx,sy", [(0, 0), (2, 2), (3, 1)])
def test_sort_by_tile_is_a_keyed_permutation(lib, sx, sy):
    """Counting sort by gather tile: the output holds exactly the input
    particles (ids travel with them) in non-decreasing tile order, and the
    output count is set."""
    from thruster_cx_code_amd import _lib
    from thruster_cx_code_amd.device import SpeciesBuffers
    print(f"[gpu] sort_by_tile {sx},{sy}", flush=True)
    nx, ny, dx, dy = 130, 61, 0.05, 0.04
    n = 300000
    st = _random_state(11, n, nx, ny, dx, dy)
    st[0][10:20] = np.float32([np.nan, np.inf, -np.inf, 1e30, -1e30, 0, 0, 0, 0, 0])
    a, b = SpeciesBuffers(n + 100, torch.device("cuda"), True), SpeciesBuffers(n + 100, torch.device("cuda"), True)
    for name, arr in zip(("x", "y", "vx", "vy", "vz"), st):
        getattr(a, name)[:n].copy_(torch.from_numpy(arr))
    tag = np.arange(n, dtype=np.uint32)
    a.tag[:n].copy_(torch.from_numpy(tag.view(np.int32)))
    a.count.fill_(n)
    b.count.fill_(-5)
    nbins = (((nx - 2) >> sx) + 1) * (((ny - 2) >> sy) + 1)
    scratch = torch.zeros(nx * ny + (nx * ny + 1023) // 1024 + 1024 + n, dtype=torch.int32, device="cuda")
    assert lib.bemcs_sort_by_tile(C.byref(a.struct()), C.byref(b.struct()), n, nx, ny, dx, dy, sx, sy,
                                  _p(scratch), _stream()) == 0
    torch.cuda.synchronize()
    assert int(b.count.item()) == n
    out = [getattr(b, k)[:n].cpu().numpy() for k in ("x", "y", "vx", "vy", "vz")]
    otag = b.tag[:n].cpu().numpy().view(np.uint32)
    assert np.array_equal(np.sort(otag), tag)
    for arr_in, arr_out in zip(st, out):
        assert np.array_equal(arr_in[otag].view(np.uint32), arr_out.view(np.uint32))
    gx = out[0] / np.float32(dx)
    gy = out[1] / np.float32(dy)
    with np.errstate(invalid="ignore"):
        i0 = np.clip(K._floor_i32(gx), 0, nx - 2)
        j0 = np.clip(K._floor_i32(gy), 0, ny - 2)
    key = (j0.astype(np.int64) >> sy) * (((nx - 2) >> sx) + 1) + (i0.astype(np.int64) >> sx)
    assert np.all(np.diff(key) >= 0) and key.max() < nbins
