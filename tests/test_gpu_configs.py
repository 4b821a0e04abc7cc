"""GPU tests at the benchmark configurations' own sizes.

C1 (256 x 96, 1e5 + 1e5): full step parity against the oracle over four
steps, odd and even.

C2 (1024 x 384, 8M + 8M held): the oracle cannot push 16M particles through
~1500 substeps in a test, so the step is checked by
* the substep counts of both species, recomputed by the oracle over all 16M
  particles after the oracle has applied the same injection and top-up;
* bit-exact parity of a random subset of ions (by id) and of electrons (by
  state) through push, hits, removal and CEX on an odd step: each particle's
  step depends only on its own state and the field;
* the Picard call of the even step against SciPy SuperLU from the same V and
  rho, within 1e-9 relative;
* bookkeeping: the NGP counts and the PPC map sum to the surviving
  populations;
* two runs giving identical bits.

C5 (2048 x 768, 32M + 32M, sim_mode Both with RF): two steps, checked for
bookkeeping, the temperature floor, the damage sign and identical bits over
two runs.
"""
import numpy as np
import pytest

from conftest import c1_params
from literal_step import electron_purge, ion_section, mcc_primary
from oracle import kernels as K
from oracle import rng
from oracle.sim import OracleSimulator
from test_gpu_step import _ions_by_pid, _mk, _steps_vs_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

F32 = np.float32


def _sim(p, **kw):
    from thruster_cx_code_amd import DigitalTwinSimulator
    g = DigitalTwinSimulator(**kw)
    g.m_e = g.m_ion / 200.0
    g.build_domain(p)
    return g


def test_c1_size_steps_parity():
    """C1 at its own size: 256 x 96 nodes, 1e5 + 1e5 preloaded, four steps."""
    p = c1_params(nx=256, ny=96, ions=100000, electrons=100000)
    g, o = _mk(p)
    _steps_vs_oracle(p, g, o, 4)


def _state_keys(x, y, vx, vy, vz):
    """One 160-bit key per particle (the bits of its five float32 values)."""
    a = np.stack([np.asarray(v, F32).view(np.uint32) for v in (x, y, vx, vy, vz)], axis=1)
    return np.ascontiguousarray(a).view(np.dtype((np.void, 20))).ravel()


def test_c2_full_size_step_properties_and_sampled_parity():
    from bench import workload_params
    p = workload_params("c2")
    print("[gpu] c2 build 8M + 8M", flush=True)
    g = _sim(p, sort_interval=20, electron_sort_interval=20)
    assert (g.num_p, g.num_e) == (8_000_000, 8_000_000)
    pre_i = [g._ion_host(k) for k in ("x", "y", "vx", "vy", "vz")] + [g._ion_host("tag").view(np.uint32)]
    pre_e = [g._e_host(k) for k in ("x", "y", "vx", "vy", "vz")]
    V0, Ex, Ey = g.V.copy(), g.Ex.copy(), g.Ey.copy()
    hold_ids = [int(v) for v in g.hold_ids.cpu().numpy().view(np.uint32)]
    print("[gpu] c2 step 1 (odd)", flush=True)
    g.step(p)
    rec = g.record()
    # bookkeeping: every survivor adds one NGP count and one PPC count
    assert int(g.cnt_i.sum().item()) == g.num_p
    assert int(g.cnt_e.sum().item()) == g.num_e
    assert int(g.current_ppc_map.sum()) == g.num_p

    # the oracle applies the same injection and top-up to the same state
    q = dict(p)
    q["preload"] = {"ions": 0, "electrons": 0}
    o = OracleSimulator()
    o.m_e = o.m_ion / 200.0
    o.build_domain(q)
    o.p_x, o.p_y, o.p_vx, o.p_vy, o.p_vz = pre_i[:5]
    o.p_pid = pre_i[5] & np.uint32(0x7FFFFFFF)
    o.p_isCEX = (pre_i[5] >> 31).astype(bool)
    o.e_x, o.e_y, o.e_vx, o.e_vy, o.e_vz = pre_e
    o.V, o.Ex, o.Ey = V0, Ex, Ey
    o.next_ion_pid = 8_000_000
    o.hold_targets, o.hold_ids = (8_000_000, 8_000_000), hold_ids
    o.iteration = 1
    o._inject(q, o.iteration * o.dt)
    o._neutralizer(q)
    o._hold(q)
    assert (o.num_p, o.num_e) == tuple(int(v) for v in g.ring_nin.view(-1, 2)[1].cpu().numpy())
    qm_i, qm_e = o.q_ion / o.m_ion, -o.q / o.m_e
    ns_i = K.substeps(o.p_x, o.p_y, o.p_vx, o.p_vy, o.p_vz, Ex, Ey, o.dx, o.dy, o.nx, o.ny, qm_i, o.dt)[0]
    ns_e = K.substeps(o.e_x, o.e_y, o.e_vx, o.e_vy, o.e_vz, Ex, Ey, o.dx, o.dy, o.nx, o.ny, qm_e, o.dt)[0]
    print(f"  n_sub {ns_i} {ns_e} (device {rec['n_sub_ion']} {rec['n_sub_e']})", flush=True)
    assert (ns_i, ns_e) == (rec["n_sub_ion"], rec["n_sub_e"])

    r = np.random.default_rng(0)
    Z = np.zeros_like(Ex)
    # ions: push, hits, removal and CEX of a random subset, by id
    sel = np.sort(r.choice(o.num_p, 40000, replace=False))
    s_i = [a[sel].copy() for a in (o.p_x, o.p_y, o.p_vx, o.p_vy, o.p_vz)]
    pid, cex = o.p_pid[sel].copy(), o.p_isCEX[sel].copy()
    x_old, y_old = s_i[0].copy(), s_i[1].copy()
    for _ in range(ns_i):
        K.push_boris(*s_i, Ex, Ey, Z, Z, Z, len(sel), o.dx, o.dy, o.nx, o.ny, o.dt / ns_i, qm_i)
    cap = dict(x_old=x_old, y_old=y_old, x=s_i[0], y=s_i[1], vx=s_i[2], vy=s_i[3], vz=s_i[4], pid=pid, cex=cex)
    outs, _ = ion_section(o, cap, q["grids"])
    keep = np.array([d["keep"] for d in outs])
    prim = np.array(mcc_primary(o, s_i[0], cex))
    P = K.mcc_prob_det64(*s_i, o.Ly, q["n0"], o.dt, o.m_ion, o.q, q["neutral_model"], q["cx_model"])
    u, _ = rng.uniforms(o.seed, rng.P_MCC, pid.astype(np.uint64), 1, 0)
    coll = keep & prim & (u < P)
    gt = g._ion_host("tag").view(np.uint32)
    gpid = gt & np.uint32(0x7FFFFFFF)
    order = np.argsort(gpid)
    pos = np.searchsorted(gpid[order], pid)
    found = (pos < len(gpid)) & (gpid[order][np.minimum(pos, len(gpid) - 1)] == pid)
    assert np.array_equal(found, keep)
    idx = order[pos[keep]]
    nc = ~coll[keep]
    for a, b in zip((g.p_x, g.p_y), s_i[:2]):
        assert np.array_equal(a[idx].view(np.uint32), b[keep].view(np.uint32))
    for a, b in zip((g.p_vx, g.p_vy, g.p_vz), s_i[2:]):
        assert np.array_equal(a[idx][nc].view(np.uint32), b[keep][nc].view(np.uint32))
    assert np.array_equal((gt[idx] >> 31).astype(bool), cex[keep] | coll[keep])
    print(f"  ions: {keep.sum()} of {len(sel)} sampled survive, {coll.sum()} CEX events", flush=True)

    # electrons: push and purge of a random subset; each survivor's state is
    # in the device's electron set
    sel = np.sort(r.choice(o.num_e, 4000, replace=False))
    s_e = [a[sel].copy() for a in (o.e_x, o.e_y, o.e_vx, o.e_vy, o.e_vz)]
    for _ in range(ns_e):
        K.push_boris(*s_e, Ex, Ey, Z, Z, Z, len(sel), o.dx, o.dy, o.nx, o.ny, o.dt / ns_e, qm_e)
    keep_e = electron_purge(o, s_e[0], s_e[1])
    dev = _state_keys(g.e_x, g.e_y, g.e_vx, g.e_vy, g.e_vz)
    want = _state_keys(*[a[keep_e] for a in s_e])
    assert np.all(np.isin(want, dev))
    print(f"  electrons: {keep_e.sum()} of {len(sel)} sampled survive, all found", flush=True)

    # even step: the Picard call against SuperLU from the same V and rho
    V1 = g.V.copy()
    print("[gpu] c2 step 2 (even)", flush=True)
    g.step(p)
    rho = g.rho.copy()
    coeff = (o.dx * 1e-3) ** 2 / o.eps0
    Vp = q["grids"][0]["V"] + q.get("V_plasma_offset", 20.0)
    want_V = o.poisson.picard(V1, rho, o.V_fixed, coeff, Vp, q["Te_up"], q["n0_plasma"], o.q, 5)
    err = np.max(np.abs(g.V - want_V)) / np.max(np.abs(want_V))
    print(f"  Picard vs SuperLU rel err {err:.2e}", flush=True)
    assert err < 1e-9
    assert int(g.cnt_i.sum().item()) == g.num_p and int(g.cnt_e.sum().item()) == g.num_e

    # identical bits from a second run
    h = _sim(p, sort_interval=20, electron_sort_interval=20)
    h.advance(p, 2)
    assert np.array_equal(h.V, g.V)
    for a, b in zip(_ions_by_pid(g.p_x, g.p_y, g.p_vx, g.p_vy, g.p_vz, g.p_pid, g.p_isCEX),
                    _ions_by_pid(h.p_x, h.p_y, h.p_vx, h.p_vy, h.p_vz, h.p_pid, h.p_isCEX)):
        assert np.array_equal(a, b)
    assert np.array_equal(np.sort(_state_keys(g.e_x, g.e_y, g.e_vx, g.e_vy, g.e_vz)),
                          np.sort(_state_keys(h.e_x, h.e_y, h.e_vx, h.e_vy, h.e_vz)))


def test_c5_full_size_steps():
    """C5 at its own size: 2048 x 768, 32M + 32M, Both with RF, two steps."""
    from bench import workload_params
    p = workload_params("c5")
    runs = []
    for k in range(2):
        print(f"[gpu] c5 run {k}", flush=True)
        g = _sim(p, sort_interval=5, electron_sort_interval=1)
        g.advance(p, 2)
        rec = g.record()
        assert int(g.cnt_i.sum().item()) == g.num_p
        assert int(g.cnt_e.sum().item()) == g.num_e
        B = g.isBound
        T = g.Tmap
        assert np.all(T[B] >= 300.0)
        assert np.all(g.damage_map >= 0.0)
        assert np.all(np.isfinite(g.V))
        assert rec["n_sub_e"] >= 1
        runs.append((g.V.copy(), T.copy(), g.damage_map.copy(),
                     _ions_by_pid(g.p_x, g.p_y, g.p_vx, g.p_vy, g.p_vz, g.p_pid, g.p_isCEX), g.num_e))
        del g
        torch.cuda.empty_cache()
    a, b = runs
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    for u, v in zip(a[3], b[3]):
        assert np.array_equal(u, v)
    assert a[4] == b[4]
