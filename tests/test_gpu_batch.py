"""Batched sweep engine: cases advanced together (one launch per phase for all
cases) give the same results as each case stepped alone.

Ions are compared by persistent id and electrons as sorted multisets (the
compaction order of one launch depends on atomic order); potentials, fields,
deposit counts and per-step diagnostics records bit for bit.
"""
import numpy as np
import pytest

from conftest import c1_params

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _cases():
    # uneven meshes, voltages, apertures, populations, seeds; one periodic
    specs = [
        dict(nx=128, ny=48, ions=6000, electrons=4000, seed=1),
        dict(nx=144, ny=52, ions=2500, electrons=3000, seed=2),
        dict(nx=160, ny=64, ions=9000, electrons=2000, seed=3, geometry="one_hole"),
    ]
    out = []
    for k, sp in enumerate(specs):
        p = c1_params(**sp)
        p["grids"][1]["V"] = -150.0 - 50.0 * k
        p["grids"][1]["r"] = 0.5 + 0.05 * k
        out.append(p)
    return out


def _build(params):
    from thruster_cx_code_amd import DigitalTwinSimulator
    sims = []
    for p in params:
        s = DigitalTwinSimulator()
        s.m_e = s.m_ion / 200.0
        s.build_domain(p)
        sims.append(s)
    return sims


def _state(s):
    tag = s._ion_host("tag").view(np.uint32)
    order = np.argsort(tag, kind="stable")
    ions = np.stack([s.p_x, s.p_y, s.p_vx, s.p_vy, s.p_vz]).astype(np.float32)[:, order]
    e = np.stack([s.e_x, s.e_y, s.e_vx, s.e_vy, s.e_vz]).astype(np.float32)
    e = e[:, np.lexsort(e[::-1])]
    return tag[order], ions, e


def test_batched_steps_match_single_case_steps():
    from thruster_cx_code_amd.batch import BatchedSimulator
    params = _cases()
    alone = _build(params)
    together = _build(params)
    batch = BatchedSimulator(together, params)
    for step in range(1, 9):
        print(f"[gpu] batch step {step}", flush=True)
        for s, p in zip(alone, params):
            s.advance(p, 1)
        batch.advance(1)
        for c, (a, b) in enumerate(zip(alone, together)):
            ra, rb = a.record(), b.record()
            for key in ("num_p", "num_e", "n_sub_ion", "n_sub_e", "entered", "crossed", "lost_to_grid",
                        "see_spawned", "mcc_events", "transmitted"):
                assert ra[key] == rb[key], (step, c, key, ra[key], rb[key])
            for key in ("min_pot", "current_div"):
                assert (np.isnan(ra[key]) and np.isnan(rb[key])) or ra[key] == rb[key], (step, c, key)
            ta, ia, ea = _state(a)
            tb, ib, eb = _state(b)
            assert np.array_equal(ta, tb), (step, c)
            assert np.array_equal(ia.view(np.uint32), ib.view(np.uint32)), (step, c)
            assert np.array_equal(ea.view(np.uint32), eb.view(np.uint32)), (step, c)
            assert np.array_equal(a.V.view(np.uint64), b.V.view(np.uint64)), (step, c)
            assert np.array_equal(a.Ex.view(np.uint32), b.Ex.view(np.uint32)), (step, c)
            assert torch.equal(a.cnt_i, b.cnt_i) and torch.equal(a.cnt_e, b.cnt_e), (step, c)
    its = batch.poisson_iter_log[-1]
    assert its == [s.poisson_iter_log[-1] for s in alone]


def _same(a, b, where):
    ra, rb = a.record(), b.record()
    for key in ("num_p", "num_e", "n_sub_ion", "n_sub_e", "entered", "crossed", "mcc_events"):
        assert ra[key] == rb[key], (where, key)
    ta, ia, ea = _state(a)
    tb, ib, eb = _state(b)
    assert np.array_equal(ta, tb) and np.array_equal(ia.view(np.uint32), ib.view(np.uint32)), where
    assert np.array_equal(ea.view(np.uint32), eb.view(np.uint32)), where
    assert np.array_equal(a.V.view(np.uint64), b.V.view(np.uint64)), where


def test_autotuned_batch_matches_untuned_batch():
    """The autotuner changes sort intervals and the count window of every
    case between steps; the results equal those of the same batch run with
    the defaults (sorting and the window change no result)."""
    from thruster_cx_code_amd.batch import BatchedSimulator
    from thruster_cx_code_amd.sweep import Autotuner
    params = _cases()
    plain = _build(params)
    tuned = _build(params)
    b_plain = BatchedSimulator(plain, params)
    b_tuned = BatchedSimulator(tuned, params)
    tuner = Autotuner(b_tuned, ion_candidates=(1, 3), electron_candidates=(1, 2),
                      window_candidates=((64, 32), (512, 4), (1, 1), (32, 16)), trial=2, every=10)
    print("[gpu] autotuned batch", flush=True)
    tuner.advance(60)  # tunes of 16 steps with 10 steps between them: two tunes
    b_plain.advance(60)
    assert len(tuner.history) == 2
    assert {tuple(h["count_window"]) for h in tuner.history} <= {(64, 32), (512, 4), (1, 1), (32, 16)}
    for c, (a, b) in enumerate(zip(plain, tuned)):
        _same(a, b, c)
        assert torch.equal(a.cnt_i, b.cnt_i) and torch.equal(a.cnt_e, b.cnt_e), c


def test_graph_replayed_batch_matches_direct_launches():
    """The per-step kernels replayed from captured HIP graphs (launch sizes
    rounded up to powers of two) give the bits of direct launches, step after
    step, across a membership change that forces a new capture; a graph is
    captured once and replayed on the following steps."""
    from thruster_cx_code_amd.batch import BatchedSimulator
    params = _cases()
    direct = _build(params)
    graphed = _build(params)
    b_direct = BatchedSimulator(direct[:2], params[:2], graphs=False)
    b_graph = BatchedSimulator(graphed[:2], params[:2], graphs=True)
    for step in range(1, 11):
        print(f"[gpu] graph batch step {step}", flush=True)
        if step == 6:  # the third case joins: new argument layout, new graph
            b_direct.add(direct[2], params[2])
            b_graph.add(graphed[2], params[2])
        b_direct.advance(1)
        b_graph.advance(1)
        for c, (a, b) in enumerate(zip(b_direct.sims, b_graph.sims)):
            _same(a, b, (step, c))
            assert torch.equal(a.cnt_i, b.cnt_i) and torch.equal(a.cnt_e, b.cnt_e), (step, c)
            assert np.array_equal(a.Ex.view(np.uint32), b.Ex.view(np.uint32)), (step, c)
    assert b_direct.graph_captures == 0
    assert 2 <= b_graph.graph_captures < 10


@pytest.mark.parametrize("coarse_max", [None, 256])
def test_large_level0_tiles_alone_and_batched(coarse_max):
    """Cases whose Poisson handles ask for the large level-0 tiles (32 x 32
    fine nodes per work-group, as sweeps and the server build them), with the
    default coarsest level or a smaller one (one more level, as the sweep's
    batched cases take), give the same bits alone and batched; their first
    potentials agree with the default handles' to the solver tolerance."""
    from thruster_cx_code_amd import DigitalTwinSimulator
    from thruster_cx_code_amd.batch import BatchedSimulator
    from thruster_cx_code_amd.sweep import build_case
    params = _cases()
    alone = [build_case(p, coarse_max=coarse_max, level0_tile=16) for p in params]
    together = [build_case(p, coarse_max=coarse_max, level0_tile=16) for p in params]
    default = _build(params)
    for a, d in zip(alone, default):
        # the initial Picard solves of build_domain, large tiles against default ones
        assert np.allclose(a.V, d.V, rtol=0.0, atol=1e-8 * np.abs(d.V).max())
    batch = BatchedSimulator(together, params)
    for step in range(1, 7):
        print(f"[gpu] large-tile batch (coarse_max {coarse_max}) step {step}", flush=True)
        for s, p in zip(alone, params):
            s.advance(p, 1)
        batch.advance(1)
        for c, (a, b) in enumerate(zip(alone, together)):
            _same(a, b, (step, c))
    assert DigitalTwinSimulator().poisson_level0_tile is None


def test_held_population_in_a_batch_matches_alone():
    """A case with a population hold (top-up to its preload each step) steps
    in a batch exactly as alone, and stays held."""
    from thruster_cx_code_amd.batch import BatchedSimulator
    params = _cases()[:2]
    params[0]["preload"]["hold"] = True
    alone = _build(params)
    together = _build(params)
    batch = BatchedSimulator(together, params)
    print("[gpu] batch with a held case", flush=True)
    for step in range(6):
        for s, p in zip(alone, params):
            s.advance(p, 1)
        batch.advance(1)
    for c, (a, b) in enumerate(zip(alone, together)):
        _same(a, b, c)
    # every step pushed at least the held populations
    from thruster_cx_code_amd.physics_engine import RING
    nin = together[0].ring_nin.view(-1, 2).cpu().numpy()
    held = (params[0]["preload"]["ions"], params[0]["preload"]["electrons"])
    for it in range(1, 7):
        assert tuple(nin[it % RING]) >= held


def test_cases_join_and_leave_between_steps():
    """A case admitted later (its own iteration counter, opposite Poisson
    parity) and a case retired midway match their alone twins."""
    from thruster_cx_code_amd.batch import BatchedSimulator
    params = _cases()[:2]
    alone = _build(params)
    together = _build(params)
    batch = BatchedSimulator([together[0]], [params[0]])
    active = {0}
    for step in range(1, 10):
        if step == 4:
            batch.add(together[1], params[1])
            active.add(1)
        if step == 8:
            batch.remove(together[0])
            active.discard(0)
        print(f"[gpu] dynamic batch step {step} cases {sorted(active)}", flush=True)
        batch.advance(1)
        for c in active:
            alone[c].advance(params[c], 1)
            _same(alone[c], together[c], (step, c))
    assert together[0].iteration == 7 and together[1].iteration == 6


class _OracleCheckedBatch:
    """Mixin for BatchedSimulator: before every batch step each case's oracle
    is fed the case's state; after it the oracle steps too and the two are
    compared (particles bit-exact on steps without a Poisson solve, within
    1e-6 otherwise; potential within 1e-9 of SuperLU on solve steps)."""

    def _check_init(self):
        self.oracles = {}
        self.checked = 0

    def advance(self, n_steps):
        from oracle.sim import OracleSimulator
        from test_gpu_step import _compare_electrons, _compare_ions, _sync
        for _ in range(int(n_steps)):
            pairs = []
            for sim, p in zip(self.sims, self.params):
                o = self.oracles.get(id(sim))
                if o is None:
                    o = OracleSimulator()
                    o.m_e = sim.m_e
                    o.build_domain(p)
                    self.oracles[id(sim)] = o
                _sync(o, sim)
                pairs.append((sim, o, p))
            super().advance(1)
            for sim, o, p in pairs:
                o.step(p)
                rec = sim.record()
                even = sim.iteration % 2 == 0
                assert rec["n_sub_ion"] == o.n_sub_ion and rec["n_sub_e"] == o.n_sub_e
                if even:
                    err = np.max(np.abs(sim.V - o.V)) / np.max(np.abs(o.V))
                    assert err < 1e-9
                _compare_ions(sim, o, exact=not even)
                _compare_electrons(sim, o, exact=not even)
                if not even:
                    assert rec["entered"] == o.entered_optics_step
                    assert rec["crossed"] == o.exit_count_step
                    assert rec["lost_to_grid"] == o.lost_to_grid_step
                    assert rec["mcc_events"] == int(np.count_nonzero(o.mcc_mask))
                self.checked += 1


def _checked_batch_class():
    from thruster_cx_code_amd.batch import BatchedSimulator

    class Checked(_OracleCheckedBatch, BatchedSimulator):
        def __init__(self, *a, **kw):
            BatchedSimulator.__init__(self, *a, **kw)
            self._check_init()
    return Checked


def test_batched_steps_match_oracle():
    """Three uneven cases (one periodic) advanced as one batch: every case's
    step equals the oracle's step on the same state, odd and even steps."""
    params = _cases()
    sims = _build(params)
    batch = _checked_batch_class()(sims, params)
    for step in range(1, 5):
        print(f"[gpu] batch vs oracle step {step}", flush=True)
        batch.advance(1)
    assert batch.checked == 12
