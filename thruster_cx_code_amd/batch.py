"""Batched sweep engine: many independent DigitalTwinSimulator cases (meshes,
voltages, apertures and particle counts may all differ) advanced in
lockstep, one kernel launch per phase of the time step serving every case.

The reference runs sweep cases one process per case
(Python/benchmarks/benchmark_perveance.py:102-113, sweep loop :11-75).  Here
each case keeps its own device state (built by its own build_domain), and a
step of the batch issues the same sequence of launches as one case's step,
each over all cases: the case index is the launch's z dimension and the
per-case argument structs are uploaded once per step from pinned memory.
The per-step kernels (step begin, substep counts, ion and electron steps,
PPC statistics, percentile, step end) run as one captured HIP graph, replayed
every step: they read their argument structs from fixed places in device
memory, and their launch sizes come from the population bounds rounded up to
a power of two, so a graph is captured again only when the membership or a
bound's power of two changes.  The Poisson update of the cases due runs as
the solver's own captured graphs.  Every case's results are bit-identical to
stepping it alone.
"""
import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _lib
from .device import EventRing, stream_ptr
from .physics_engine import RING, DigitalTwinSimulator, check


class _ArgRing:
    """Per-step argument arrays: ctypes arrays built in pinned host memory
    (a ring of buffers, each reused only after the event of its copy) and
    uploaded to one device buffer before the step's first launch."""

    def __init__(self, device, nbytes=1 << 20, depth=4):
        self.device = device
        self.depth = depth
        self.host = [torch.empty(nbytes, dtype=torch.uint8, pin_memory=True) for _ in range(depth)]
        self.events = [None] * depth
        self.dev = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.slot = -1
        self.off = 0

    def begin(self):
        self.slot = (self.slot + 1) % self.depth
        ev = self.events[self.slot]
        if ev is not None:
            ev.synchronize()
        self.off = 0

    def alloc(self, ctype, n):
        """A ctypes array of n ctype in the current host buffer, and the
        device address it will have after commit()."""
        size = C.sizeof(ctype) * n
        off = (self.off + 255) & ~255
        if off + size > self.host[0].numel():
            raise RuntimeError("batch argument buffer too small")
        self.off = off + size
        addr = self.host[self.slot].data_ptr() + off
        C.memset(addr, 0, size)  # fields a caller does not set are zero (null pointers)
        arr = (ctype * n).from_address(addr)
        return arr, C.c_void_p(self.dev.data_ptr() + off)

    def commit(self):
        n = (self.off + 255) & ~255
        self.dev[:n].copy_(self.host[self.slot][:n], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self.events[self.slot] = ev


_COUNT_POOL = None


def _count_pool():
    global _COUNT_POOL
    if _COUNT_POOL is None:
        _COUNT_POOL = ThreadPoolExecutor(max_workers=max(1, min(8, (os.cpu_count() or 1) // 2)))
    return _COUNT_POOL


class BatchedSimulator:
    """Time steps of several built DigitalTwinSimulator cases, one launch per
    phase serving all of them.

    sims: the cases (each already build_domain'ed, on the same device);
    params: the params dict of each case (the step reads them like step()).
    Cases keep their own iteration counters, so they may join (add) and
    leave (remove) between steps: each case runs its Poisson update on its own
    even steps and sorts on its own schedule, as when stepped alone; the
    Poisson solves of the cases due in a step form one batched solve.
    """

    def __init__(self, sims=(), params=(), graphs=None):
        if len(params) != len(sims):
            raise ValueError("one params dict per case")
        self.sims, self.params = [], []
        self.device = None
        self.lib = _lib.load()
        self._sets = {}
        # per-step kernels as captured HIP graphs (BEMCS_BATCH_GRAPHS=0: direct launches)
        self.graphs = os.environ.get("BEMCS_BATCH_GRAPHS", "1") != "0" if graphs is None else bool(graphs)
        self._graph_cache = {}
        self._cap_stream = None
        self._divisors_checked = False
        self.graph_captures = 0
        self.args = None
        self.poisson_iter_log = []
        self.kernel_timing = None
        self._events = None  # count-readback events, made at the first step
        self._ahead = []  # pool futures: the next step's beam-ion counts
        for s, p in zip(sims, params):
            self.add(s, p)

    @property
    def n(self):
        return len(self.sims)

    def add(self, sim, params):
        """Admit a built case (between steps)."""
        if not sim._built:
            raise ValueError("every case must be built (build_domain) first")
        if params.get('sim_mode', 'Both') != 'Trajectories' or params.get('rf_enable'):
            raise ValueError("batched cases run sim_mode 'Trajectories' without RF; step thermal, "
                             "erosion and RF cases with DigitalTwinSimulator.step/advance")
        if sim.deposit_shape != 'ngp':
            raise ValueError("batched cases use the NGP deposit; step bilinear-deposit cases with "
                             "DigitalTwinSimulator.step/advance")
        sim._resolve_surface()
        if self.device is None:
            self.device = sim.device
        sim._prepare_step_params(params)
        self.sims.append(sim)
        self.params.append(params)
        self._membership_changed()

    def remove(self, sim):
        """Retire a case (between steps); its state stays valid for reading."""
        k = self.sims.index(sim)
        del self.sims[k]
        del self.params[k]
        self._membership_changed()

    def _membership_changed(self):
        self._free_sets()
        self._free_graphs()
        self._divisors_checked = False
        if self.device is None:
            return
        n = max(1, self.n)
        if self.args is None or self.args.dev.numel() < 8192 * n:
            self.args = _ArgRing(self.device, nbytes=max(1 << 20, 16384 * n))
        self._div = (C.c_float * (2 * n))()
        for c, s in enumerate(self.sims):
            self._div[2 * c] = float(np.float32(s.dx))
            self._div[2 * c + 1] = float(np.float32(s.dy))
        self.max_nodes = max((s.nx * s.ny for s in self.sims), default=1)

    def _free_sets(self):
        for b in self._sets.values():
            try:
                self.lib.bemcs_poisson_batch_free(b)
            except Exception:
                pass
        self._sets = {}

    def _free_graphs(self):
        if self._graph_cache and self.device is not None:
            torch.cuda.synchronize(self.device)  # no replay of them is still queued
        for g in self._graph_cache.values():
            try:
                self.lib.bemcs_graph_destroy(g)
            except Exception:
                pass
        self._graph_cache = {}

    def __del__(self):
        try:
            self._free_sets()
            self._free_graphs()
        except Exception:
            pass

    # ------------------------------------------------------------ step
    def advance(self, n_steps):
        for _ in range(int(n_steps)):
            self._step()

    @staticmethod
    def _chunks(items, weight, nchunk):
        """items split into nchunk lists of similar total weight (largest
        first, each to the lightest list)."""
        chunks = [[] for _ in range(nchunk)]
        load = [0.0] * nchunk
        for x in sorted(items, key=lambda x: -weight(x)):
            j = load.index(min(load))
            chunks[j].append(x)
            load[j] += weight(x) + 1.0
        return chunks

    def _ion_counts(self, lams):
        """Every case's beam-ion count of the step: the sample drawn ahead
        during the last step (_sample_ahead) where it is of this case, step
        and rate, else drawn now.  The sampler costs O(lambda) (one uniform per
        unit of the mean), so batches whose cases inject thousands of ions per
        step run it on host threads (the C sampler releases the GIL); the
        counts are the same either way."""
        ahead = {}
        for f in self._ahead:
            for sim, it, lam, cnt in f.result():
                ahead[id(sim)] = (sim, it, lam, cnt)
        self._ahead = []
        out = [None] * len(lams)
        todo = []
        for c, lam in enumerate(lams):
            if lam is None:
                continue
            sim = self.sims[c]
            a = ahead.get(id(sim))
            if a is not None and a[0] is sim and a[1] == sim.iteration and a[2] == lam:
                out[c] = a[3]
            else:
                todo.append(c)
        work = sum(lams[c] for c in todo)
        if work < 20000 or len(todo) < 2:
            for c in todo:
                out[c] = self.sims[c]._ion_injection_count(lams[c])
            return out
        pool = _count_pool()
        chunks = self._chunks(todo, lambda c: lams[c], min(len(todo), pool._max_workers))

        def run(chunk):
            return [(c, self.sims[c]._ion_injection_count(lams[c])) for c in chunk]
        for res in pool.map(run, chunks):
            for c, n in res:
                out[c] = n
        return out

    def _sample_ahead(self):
        """Draw the next step's beam-ion counts on the pool's threads.  They
        depend only on each case's seed, step and rate, so they can be drawn
        while this thread waits in the library (the Picard call releases the
        GIL); the next step takes them (_ion_counts) and draws itself any it
        does not find (a case admitted in between)."""
        jobs = []
        for sim, p in zip(self.sims, self.params):
            it = sim.iteration + 1
            lam = sim._ion_injection_lambda(p, it * sim.dt)
            if lam is not None:
                jobs.append((sim, it, lam))
        if not jobs:
            return
        pool = _count_pool()

        def run(chunk):
            return [(sim, it, lam, sim._ion_injection_count(lam, it)) for sim, it, lam in chunk]
        self._ahead = [pool.submit(run, ch)
                       for ch in self._chunks(jobs, lambda j: j[2], min(len(jobs), pool._max_workers))]

    def _picard(self, idx):
        """Batched Picard update of the cases idx (their own even steps)."""
        key = tuple(self.sims[c]._poisson.value for c in idx)
        b = self._sets.get(key)
        if b is None:
            handles = (C.c_void_p * len(idx))(*key)
            b = C.c_void_p()
            check(self.lib.bemcs_poisson_batch_create(handles, len(idx), C.byref(b)), "poisson_batch_create")
            self._sets[key] = b
        m = len(idx)
        ps = (_lib.Picard * m)(*[self.sims[c]._picard_params(5, self.params[c]) for c in idx])
        its = (C.c_int32 * m)()
        rel = (C.c_double * m)()
        check(self.lib.bemcs_poisson_batch_picard(b, ps, its, rel, stream_ptr()), "poisson_batch_picard")
        for c, it, r in zip(idx, its, rel):
            s = self.sims[c]
            s.last_poisson_iters, s.last_poisson_relres = int(it), float(r)
            s.poisson_iter_log.append(int(it))
        self.poisson_iter_log.append(list(its))

    def _step(self):
        sims, n, lib = self.sims, self.n, self.lib
        if n == 0:
            return
        s = stream_ptr()
        A = self.args
        A.begin()
        # the argument structs of the per-step kernels come first in the
        # buffer, so their device addresses (the captured graphs' arguments)
        # depend only on the number of cases
        dsz = C.sizeof(_lib.IonDiag)
        sb, sb_d = A.alloc(_lib.StepBegin, n)
        sub, sub_d = A.alloc(_lib.SubstepsCase, 2 * n)
        ic, ic_d = A.alloc(_lib.IonCase, n)
        ec, ec_d = A.alloc(_lib.ElectronCase, n)
        pc, pc_d = A.alloc(_lib.PpcCase, n)
        qc, qc_d = A.alloc(_lib.PercentileCase, n)
        se, se_d = A.alloc(_lib.StepEnd, n)
        # ---- host planning: bounds, injection plans, capacity (before any launch)
        lams = []
        for sim, p in zip(sims, self.params):
            sim.iteration += 1
            lams.append(sim._ion_injection_lambda(p, sim.iteration * sim.dt))
        counts = self._ion_counts(lams)
        plans = []
        for sim, p, n_ion in zip(sims, self.params, counts):
            sim._refresh_bounds()
            pl = sim._inject_plan(p, sim.iteration * sim.dt, n_ion=n_ion) + sim._hold_plans(p)
            sim._reserve_step(pl)
            plans.append(pl)
        # injection groups: the k-th injection into one species of every case
        groups = {}
        for c, pl in enumerate(plans):
            seen = {True: 0, False: 0}
            for is_ion, ip, cnt in pl:
                groups.setdefault((is_ion, seen[is_ion]), []).append((c, ip, cnt))
                seen[is_ion] += 1
        inj_launch = []
        for key in sorted(groups, key=lambda k: (not k[0], k[1])):
            members = groups[key]
            arr, dptr = A.alloc(_lib.InjectCase, len(members))
            max_n = 0
            for j, (c, ip, cnt) in enumerate(members):
                sim = sims[c]
                sp = sim.ions[0] if key[0] else sim.elec[0]
                arr[j].sp = sp.struct()
                arr[j].p = ip
                arr[j].n = int(cnt)
                arr[j].n_bound_before = int(sim.bound_i if key[0] else sim.bound_e)
                cb = sim._bi if key[0] else sim._be
                if ip.target > 0:
                    cb.top_up(ip.target)  # population hold: count' = max(count, target)
                else:
                    cb.grow(cnt)
                max_n = max(max_n, int(cnt))
            inj_launch.append((dptr, len(members), max_n))
        # ---- per-step argument structs of every case
        sort_i = [c for c, m in enumerate(sims) if m.sort_interval > 0 and m.iteration % m.sort_interval == 0]
        sort_e = [c for c, m in enumerate(sims)
                  if m.electron_sort_interval > 0 and m.iteration % m.electron_sort_interval == 0]
        si = A.alloc(_lib.SortCase, len(sort_i)) if sort_i else None
        sx = A.alloc(_lib.SortCase, len(sort_e)) if sort_e else None
        poisson = [c for c, m in enumerate(sims) if m.iteration % 2 == 0]
        max_bi = max_be = 0
        for c, sim in enumerate(sims):
            it = sim.iteration
            slot = it % RING
            I0, I1 = sim.ions
            E0, E1 = sim.elec
            bi, be = sim.bound_i, sim.bound_e
            max_bi, max_be = max(max_bi, bi), max(max_be, be)
            diag_ptr = sim.ring_diag.data_ptr() + slot * dsz
            nsb = sim.n_sub_d.data_ptr()
            dsb = sim.ds_bits.data_ptr()
            b = sb[c]
            b.ds_bits, b.count_out_i, b.count_out_e = dsb, I1.count.data_ptr(), E1.count.data_ptr()
            b.node_zero[0], b.node_zero[1], b.node_zero[2] = (sim.cnt_i_next.data_ptr(),
                                                              sim.cnt_e_next.data_ptr(), sim.ppc_d.data_ptr())
            b.n_nodes, b.diag, b.diag_bytes = sim.nx * sim.ny, diag_ptr, dsz
            b.ppc_stats = sim.ring_ppc.data_ptr() + 8 * slot
            b.count_in_i, b.count_in_e = I0.count.data_ptr(), E0.count.data_ptr()
            b.nin_out = sim.ring_nin.data_ptr() + 8 * slot
            b.count_out_x = None
            for k, (sp, qm, off) in enumerate(((I0, sim.q_ion / sim.m_ion, 0), (E0, -sim.q / sim.m_e, 8))):
                u = sub[2 * c + k]
                u.x, u.y, u.vx, u.vy, u.vz = (sp.x.data_ptr(), sp.y.data_ptr(), sp.vx.data_ptr(),
                                              sp.vy.data_ptr(), sp.vz.data_ptr())
                u.n_dev, u.Exy, u.nx, u.ny = sp.count.data_ptr(), sim.Exy.data_ptr(), sim.nx, sim.ny
                u.dx, u.dy, u.q_m, u.dt, u.dt2, u.frac = sim.dx, sim.dy, qm, sim.dt, sim.dt ** 2, 0.25
                u.ds_max_bits, u.n_sub_out = dsb + off, nsb + off // 2
            ip = sim._ion_params
            exy, isb = sim.Exy.data_ptr(), sim.is_bound_d.data_ptr()
            cen = sim.cnt_e_next.data_ptr()
            ip.Exy, ip.is_bound, ip.n_sub = exy, isb, nsb
            ip.cnt_next, ip.ppc = sim.cnt_i_next.data_ptr(), sim.ppc_d.data_ptr()
            ip.crossed_angle, ip.diag = sim.crossed_angle.data_ptr(), diag_ptr
            ip.e_cnt_next = cen
            ip.step = it & 0xFFFFFFFF
            setattr(ic[c], "in", I0.struct())
            ic[c].out, ic[c].e_out, ic[c].p = I1.struct(), E1.struct(), ip
            ep = sim._elec_params
            ep.Exy, ep.is_bound, ep.n_sub = exy, isb, nsb + 4
            ep.cnt_next, ep.overflow = cen, sim.overflow_d.data_ptr()
            setattr(ec[c], "in", E0.struct())
            ec[c].out, ec[c].p = E1.struct(), ep
            # after the step the output buffers become the inputs
            pc[c].ppc, pc[c].n_nodes = sim.ppc_d.data_ptr(), sim.nx * sim.ny
            pc[c].accum, pc[c].stats = sim.accum_ppc_d.data_ptr(), sim.ring_ppc.data_ptr() + 8 * slot
            q = qc[c]
            q.angles, q.n_dev, q.q_percent, q.min_count = sim.crossed_angle.data_ptr(), diag_ptr + 4, 95.0, 5
            q.out = sim.ring_theta.data_ptr() + 8 * slot
            e = se[c]
            e.count_i, e.count_e = I1.count.data_ptr(), E1.count.data_ptr()
            e.n_sub, e.counts_out = nsb, sim.ring_counts.data_ptr() + 8 * slot
            e.nsub_out = sim.ring_nsub.data_ptr() + 8 * slot
            e.V, e.saddle_idx, e.n_saddle = sim.V_d.data_ptr(), sim._saddle_idx.data_ptr(), sim._saddle_idx.numel()
            e.minpot_out = sim.ring_minpot.data_ptr() + 8 * slot
        # bounds after the step: SEE adds at most one electron per ion
        for sim in sims:
            sim._be.grow(sim.bound_i)
        post_bi = max(sim.bound_i for sim in sims)
        post_be = max(sim.bound_e for sim in sims)
        # sort cases read the post-step buffers (index 1 before the swap)
        for j, c in enumerate(sort_i):
            sim = sims[c]
            o = si[0][j]
            setattr(o, "in", sim.ions[1].struct())
            o.out, o.n_bound = sim.ions[0].struct(), sim.bound_i
            o.nx, o.ny, o.dx, o.dy = sim.nx, sim.ny, sim.dx, sim.dy
            o.scratch = sim.sort_scratch_for(max(sim.bound_i, sim.bound_e)).data_ptr()
        for j, c in enumerate(sort_e):
            sim = sims[c]
            o = sx[0][j]
            setattr(o, "in", sim.elec[1].struct())
            o.out, o.n_bound = sim.elec[0].struct(), sim.bound_e
            o.nx, o.ny, o.dx, o.dy = sim.nx, sim.ny, sim.dx, sim.dy
            o.scratch = sim.sort_scratch_for(max(sim.bound_i, sim.bound_e)).data_ptr()
        A.commit()
        # ---- launches
        for dptr, cnt, max_n in inj_launch:
            check(lib.bemcs_inject_batched(dptr, cnt, max_n, s), "inject_batched")
        self._sample_ahead()
        if poisson:
            kt = self.kernel_timing
            ev = DigitalTwinSimulator._timed_begin(kt)
            self._picard(poisson)
            DigitalTwinSimulator._timed_end(kt, "poisson_picard", ev)
        core = (sb_d, sub_d, ic_d, ec_d, pc_d, qc_d, se_d)
        if self.graphs and self.kernel_timing is None:
            self._core_graph(core, n, max_bi, max_be, s)
        else:
            self._core_launches(core, n, max_bi, max_be, s, self.kernel_timing)
        for sim in sims:
            sim.ions.reverse()
            sim.elec.reverse()
            sim.cnt_i, sim.cnt_i_next = sim.cnt_i_next, sim.cnt_i
            sim.cnt_e, sim.cnt_e_next = sim.cnt_e_next, sim.cnt_e
            sim.ppc_steps_count += 1
        if sort_i:
            check(lib.bemcs_sort_by_cell_batched(si[1], len(sort_i), post_bi, self.max_nodes, s), "sort_batched")
            for c in sort_i:
                sims[c].ions.reverse()
        if sort_e:
            check(lib.bemcs_sort_by_cell_batched(sx[1], len(sort_e), post_be, self.max_nodes, s), "sort_batched")
            for c in sort_e:
                sims[c].elec.reverse()
        # the counts reach each case's pinned ring asynchronously; one event
        # for all cases tells the host bounds when
        if self._events is None:
            self._events = EventRing()
        for sim in sims:
            if len(sim._pending) >= len(self._events) - 2:
                sim._refresh_bounds(block=True)
        ev = self._events.next()
        ev.record()
        for sim in sims:
            sim._pending.append((ev, sim.iteration % RING, sim._bi.snapshot(), sim._be.snapshot(), sim.iteration))

    def _core_launches(self, core, n, bi, be, s, kt=None):
        """The per-step kernels of every case: clears and records, substep
        counts, the fused ion and electron steps, PPC statistics, the
        divergence percentile and the step-end records.  bi and be bound the
        ion and electron counts (they size the launches; the kernels read the
        counts on the device)."""
        lib = self.lib
        sb_d, sub_d, ic_d, ec_d, pc_d, qc_d, se_d = core
        check(lib.bemcs_step_begin_batched(sb_d, n, self.max_nodes, s), "step_begin_batched")
        check(lib.bemcs_substeps_batched(sub_d, 2 * n, max(bi, be), s), "substeps_batched")
        ev = DigitalTwinSimulator._timed_begin(kt)
        check(lib.bemcs_ion_step_batched(ic_d, n, bi, self._div, s), "ion_step_batched")
        DigitalTwinSimulator._timed_end(kt, "ion_step", ev)
        ev = DigitalTwinSimulator._timed_begin(kt)
        check(lib.bemcs_electron_step_batched(ec_d, n, be, self._div, s), "electron_step_batched")
        DigitalTwinSimulator._timed_end(kt, "electron_step", ev)
        check(lib.bemcs_ppc_stats_batched(pc_d, n, s), "ppc_stats_batched")
        check(lib.bemcs_percentile_batched(qc_d, n, s), "percentile_batched")
        check(lib.bemcs_step_end_batched(se_d, n, s), "step_end_batched")

    @staticmethod
    def _pow2(b):
        b = int(b)
        return 0 if b <= 0 else 1 << (b - 1).bit_length()

    def _core_graph(self, core, n, bi, be, s):
        """Replay (capturing first if needed) the graph of _core_launches for
        the bounds rounded up to powers of two.  A larger launch only adds
        work-groups that find no particle; the results do not change."""
        key = (self._pow2(bi), self._pow2(be))
        g = self._graph_cache.get(key)
        if g is None:
            lib = self.lib
            if not self._divisors_checked:
                # the division check of a new divisor synchronises: never inside a capture
                for d in self._div:
                    lib.bemcs_fastdiv_enabled(d, s)
                self._divisors_checked = True
            if self._cap_stream is None:
                # capture cannot run on the legacy default stream; the graph is
                # launched on the step's stream
                self._cap_stream = torch.cuda.Stream(device=self.device)
            cs = C.c_void_p(self._cap_stream.cuda_stream)
            if len(self._graph_cache) >= 8:
                self._free_graphs()
            check(lib.bemcs_graph_capture_begin(cs), "graph_capture_begin")
            try:
                self._core_launches(core, n, key[0], key[1], cs)
            except BaseException:
                lib.bemcs_graph_capture_end(cs, 1, C.byref(C.c_void_p()))
                raise
            g = C.c_void_p()
            check(lib.bemcs_graph_capture_end(cs, 0, C.byref(g)), "graph_capture_end")
            self._graph_cache[key] = g
            self.graph_captures += 1
        check(self.lib.bemcs_graph_launch(g, s), "graph_launch")

    # ------------------------------------------------------ diagnostics
    def records(self):
        """Diagnostics of the last step of every case (synchronising)."""
        return [s.record() for s in self.sims]
