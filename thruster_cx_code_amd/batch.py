"""Batched sweep engine: many independent DigitalTwinSimulator cases (meshes,
voltages, apertures and particle counts may all differ) advanced in
lockstep, one kernel launch per phase of the time step serving every case.

The reference runs sweep cases one process per case
(Python/benchmarks/benchmark_perveance.py:102-113, sweep loop :11-75).  Here
each case keeps its own device state (built by its own build_domain), and a
step of the batch issues the same sequence of launches as one case's step,
each over all cases: the case index is the launch's z dimension and the
per-case argument structs are uploaded once per step from pinned memory.
Every case's results are bit-identical to stepping it alone.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from .device import ptr, stream_ptr
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
        arr = (ctype * n).from_address(self.host[self.slot].data_ptr() + off)
        return arr, C.c_void_p(self.dev.data_ptr() + off)

    def commit(self):
        n = (self.off + 255) & ~255
        self.dev[:n].copy_(self.host[self.slot][:n], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self.events[self.slot] = ev


class BatchedSimulator:
    """Lockstep time steps of several built DigitalTwinSimulator cases.

    sims: the cases (each already build_domain'ed, on the same device);
    params: the params dict of each case (the step reads them like step()).
    Poisson runs on every case on the same (even) steps, as in the reference
    schedule; all cases share the iteration counter parity because they
    advance together.
    """

    def __init__(self, sims, params):
        if not sims:
            raise ValueError("BatchedSimulator needs at least one case")
        if len(params) != len(sims):
            raise ValueError("one params dict per case")
        self.sims = list(sims)
        self.params = list(params)
        self.n = len(self.sims)
        self.device = self.sims[0].device
        self.lib = _lib.load()
        its = {s.iteration for s in self.sims}
        if len(its) != 1:
            raise ValueError("cases must start at the same iteration")
        for s, p in zip(self.sims, self.params):
            if not s._built:
                raise ValueError("every case must be built (build_domain) first")
            s._prepare_step_params(p)
        handles = (C.c_void_p * self.n)(*[s._poisson.value for s in self.sims])
        self._batch = C.c_void_p()
        check(self.lib.bemcs_poisson_batch_create(handles, self.n, C.byref(self._batch)),
              "poisson_batch_create")
        self.args = _ArgRing(self.device, nbytes=max(1 << 20, 8192 * self.n))
        self._div = (C.c_float * (2 * self.n))()
        for c, s in enumerate(self.sims):
            self._div[2 * c] = float(np.float32(s.dx))
            self._div[2 * c + 1] = float(np.float32(s.dy))
        self.max_nodes = max(s.nx * s.ny for s in self.sims)
        self.poisson_iter_log = []
        self.kernel_timing = None

    def __del__(self):
        b = getattr(self, "_batch", None)
        if b is not None and b.value:
            try:
                self.lib.bemcs_poisson_batch_free(b)
            except Exception:
                pass

    # ------------------------------------------------------------ step
    def advance(self, n_steps):
        for _ in range(int(n_steps)):
            self._step()

    def _picard(self):
        ps = (_lib.Picard * self.n)(*[s._picard_params(5, p) for s, p in zip(self.sims, self.params)])
        its = (C.c_int32 * self.n)()
        rel = (C.c_double * self.n)()
        check(self.lib.bemcs_poisson_batch_picard(self._batch, ps, its, rel, stream_ptr()),
              "poisson_batch_picard")
        for s, it, r in zip(self.sims, its, rel):
            s.last_poisson_iters, s.last_poisson_relres = int(it), float(r)
            s.poisson_iter_log.append(int(it))
        self.poisson_iter_log.append(list(its))

    def _step(self):
        sims, n, lib = self.sims, self.n, self.lib
        s = stream_ptr()
        A = self.args
        A.begin()
        # ---- host planning: bounds, injection plans, capacity (before any launch)
        plans = []
        for sim, p in zip(sims, self.params):
            sim.iteration += 1
            sim._refresh_bounds()
            pl = sim._inject_plan(p, sim.iteration * sim.dt)
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
                (sim._bi if key[0] else sim._be).grow(cnt)
                max_n = max(max_n, int(cnt))
            inj_launch.append((dptr, len(members), max_n))
        # ---- per-step argument structs of every case
        it = sims[0].iteration
        slot = it % RING
        dsz = C.sizeof(_lib.IonDiag)
        sb, sb_d = A.alloc(_lib.StepBegin, n)
        sub, sub_d = A.alloc(_lib.SubstepsCase, 2 * n)
        ic, ic_d = A.alloc(_lib.IonCase, n)
        ec, ec_d = A.alloc(_lib.ElectronCase, n)
        pc, pc_d = A.alloc(_lib.PpcCase, n)
        qc, qc_d = A.alloc(_lib.PercentileCase, n)
        se, se_d = A.alloc(_lib.StepEnd, n)
        sort_i = sims[0].sort_interval > 0 and it % sims[0].sort_interval == 0
        sort_e = sims[0].electron_sort_interval > 0 and it % sims[0].electron_sort_interval == 0
        si = A.alloc(_lib.SortCase, n) if sort_i else None
        sx = A.alloc(_lib.SortCase, n) if sort_e else None
        max_bi = max_be = 0
        for c, sim in enumerate(sims):
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
            for k, (sp, qm, off) in enumerate(((I0, sim.q_ion / sim.m_ion, 0), (E0, -sim.q / sim.m_e, 8))):
                u = sub[2 * c + k]
                u.x, u.y, u.vx, u.vy, u.vz = (sp.x.data_ptr(), sp.y.data_ptr(), sp.vx.data_ptr(),
                                              sp.vy.data_ptr(), sp.vz.data_ptr())
                u.n_dev, u.Exy, u.nx, u.ny = sp.count.data_ptr(), sim.Exy.data_ptr(), sim.nx, sim.ny
                u.dx, u.dy, u.q_m, u.dt, u.dt2, u.frac = sim.dx, sim.dy, qm, sim.dt, sim.dt ** 2, 0.25
                u.ds_max_bits, u.n_sub_out = dsb + off, nsb + off // 2
            ip = sim._ion_params
            ip.Exy, ip.is_bound, ip.n_sub = ptr(sim.Exy), ptr(sim.is_bound_d), C.c_void_p(nsb)
            ip.cnt_next, ip.ppc = ptr(sim.cnt_i_next), ptr(sim.ppc_d)
            ip.crossed_angle, ip.diag = ptr(sim.crossed_angle), C.c_void_p(diag_ptr)
            ip.e_cnt_next = ptr(sim.cnt_e_next)
            ip.step = it & 0xFFFFFFFF
            setattr(ic[c], "in", I0.struct())
            ic[c].out, ic[c].e_out, ic[c].p = I1.struct(), E1.struct(), ip
            ep = sim._elec_params
            ep.Exy, ep.is_bound, ep.n_sub = ptr(sim.Exy), ptr(sim.is_bound_d), C.c_void_p(nsb + 4)
            ep.cnt_next, ep.overflow = ptr(sim.cnt_e_next), ptr(sim.overflow_d)
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
        for c, sim in enumerate(sims):
            # sort cases read the post-step buffers (index 1 before the swap)
            if sort_i:
                o = si[0][c]
                setattr(o, "in", sim.ions[1].struct())
                o.out, o.n_bound = sim.ions[0].struct(), sim.bound_i
                o.nx, o.ny, o.dx, o.dy = sim.nx, sim.ny, sim.dx, sim.dy
                o.scratch = sim.sort_scratch_for(max(sim.bound_i, sim.bound_e)).data_ptr()
            if sort_e:
                o = sx[0][c]
                setattr(o, "in", sim.elec[1].struct())
                o.out, o.n_bound = sim.elec[0].struct(), sim.bound_e
                o.nx, o.ny, o.dx, o.dy = sim.nx, sim.ny, sim.dx, sim.dy
                o.scratch = sim.sort_scratch_for(max(sim.bound_i, sim.bound_e)).data_ptr()
        A.commit()
        # ---- launches
        for dptr, cnt, max_n in inj_launch:
            check(lib.bemcs_inject_batched(dptr, cnt, max_n, s), "inject_batched")
        if it % 2 == 0:
            kt = self.kernel_timing
            ev = DigitalTwinSimulator._timed_begin(kt)
            self._picard()
            DigitalTwinSimulator._timed_end(kt, "poisson_picard", ev)
        check(lib.bemcs_step_begin_batched(sb_d, n, self.max_nodes, s), "step_begin_batched")
        check(lib.bemcs_substeps_batched(sub_d, 2 * n, max(max_bi, max_be), s), "substeps_batched")
        kt = self.kernel_timing
        ev = DigitalTwinSimulator._timed_begin(kt)
        check(lib.bemcs_ion_step_batched(ic_d, n, max_bi, self._div, s), "ion_step_batched")
        DigitalTwinSimulator._timed_end(kt, "ion_step", ev)
        ev = DigitalTwinSimulator._timed_begin(kt)
        check(lib.bemcs_electron_step_batched(ec_d, n, max_be, self._div, s), "electron_step_batched")
        DigitalTwinSimulator._timed_end(kt, "electron_step", ev)
        check(lib.bemcs_ppc_stats_batched(pc_d, n, s), "ppc_stats_batched")
        check(lib.bemcs_percentile_batched(qc_d, n, s), "percentile_batched")
        check(lib.bemcs_step_end_batched(se_d, n, s), "step_end_batched")
        for sim in sims:
            sim.ions.reverse()
            sim.elec.reverse()
            sim.cnt_i, sim.cnt_i_next = sim.cnt_i_next, sim.cnt_i
            sim.cnt_e, sim.cnt_e_next = sim.cnt_e_next, sim.cnt_e
            sim.ppc_steps_count += 1
        if sort_i:
            check(lib.bemcs_sort_by_cell_batched(si[1], n, post_bi, self.max_nodes, s), "sort_batched")
            for sim in sims:
                sim.ions.reverse()
        if sort_e:
            check(lib.bemcs_sort_by_cell_batched(sx[1], n, post_be, self.max_nodes, s), "sort_batched")
            for sim in sims:
                sim.elec.reverse()
        # asynchronous count readback for the host bounds (one event for all cases)
        ev = torch.cuda.Event()
        cnts = []
        for sim in sims:
            h = torch.empty(2, dtype=torch.int32, pin_memory=True)
            h.copy_(sim.ring_counts[2 * slot:2 * slot + 2], non_blocking=True)
            cnts.append(h)
        ev.record()
        for sim, h in zip(sims, cnts):
            sim._pending.append((ev, h, sim._bi.snapshot(), sim._be.snapshot(), it))

    # ------------------------------------------------------ diagnostics
    def records(self):
        """Diagnostics of the last step of every case (synchronising)."""
        return [s.record() for s in self.sims]
