"""Run histories and their exports, as the GUI records and writes them
(reference Python/main.py:2172-2193 for the per-step records, :2316-2395 for
the history CSV, :2397-2470 for the particle-tracking CSV).

``StepHistory.record(sim, step_result)`` appends one step's values (after
``sim.step``); ``export_csv`` writes the GUI's columns; ``TrackingBuffer``
keeps ``get_particle_kinematics`` snapshots and writes them with the GUI's
8-column header.  Energy conservation tracking is
``DigitalTwinSimulator.get_total_energy`` (history lists and the 5 % warning
live on the simulator, as in the reference).
"""
import csv

import numpy as np

HISTORY_HEADER = ['iteration', 't_sim_s', 'minpotential', 'beamdivergence°', 'transparency',
                  'total_active_cells', 'cells_less_than_3_macroparticles']
TRACKING_HEADER = ['time_s', 'x_mm', 'y_mm', 'vx_ms', 'vy_ms', 'vz_ms', 'energy_eV', 'type']


class StepHistory:
    def __init__(self):
        self.iter_history, self.time_history = [], []
        self.ebs_history, self.div_history = [], []
        self.transparency_history, self.transparency3_history = [], []
        self.active_cells_history, self.low_ppc_cells_history = [], []
        self.T_histories = []

    def record(self, sim, step_result):
        """step_result: the 5-tuple of DigitalTwinSimulator.step."""
        _, min_pot, current_div, T_grids, trans_last_frame = step_result
        self.iter_history.append(sim.iteration)
        self.time_history.append(sim.iteration * sim.dt)
        self.ebs_history.append(min_pot)
        self.div_history.append(current_div)
        self.transparency_history.append(sim.get_transparency())
        self.transparency3_history.append(trans_last_frame)
        self.active_cells_history.append(sim.total_active_cells)
        self.low_ppc_cells_history.append(sim.low_ppc_cells)
        while len(self.T_histories) < len(T_grids):
            self.T_histories.append([])
        for i, T in enumerate(T_grids):
            self.T_histories[i].append(T)

    def __len__(self):
        return len(self.iter_history)

    def export_csv(self, path):
        """History CSV with the GUI's header and row layout; missing values
        are written as empty fields."""
        n = len(self.iter_history)
        if n == 0:
            raise ValueError("no simulation history to export")

        def at(seq, j):
            return seq[j] if j < len(seq) else ''
        with open(path, 'w', newline='', encoding='utf-8') as f:
            w = csv.writer(f)
            w.writerow(HISTORY_HEADER + [f"grid_{i + 1}_temp_K" for i in range(len(self.T_histories))])
            for j in range(n):
                row = [self.iter_history[j], at(self.time_history, j), self.ebs_history[j],
                       self.div_history[j], at(self.transparency_history, j),
                       at(self.active_cells_history, j), at(self.low_ppc_cells_history, j)]
                row += [at(Th, j) for Th in self.T_histories]
                w.writerow(row)
        return path


class TrackingBuffer:
    """Snapshots of get_particle_kinematics (time, x, y, vx, vy, vz,
    energy_eV, type: 0 primary ion, 1 CEX ion, 2 electron x <= 4 mm,
    3 electron beyond)."""

    def __init__(self):
        self.chunks = []

    def append(self, sim):
        data = sim.get_particle_kinematics()
        if data.size > 0:
            self.chunks.append(data)

    def export_csv(self, path):
        if not self.chunks:
            raise ValueError("no particle tracking data recorded")
        data = np.vstack(self.chunks)
        ncols = data.shape[1] if data.ndim > 1 else 0
        if ncols == 8:
            header = TRACKING_HEADER
        elif ncols == 6:
            header = ['x_mm', 'y_mm', 'vx_ms', 'vy_ms', 'vz_ms', 'isCEX']
        else:
            header = [f"col_{c}" for c in range(ncols)]
        with open(path, 'w', newline='', encoding='utf-8') as f:
            w = csv.writer(f)
            w.writerow(header)
            w.writerows(data.tolist())
        return path
