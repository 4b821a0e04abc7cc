"""Text summaries of rocprofv3 output (rocpd SQLite databases) for the
profiles kept with the benchmark.

    python -m thruster_cx_code_amd.rocprof_summary kernels DB --cmd "..." > out.txt
    python -m thruster_cx_code_amd.rocprof_summary traffic --fetch DB --write DB \
        --copy-fetch DB --copy-write DB --kernel k_electron_step --last 100 --cmd "..." > out.txt

`kernels` lists per-kernel dispatch counts and durations from the kernel
trace (the figures `--stats` prints).  `traffic` averages the
FETCH_SIZE and WRITE_SIZE counters (KiB per dispatch) over the last `--last`
dispatches of one kernel and scales each by the ratio of true bytes to counted
bytes measured on a device-to-device copy of known size (`bench.py
--copy-calibration`), giving HBM bytes per launch.
"""
import argparse
import glob
import os
import sqlite3
import sys


def _db(path):
    if os.path.isdir(path):
        found = sorted(glob.glob(os.path.join(path, "**", "*.db"), recursive=True))
        if not found:
            sys.exit(f"no .db under {path}")
        return found[0]
    return path


def kernels(db, top=30, tail=None):
    con = sqlite3.connect(_db(db))
    rows = con.execute("select name, count(*), sum(duration), avg(duration), min(duration), "
                       "max(duration) from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows)
    lines = [f"{'kernel':70s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>9s} {'min_us':>9s} "
             f"{'max_us':>9s} {'pct':>6s}"]
    for name, calls, tot, avg, mn, mx in rows[:top]:
        lines.append(f"{name[:70]:70s} {calls:7d} {tot / 1e6:10.3f} {avg / 1e3:9.2f} {mn / 1e3:9.2f} "
                     f"{mx / 1e3:9.2f} {100.0 * tot / total:6.2f}")
    lines.append(f"{'(all kernels)':70s} {sum(r[1] for r in rows):7d} {total / 1e6:10.3f}")
    if tail:
        sub, n = tail.rsplit(":", 1)
        d = [r[0] for r in con.execute("select duration from kernels where name like ? "
                                       "order by start", (f"%{sub}%",))][-int(n):]
        if d:
            lines.append(f"{sub}: average of the last {len(d)} dispatches (the timed region): "
                         f"{sum(d) / len(d) / 1e3:.2f} us")
    return lines


def _counter(db, counter, kernel_sub):
    con = sqlite3.connect(_db(db))
    rows = con.execute("select kernel_name, dispatch_id, value from counters_collection "
                       "where counter_name = ? order by dispatch_id", (counter,)).fetchall()
    return [(n, d, v) for n, d, v in rows if kernel_sub is None or kernel_sub in n]


def _copy_factor(db, counter, true_bytes):
    rows = _counter(db, counter, None)
    if not rows:
        sys.exit(f"no {counter} rows in {db}")
    # the calibration copies are the largest dispatches of the run
    vals = sorted((v for _, _, v in rows), reverse=True)
    reps = max(1, min(len(vals), 10))
    counted = sum(vals[:reps]) / reps * 1024.0
    return true_bytes / counted, counted


def traffic(a):
    lines = []
    f = _counter(a.fetch, "FETCH_SIZE", a.kernel)[-a.last:]
    w = _counter(a.write, "WRITE_SIZE", a.kernel)[-a.last:]
    if not f or not w:
        sys.exit("kernel not found in the counter databases")
    fk = sum(v for _, _, v in f) / len(f)
    wk = sum(v for _, _, v in w) / len(w)
    ff, fcnt = _copy_factor(a.copy_fetch, "FETCH_SIZE", a.copy_bytes)
    wf, wcnt = _copy_factor(a.copy_write, "WRITE_SIZE", a.copy_bytes)
    t = fk * 1024.0 * ff + wk * 1024.0 * wf
    lines.append(f"kernel: {f[0][0]}")
    lines.append(f"dispatches averaged: last {len(f)} (FETCH_SIZE), last {len(w)} (WRITE_SIZE)")
    lines.append(f"FETCH_SIZE_KiB_per_dispatch: {fk:.1f}")
    lines.append(f"WRITE_SIZE_KiB_per_dispatch: {wk:.1f}")
    lines.append(f"copy calibration: {a.copy_bytes} B read and written per copy; FETCH_SIZE counts "
                 f"{fcnt:.0f} B -> factor {ff:.3f}; WRITE_SIZE counts {wcnt:.0f} B -> factor {wf:.3f}")
    lines.append(f"read_bytes_per_launch: {fk * 1024.0 * ff:.0f}")
    lines.append(f"write_bytes_per_launch: {wk * 1024.0 * wf:.0f}")
    lines.append(f"traffic_bytes_per_launch: {t:.0f}")
    return lines


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("--cmd", required=True)
    k.add_argument("--top", type=int, default=30)
    k.add_argument("--tail", help="KERNEL:N -- also average the last N dispatches of KERNEL")
    t = sub.add_parser("traffic")
    for name in ("fetch", "write", "copy_fetch", "copy_write"):
        t.add_argument("--" + name.replace("_", "-"), dest=name, required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--last", type=int, default=100)
    t.add_argument("--copy-bytes", type=int, default=512 << 20)
    t.add_argument("--cmd", required=True)
    a = ap.parse_args()
    out = [f"command: {a.cmd}"]
    out += kernels(a.db, a.top, a.tail) if a.mode == "kernels" else traffic(a)
    print("\n".join(out))


if __name__ == "__main__":
    main()
