"""Prints the FP32-vs-FP64 table of the sweep observables: the oracle with
float32 particle state against float64 state, free-running on the same seeds
(python tests/precision_table.py [seeds] [steps])."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from ensemble import OBSERVABLES, TOLERANCE, compare, run_oracle, window_means  # noqa: E402


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    steps = int(sys.argv[2]) if len(sys.argv) > 2 else 60
    seeds = list(range(1, n + 1))
    a = [window_means(run_oracle(s, steps, "ref32"), 10) for s in seeds]
    b = [window_means(run_oracle(s, steps, "fp64"), 10) for s in seeds]
    res = compare(a, b)
    print(f"| observable | FP32 state | FP64 state | mean paired difference | relative | bound "
          f"(max(3 SE, {'/'.join(f'{TOLERANCE[k]:g}' for k in OBSERVABLES)})) |")
    print("|---|---|---|---|---|---|")
    for k in OBSERVABLES:
        m32, m64, d, bound, ok = res[k]
        rel = abs(d) / abs(m64) if m64 else float("nan")
        print(f"| {k} | {m32:.6g} | {m64:.6g} | {d:.3g} | {rel:.2g} | {bound:.3g} |")
    print(f"({n} seeds, steps 11-{steps} of each run, 128 x 48 nodes, 50000 + 5000 preloaded)")


if __name__ == "__main__":
    main()
