"""FP32 against FP64: the oracle with float32 particle state (the reference's
NumPy dtype semantics) and with float64 state, free-running on the same seeds
and random streams, agree on the sweep observables (beamlet current, theta95,
saddle potential, accel-grid impingement) within max(3 standard errors,
stated tolerance); see tests/ensemble.py."""
from ensemble import compare, run_oracle, window_means


def test_fp32_state_matches_fp64_state_in_ensemble():
    seeds = [1, 2, 3, 4]
    a = [window_means(run_oracle(s, 50, "ref32"), 10) for s in seeds]
    b = [window_means(run_oracle(s, 50, "fp64"), 10) for s in seeds]
    for k, (ma, mb, d, bound, ok) in compare(a, b).items():
        assert ok, (k, ma, mb, d, bound)
