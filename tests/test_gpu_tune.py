"""The autotuner's learned cost model on the GPU: trained on measured QC
encode times over a grid of (size, knob) points, it picks knobs for held-out
sizes whose measured time is within 15% of the exhaustive-search best and no
worse than the built-in rule, while replacing 16 timed runs per new size by
one model evaluation.  The gain over the rule, over an untrained network's
pick and over exhaustive search (tuning time) is printed."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_tuned_qc_knobs_near_exhaustive(ctx):
    from cuot_amd.tune import QC_DEFAULT, Autotuner, CostModel, QC_SPACE, qc_features

    tu = Autotuner(ctx, reps=3)
    train_sizes = [1 << 14, 3 << 13, 1 << 16, 3 << 15, 1 << 18, 3 << 17, 1 << 19]
    print("collecting QC timings", flush=True)
    tu.collect_qc(train_sizes)
    loss = tu.train("qc", epochs=1500)
    print(f"cost model training loss {loss:.4f}", flush=True)
    untrained = CostModel(ctx, 4, 32, seed=123)
    untrained._ensure_norm(np.array(tu.data["qc"][0]), np.array(tu.data["qc"][1]))
    worst = 0.0
    for n in [5 << 14, 5 << 16, 7 << 16]:
        t0 = time.perf_counter()
        k = tu.choose_qc(n)
        t_choose = time.perf_counter() - t0
        ku = QC_SPACE[int(np.argmin(untrained.predict(np.array([qc_features(n, q) for q in QC_SPACE]))))]
        best_k, best_ms, all_ms, t_exh = tu.exhaustive_qc(n)
        tuned = tu.measure_qc(n, k)
        default = tu.measure_qc(n, QC_DEFAULT)
        untr = tu.measure_qc(n, ku)
        print(f"n={n}: tuned {k} {tuned:.3f} ms | rule {default:.3f} ms | untrained pick {untr:.3f} ms | "
              f"exhaustive {best_k} {best_ms:.3f} ms (median of all {np.median(all_ms):.3f}) | "
              f"choice {t_choose * 1e3:.2f} ms vs exhaustive {t_exh:.2f} s", flush=True)
        worst = max(worst, tuned / best_ms)
        assert tuned <= 1.15 * best_ms
        assert tuned <= 1.05 * default
    print(f"worst tuned/best ratio {worst:.3f}", flush=True)
