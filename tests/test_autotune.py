"""The autotuner's model and its choice of class bounds (CPU; the
measurements themselves need the GPU)."""
import numpy as np

from gpupairhmm_amd import autotune as A


def _rates(base=7.5e12):
    # a plausible slot-rate model: larger tiles a little faster per slot
    r = {}
    for rows in A.ROWS:
        for form, k in (("fast", 1.0), ("safe", 0.85)):
            r[(rows, form)] = base * k * (0.9 + 0.005 * min(rows, 20))
    return r


def test_candidates_hold_the_class():
    for b in (15, 100, 151, 300, 1023):
        for g, r in A.candidates(b):
            assert 1 <= g <= 64 and g * r >= b + 1 and r in A.ROWS


def test_workload_lengths_is_a_distribution():
    p = A.workload_lengths()
    assert len(p) == 1024 and abs(p.sum() - 1) < 1e-12 and p[0] == 0 and np.all(p >= 0)
    assert p[150] > p[600] and p[1023] > 0  # the mix is heavy at 150, every length served


def test_choose_bounds_covers_every_length_with_fitting_tiles():
    rates = _rates()
    p = A.workload_lengths()
    table, cost = A.choose_bounds(rates, p, k=12)
    bounds = [b for b, _, _ in table]
    assert bounds == sorted(set(bounds)) and bounds[-1] == 1023 and len(table) <= 12
    for b, g, r in table:
        assert g * r >= b + 1 and 1 <= g <= 64
    # no class spans lengths below its floor, whatever the weights
    for (a, _, _), (b, _, _) in zip([(0, 0, 0)] + table[:-1], table):
        assert a >= A.class_floor(b)
    # the chosen table is no worse under the model than any fixed table,
    # e.g. the probe bounds with their cheapest tiles
    probe = []
    for b in A.BOUNDS:
        g, r = min(A.candidates(b), key=lambda c: A.model_cost(c[0], c[1], "fast", rates))
        probe.append((g, r, b))
    assert cost <= A.table_cost(probe, rates, p) * (1 + 1e-9) or len(A.BOUNDS) > 12
    # more classes never cost more
    _, cost32 = A.choose_bounds(rates, p, k=32)
    assert cost32 <= cost * (1 + 1e-12)
    assert cost32 <= A.table_cost(probe, rates, p) * (1 + 1e-9)


def test_fit_rates_inverts_the_model():
    # measurements made from a known slot rate come back as that rate
    sr = 7.0e12
    meas = []
    for b in (95, 151, 255):
        for g, r in A.candidates(b):
            for hl in (150, 1000):
                gcups = sr * A.lane_eff(g) * (hl / (hl + g)) * b / (g * r) / 1e9
                meas.append({"bound": b, "lanes": g, "rows": r, "hap_len": hl, "form": "fast",
                             "gcups": gcups, "seconds": 1.0})
    fitted = A.fit_rates(meas)  # in Gcells/s, like the measured GCUPS
    for (r, f), v in fitted.items():
        assert abs(v * 1e9 - sr) / sr < 1e-9
