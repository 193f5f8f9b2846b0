"""The autotuner's accuracy model (esp_hoomd_amd/tune.py) against the committed scan of the
oracle's FP64 ESP against its direct Ewald sum (tests/golden/accuracy_scan.json, written by
tests/golden/make_accuracy_scan.py on random +-1 electrolytes at density 0.1, 200
configurations of c, P and sigma with the default reciprocal cutoff)."""
import json
from pathlib import Path

import numpy as np

from esp_hoomd_amd import _lib, tune

SCAN = json.loads((Path(__file__).resolve().parent / "golden" / "accuracy_scan.json").read_text())


def test_split_error_decays_as_the_fitted_exponential():
    """At P = 12 and sigma ~ 2 (window error negligible) the error falls by 10^0.416 per unit
    of the prolate bandwidth c, as SPLIT_FIT states."""
    pts = [p for p in SCAN["points"] if p["P"] == 12 and p["sigma"] > 1.8 and p["c"] >= 8]
    c = np.array([p["c"] for p in pts], float)
    e = np.array([p["error"] for p in pts])
    slope, intercept = np.polyfit(c, np.log10(e), 1)
    assert abs(slope - tune.SPLIT_FIT[0]) < 0.03, slope
    # the model's split term (times its cutoff factor) is within 40 % of every point
    ratio = tune.CUT_FACTOR * np.array([tune.split_error(x) for x in c]) / e
    assert np.all((ratio > 0.6) & (ratio < 1.4)), ratio


def test_model_brackets_the_measured_errors():
    """Over the tuner's working range (errors 1e-7 to 1e-3, P >= 6) the model's prediction is
    within 0.6x-4x of the measurement (it leans conservative: median ~1.4x)."""
    ratios = []
    for p in SCAN["points"]:
        if not (1e-7 < p["error"] < 1e-3 and p["P"] >= 6):
            continue
        pred = tune.predicted_error(p["c"], p["sigma"], p["P"], _lib.ESP_PREC_FP64)
        ratios.append(pred / p["error"])
    ratios = np.array(ratios)
    assert len(ratios) > 80
    assert ratios.min() > 0.6 and ratios.max() < 4.0, (ratios.min(), ratios.max())
    assert 1.0 < np.median(ratios) < 2.0


def test_reciprocal_cutoff_costs_about_half_again_at_high_oversampling():
    """At sigma = 2 the default cutoff |k| <= 1.2 kappa raises the error by ~1.5x (nearer the
    mesh's Nyquist limit, sigma ~ 1.2 where the tuner works, the mesh cannot carry those modes
    anyway; the model's CUT_FACTOR 1.2 is the working-range value)."""
    r = np.array([x["ratio"] for x in SCAN["cut"]])
    assert np.all((r > 1.3) & (r < 1.8)), r
