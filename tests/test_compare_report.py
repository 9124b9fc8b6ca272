"""The `compare` A/B error report (absF2, rFavg, rFmax, rPE and the averaged
summary of CompareForce, protomol/force/CompareForce.cpp:33-130) against the
report the reference build prints for the same two forces on the same
system (fixtures written by tests/golden/make_fixtures.py: the reference's
own SystemCompareForce pair, O/H masses)."""
import os
import re

import numpy as np
import pytest

from protomol3_amd import outputs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _numbers(line):
    return [float(v) for v in re.findall(r"(?:absF2|rFavg|rFmax|rPE) ([0-9.eE+-]+)", line)]


@pytest.mark.parametrize("name", ["compare_water1536_ljc2", "compare_water1536_pme"])
def test_compare_report_matches_the_reference(name):
    fx = np.load(os.path.join(GOLDEN, name + ".npz"))
    lines = [str(x) for x in fx["report_lines"]]
    r = outputs.compare_forces(fx["forces_a"], float(fx["pe_a"]), fx["forces_b"], float(fx["pe_b"]), fx["mass"])
    ref = _numbers(lines[0])
    ours = [r["absF2"], r["rFavg"], r["rFmax"], r["rPE"]]
    print(outputs.compare_line(0, r), "\n", lines[0])
    # the reference prints 15 significant digits
    assert np.allclose(ours, ref, rtol=2e-14, atol=0)
    assert outputs.compare_line(0, r).split(",")[0] == lines[0].split(",")[0]
    summary = outputs.compare_summary([r, r])
    ref_sum = _numbers(lines[-1])
    assert summary["n"] == int(lines[-1].rsplit("n=", 1)[1])
    assert np.allclose([summary[k][0] for k in ("absF2", "rFavg", "rFmax", "rPE")], ref_sum, rtol=2e-14, atol=0)
    assert all(summary[k][1] == 0.0 for k in ("absF2", "rFavg", "rFmax", "rPE"))


def test_compare_report_mass_weighting_and_spread():
    """Closed forms: equal masses drop out of the relative figures; a heavier
    atom's error counts 1/sqrt(m) as much; the summary's spread is the sample
    standard deviation."""
    rng = np.random.default_rng(1)
    f_ref = rng.normal(size=(10, 3))
    f = f_ref + 0.01 * rng.normal(size=(10, 3))
    r1 = outputs.compare_forces(f_ref, 2.0, f, 2.5, np.ones(10))
    r4 = outputs.compare_forces(f_ref, 2.0, f, 2.5, 4.0 * np.ones(10))
    assert np.isclose(r1["rFavg"], r4["rFavg"]) and np.isclose(r1["rFmax"], r4["rFmax"])
    assert r1["rPE"] == pytest.approx(0.2) and r1["absF2"] == pytest.approx(np.sqrt(((f - f_ref) ** 2).sum(1).max()))
    m = np.ones(10)
    m[0] = 100.0
    d = np.sqrt(((f - f_ref) ** 2).sum(1))
    g = np.sqrt((f ** 2).sum(1))
    w = 1.0 / np.sqrt(m)
    assert outputs.compare_forces(f_ref, 1.0, f, 1.0, m)["rFavg"] == pytest.approx((d * w).sum() / (g * w).sum())
    s = outputs.compare_summary([r1, dict(r1, rPE=0.4)])
    assert s["rPE"] == pytest.approx((0.3, np.std([0.2, 0.4], ddof=1)))
