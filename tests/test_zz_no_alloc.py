"""pmx_sys_compute_noalloc (the force pass with its per-pass scratch in the
library's static per-device array, no device allocation) against
pmx_sys_compute on the same systems, called through the C ABI: two calls in
a row, the second on a smaller system, then the first system again (the
array keeps the previous call's contents).  The mixed form is bit-exact
(fixed-point accumulation, fixed reduction order); the FP64 form sums forces
with FP64 atomics in arrival order, so it is compared at 1e-12 relative."""
import numpy as np
import pytest

from protomol3_amd import pmx, systems

pytestmark = pytest.mark.gpu


def _spec_pme(rc, grid):
    return pmx.ForceSpec(6.0, lj=(pmx.SW_C2, rc - 2.0, rc), pme=dict(rc=rc, grid=grid, order=4, alpha=-1))


def _run(sy, pos, noalloc):
    sy.set_positions(pos)
    if noalloc:
        sy.compute_noalloc(pmx.EVAL_PAIRSTATS)
    else:
        sy.compute(pmx.EVAL_PAIRSTATS)
    f, ec, elj = sy.get_forces()
    return f, ec, elj, sy.pair_stats()


@pytest.mark.parametrize("precision", [pmx.MIXED, pmx.FP64])
def test_noalloc_matches_compute_twice_in_a_row(precision):
    big = systems.polymer_water(n_total=3200, box=32.0, n_beads=50, seed=4, scaled14=0.5)
    small = systems.water_box(seed=9)
    ctx = pmx.Context(0)
    try:
        sa = pmx.System(ctx, big, _spec_pme(12.0, (32, 32, 32)), precision)
        sb = pmx.System(ctx, small, _spec_pme(10.0, (24, 24, 24)), precision)
        try:
            ref_a = _run(sa, big.pos, False)
            ref_b = _run(sb, small.pos, False)
            got = [("big", _run(sa, big.pos, True), ref_a), ("small", _run(sb, small.pos, True), ref_b),
                   ("big again", _run(sa, big.pos, True), ref_a)]
        finally:
            sa.close()
            sb.close()
    finally:
        ctx.close()
    for what, (f, ec, elj, st), (fr, ecr, eljr, str_) in got:
        print(f"noalloc {what} precision={precision} Ecoul {ec!r} vs {ecr!r} pairs {st[0]} vs {str_[0]}",
              flush=True)
        assert st == str_
        if precision == pmx.MIXED:
            assert np.array_equal(f, fr)
            assert ec == ecr and elj == eljr
        else:
            assert np.abs(f - fr).max() <= 1e-12 * np.abs(fr).max()
            assert abs(ec - ecr) <= 1e-12 * abs(ecr) and abs(elj - eljr) <= 1e-12 * abs(eljr)


def test_noalloc_rejects_timing_and_uncached():
    s = systems.water_box(seed=9)
    ctx = pmx.Context(0)
    try:
        sy = pmx.System(ctx, s, _spec_pme(10.0, (24, 24, 24)), pmx.MIXED)
        try:
            sy.set_positions(s.pos)
            with pytest.raises(pmx.PmxError) as e:
                sy.compute_noalloc(pmx.EVAL_TIMING)
            assert e.value.code == -1
            ctx.lib.pmx_sys_uncache(sy.h)
            with pytest.raises(pmx.PmxError) as e:
                sy.compute_noalloc(0)
            assert e.value.code == -1
            # the allocating form rebuilds the tables; then the no-allocation form runs again
            sy.compute(0)
            sy.get_forces()
            sy.compute_noalloc(0)
            sy.get_forces()
        finally:
            sy.close()
    finally:
        ctx.close()
