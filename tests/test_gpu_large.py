"""The large water box at full size on the GPU (999,000 atoms in a 216 A box,
LJ C2 8-10 A, PME real Cutoff 10 A, 256^3 grid, order 6, alpha 0.34) against
the oracle through periodic replication.

The box is 2 x 2 x 2 copies of a 108 A box of 41,625 waters (124,875 atoms),
so every atom's force equals the force on its image in the small box and
every energy is 8 times the small box's: the cutoff (10 A) is below half the
small box, and the tiled charge grid (256^3 at the same spacing as the small
box's 128^3) has structure factors only at even indices, where they are 8
times the small grid's, with the same B-spline moduli and reciprocal vectors
(the identity itself is checked on the CPU by the oracle at a small size in
test_oracle_math).  The oracle evaluates the small box; the GPU evaluates
the full box.  Cell assignment is compared with the oracle on the full box.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from protomol3_amd import pmx, systems

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-9
MIXED_RMS = 1e-5
# Mixed-form energies at this size: about 2e8 pair terms and 1.7e7 grid
# points, each rounded in FP32 (~6e-8 relative) with both signs, against
# totals that are small differences of large terms; the bound is per atom.
MIXED_E_PER_ATOM = 1e-6
LJ = (O.SW_C2, 8.0, 10.0)
PME_SMALL = dict(rc=10.0, grid=(128, 128, 128), order=6, alpha=0.34)
PME_BIG = dict(rc=10.0, grid=(256, 256, 256), order=6, alpha=0.34)
CELLSIZE = 5.0


@pytest.fixture(scope="module")
def ctx():
    c = pmx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def boxes():
    small = systems.big_water(n_water=41625, box=108.0, seed=2)
    big = systems.tiled(small, (2, 2, 2))
    ospec = O.make_spec(CELLSIZE, lj=LJ, pme=dict(PME_SMALL, real=True, recip=True, corr=True))
    f, ec, elj = O.evaluate(small, ospec, nthreads=min(16, os.cpu_count() or 1))
    npairs = len(O.pairs(small, CELLSIZE, 10.0))
    print(f"oracle small box: n {small.n} Ecoul {ec:.12g} Elj {elj:.12g} pairs {npairs}", flush=True)
    return big, np.tile(f, (8, 1)), 8.0 * ec, 8.0 * elj, 8 * npairs


def rms_rel(f, ref):
    return np.sqrt(np.mean(np.sum((f - ref) ** 2, 1))) / np.sqrt(np.mean(np.sum(ref ** 2, 1)))


@pytest.mark.parametrize("precision", [pmx.FP64, pmx.MIXED])
def test_large_box_matches_replicated_oracle(ctx, boxes, precision):
    big, f_ref, ec_ref, elj_ref, np_ref = boxes
    assert big.n == 999000
    spec = pmx.ForceSpec(CELLSIZE, lj=LJ, pme=PME_BIG)
    sy = pmx.System(ctx, big, spec, precision)
    try:
        f, ec, elj = sy.evaluate(big.pos, flags=pmx.EVAL_PAIRSTATS)
        n, _ = sy.pair_stats()
        cells, dims = sy.cells() if precision == pmx.FP64 else (None, None)
    finally:
        sy.close()
    err = rms_rel(f, f_ref)
    print(f"large box prec={precision} pairs {n} vs {np_ref} Ecoul {ec:.12g} vs {ec_ref:.12g} "
          f"Elj {elj:.12g} vs {elj_ref:.12g} rms_rel {err:.3e} maxabs {np.abs(f - f_ref).max():.3e}", flush=True)
    assert n == np_ref
    if precision == pmx.FP64:
        assert np.abs(f - f_ref).max() <= FP64_RTOL * np.abs(f_ref).max()
        assert abs(ec - ec_ref) <= FP64_RTOL * big.n
        assert abs(elj - elj_ref) <= FP64_RTOL * max(1.0, abs(elj_ref))
        ocells, odims = O.cells(big, CELLSIZE)
        assert dims == odims
        assert np.array_equal(cells, ocells)
    else:
        assert err <= MIXED_RMS
        assert abs(ec - ec_ref) <= MIXED_E_PER_ATOM * big.n
        assert abs(elj - elj_ref) <= MIXED_E_PER_ATOM * big.n
