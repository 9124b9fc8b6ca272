"""GPU parity against the reference fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py from the reference compiled from its own
sources): every fixture, both numeric forms, through the C ABI.

Checked per fixture: forces at the fixture's atoms (all atoms, or a stride
for the 26,000-atom workload), Coulomb and Lennard-Jones energies, the PME
alpha, the cell of every atom, and the count and set hash of the in-cutoff
non-excluded pairs (bit-exact in both forms).  The fixtures include the
26,000-atom polymer/water workload with PME 64^3, scaled 1-4 pairs
(MODIFIED class, coulombScalingFactor 0.5, separate A14/B14), a system with
net charge (charged-system energy term) and LJ and Coulomb cutoffs that
differ."""
import numpy as np
import pytest

from golden import fixtures as FX
from protomol3_amd import pmx

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-9
MIXED_RMS = 1e-5     # RMS relative force error of the mixed form
MIXED_ERTOL = 3e-5   # relative energy error of the mixed form (FP32 pair energies summed)


@pytest.fixture(scope="module")
def ctx():
    c = pmx.Context(0)
    yield c
    c.close()


def _rms_rel(f, ref):
    return np.sqrt(np.mean(np.sum((f - ref) ** 2, 1)) / np.mean(np.sum(ref ** 2, 1)))


@pytest.mark.parametrize("name", FX.names())
@pytest.mark.parametrize("precision", [pmx.FP64, pmx.MIXED])
def test_fixture_on_gpu(ctx, name, precision):
    fx = FX.load(name)
    s = FX.system(fx)
    sp = FX.spec(fx)
    gspec = pmx.ForceSpec(sp["cellsize"], lj=sp["lj"], coul=sp["coul"], pme=sp["pme"])
    print(f"fixture {name} n={s.n} precision={'mixed' if precision else 'fp64'}", flush=True)
    sy = pmx.System(ctx, s, gspec, precision)
    try:
        f, ec, elj = sy.evaluate(s.pos, flags=pmx.EVAL_PAIRSTATS)
        npairs, phash = sy.pair_stats()
        cells, dims = sy.cells()
        alpha = sy.alpha() if sp["pme"] is not None else 0.0
    finally:
        sy.close()
    idx = fx["force_index"]
    fref = fx["forces"]
    ec_ref, elj_ref = float(fx["e_coulomb"]), float(fx["e_lj"])
    rms = _rms_rel(f[idx], fref)
    print(f"  Ecoul {ec:.12g} vs {ec_ref:.12g}  Elj {elj:.12g} vs {elj_ref:.12g}  rms_rel {rms:.3e}  "
          f"pairs {npairs} vs {int(fx['npairs'])}", flush=True)
    # membership and binning: bit-exact in both forms
    assert npairs == int(fx["npairs"])
    assert phash == int(fx["pair_hash"])
    assert dims == tuple(int(v) for v in fx["cell_dims"])
    assert np.array_equal(cells, fx["cells"])
    if sp["pme"] is not None:
        assert abs(alpha - float(fx["alpha"])) <= 1e-14 * float(fx["alpha"])
    if precision == pmx.FP64:
        scale = max(1.0, np.abs(fref).max())
        assert np.abs(f[idx] - fref).max() <= FP64_RTOL * scale
        assert abs(ec - ec_ref) <= FP64_RTOL * max(1.0, abs(ec_ref))
        assert abs(elj - elj_ref) <= FP64_RTOL * max(1.0, abs(elj_ref))
    else:
        assert rms <= MIXED_RMS
        assert abs(ec - ec_ref) <= MIXED_ERTOL * max(1.0, abs(ec_ref))
        assert abs(elj - elj_ref) <= MIXED_ERTOL * max(1.0, abs(elj_ref))
