"""The CHARMM topology builder (protomol3_amd/charmm.py) against the
reference's own (BuildTopology.cpp, compiled from the reference sources
into oracle/_ref/pm_ref_topo by tests/golden/make_fixtures.py) on the
reference's test molecules: the ww_domain protein (dihedrals, impropers,
five-membered rings, 32 LJ types) and the alanine dipeptide in a water
droplet, for several `exclude` keywords.  Integer data must match exactly;
scaled charges and the LJ table to rounding."""
import os

import numpy as np
import pytest

from golden import fixtures as FX
from protomol3_amd import charmm

PAR27 = os.path.join(FX.W423, "par_all27_prot_lipid.inp")


def reference_exclusion_rules(entries):
    """ExclusionTable::add (ExclusionTable.cpp:103-120) applied to a list of
    (a1, a2, class): the correction table (each pair once, first class, in
    insertion order) and the check() class of each pair (the last added)."""
    table, seen, check = [], set(), {}
    for a, b, c in entries:
        a, b = min(a, b), max(a, b)
        if c != 0 and a != b and (a, b) not in seen:
            table.append((a, b, c))
            seen.add((a, b))
        check[(a, b)] = c
    return table, check


@pytest.mark.parametrize("name", FX.topo_names())
def test_charmm_topology_matches_reference_builder(name):
    fx = FX.load(name)
    stem, exclude = str(fx["stem"]), str(fx["exclude"])
    d = os.path.join(FX.HERE, "charmm")
    s = charmm.build_system(os.path.join(d, stem + ".psf"), os.path.join(d, stem + ".pdb"), PAR27,
                            box=(100.0, 100.0, 100.0), exclude=exclude)
    assert np.array_equal(s.atype, fx["type"])
    assert np.array_equal(s.molecule, fx["molecule"])
    assert np.abs(s.charge - fx["scaled_charge"]).max() <= 1e-15 * np.abs(fx["scaled_charge"]).max()
    assert s.lj.shape == fx["lj"].shape
    assert np.allclose(s.lj, fx["lj"], rtol=1e-14, atol=0.0)
    table, check = reference_exclusion_rules(s.exclusions.tolist())
    ref = fx["table"]
    assert [tuple(t) for t in table] == [tuple(r[:3]) for r in ref.tolist()]
    assert all(check[(int(r[0]), int(r[1]))] == int(r[3]) for r in ref)
    nb, na, nd, ni = fx["counts"].tolist()
    assert (len(s.bonds), len(s.angles), len(s.dihedrals), len(s.impropers)) == (nb, na, nd, ni)


def test_ring_pairs_take_both_classes():
    """ww_domain's five-membered rings hold pairs that are 1-3 one way round
    and 1-4 the other: under scaled1-4 the reference's table keeps FULL
    (added first, from the angle) while check() answers MODIFIED (added last,
    from the dihedral).  The fixture shows such pairs exist, so the engine's
    handling of them is exercised by the force fixtures of this molecule."""
    ref = FX.load("topo_ww_domain_scaled14")["table"]
    split = ref[(ref[:, 2] == 1) & (ref[:, 3] == 2)]
    assert len(split) > 0
