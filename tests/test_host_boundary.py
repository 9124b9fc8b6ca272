"""The C++ ProtoMol boundary (ForceFactory -> GpuNonbondedSystemForce ->
pmx C ABI), driven the way the reference's regression tests drive ProtoMol:
force lines in .conf grammar, a system, an output file compared with the
reference's numbers."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from golden import fixtures as FX
from protomol3_amd import build, sysfile

EXE = build.HOST_EXE


def force_lines(sp):
    out = []
    sw = {1: "Cutoff", 2: "C1", 3: "C2"}
    if sp["lj"] is not None:
        k, on, rc = sp["lj"]
        out.append(f"LennardJones -algorithm NonbondedCutoff -switchingFunction {sw[k]}"
                   + (f" -switchon {on!r}" if k == 3 else "") + f" -cutoff {rc!r}")
    if sp["coul"] is not None:
        k, on, rc = sp["coul"]
        out.append(f"Coulomb -algorithm NonbondedCutoff -switchingFunction {sw[k]}"
                   + (f" -switchon {on!r}" if k == 3 else "") + f" -cutoff {rc!r}")
    if sp["pme"] is not None:
        p = sp["pme"]
        terms = "".join(t for b, t in ((1, " -real"), (2, " -reciprocal"), (4, " -correction")) if p["terms"] & b)
        g = p["grid"]
        out.append(f"Coulomb -algorithm PMEwald{terms} -interpolation BSpline -switchingFunction "
                   f"{sw[p['switch']]} -gridsize {g[0]} {g[1]} {g[2]} -cutoff {p['rc']!r} -order {p['order']} "
                   f"-accuracy {p['accuracy']!r} -alpha {p['alpha']!r}")
    return out


def run_exe(s, cellsize, lines, mixed=False, extra=()):
    with tempfile.TemporaryDirectory() as td:
        sf, of = os.path.join(td, "sys"), os.path.join(td, "out")
        sysfile.write_sysfile(sf, s)
        args = [EXE, sf, of, "cellsize", repr(cellsize)] + (["--mixed"] if mixed else []) + list(extra)
        for ln in lines:
            args += ["--force", ln]
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        if r.returncode != 0:
            return r.returncode, r.stderr, None
        return 0, r.stderr, sysfile.read_forces(of)


@pytest.mark.parametrize("line,msg", [
    ("LennardJones -algorithm NonbondedFull -cutoff 8", "not served"),
    ("Coulomb -algorithm PMEwald -interpolation Hermite -gridsize 8 8 8 -cutoff 5", "BSpline"),
    ("LennardJones -algorithm NonbondedCutoff -switchingFunction C7 -cutoff 8", "switching function"),
    ("Gravity -algorithm NonbondedCutoff -cutoff 8", "unknown force"),
])
def test_force_line_grammar_errors_are_reported(line, msg):
    fx = FX.load("water423_ljc2_coulc1")
    rc, err, _ = run_exe(FX.system(fx), 6.5, [line])
    assert rc == 3 and msg in err


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["water423_ljc2_coulc1", "water1536_ljc2_pme", "water1536_pme_c1",
                                  "polymer3200_pme_auto", "polymer26000_pme_auto", "polymer3200_mod14_pme",
                                  "polymer3200_mod14_ljc2_coulc2", "polymer3201_charged_pme"])
def test_conf_force_lines_on_gpu_match_reference_fixture(name):
    fx = FX.load(name)
    s = FX.system(fx)
    sp = FX.spec(fx)
    rc, err, out = run_exe(s, sp["cellsize"], force_lines(sp))
    assert rc == 0, err
    ec, elj, f = out
    idx = fx["force_index"]
    fref = fx["forces"]
    assert np.abs(f[idx] - fref).max() <= 1e-9 * max(1.0, np.abs(fref).max())
    assert abs(ec - float(fx["e_coulomb"])) <= 1e-9 * max(1.0, abs(float(fx["e_coulomb"])))
    assert abs(elj - float(fx["e_lj"])) <= 1e-9 * max(1.0, abs(float(fx["e_lj"])))


@pytest.mark.gpu
@pytest.mark.parametrize("ranks", [1, 3])
def test_parallel_evaluate_gives_one_rank_the_whole_force(ranks):
    """Under ProtoMol's static work distribution the GPU force is one work
    package: exactly one of the ranks evaluates it (the others add nothing),
    and the ranks' summed forces and energies are the sequential ones
    (NonbondedCutoffSystemForce.h:57-77, SystemForce.h:107-112)."""
    fx = FX.load("polymer3200_pme_auto")
    s = FX.system(fx)
    sp = FX.spec(fx)
    rc, err0, seq = run_exe(s, sp["cellsize"], force_lines(sp), mixed=True)
    assert rc == 0, err0
    rc, err, par = run_exe(s, sp["cellsize"], force_lines(sp), mixed=True, extra=["--ranks", str(ranks)])
    assert rc == 0, err
    print(err, flush=True)
    assert err.count("evaluated") == 1 and err.count("nothing") == ranks - 1
    assert err.count("blocks 1,") == ranks
    assert np.array_equal(par[2], seq[2]) and par[0] == seq[0] and par[1] == seq[1]
