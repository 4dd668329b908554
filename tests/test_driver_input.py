"""The Fortran driver's input checks, which run before any GPU call: unrecognised keywords are
fatal and print the offending lines, as in the reference (parameters.F90:999-1030); reference
keywords of features the driver does not run are refused when switched on."""
import os
import re
import subprocess

import pytest

from tests.conftest import GOLDEN, ROOT

EXE = os.path.join(ROOT, "espins_amd", "driver", "mc_gpu.x")
pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built (no amdflang)")


def _run(tmp_path, text):
    (tmp_path / "x.mcin").write_text(text)
    return subprocess.run([EXE, "x"], cwd=tmp_path, capture_output=True, text=True, timeout=60)


def test_unrecognised_keyword_is_fatal(tmp_path):
    text = open(os.path.join(GOLDEN, "examples", "ex01_cubic.mcin")).read()
    r = _run(tmp_path, text + "\n steps_warmpu = 10\n")
    assert r.returncode != 0
    assert "contained unrecognised keywords" in r.stdout
    assert "steps_warmpu = 10" in r.stdout
    assert "Unrecognised keyword(s) in input file" in r.stdout


def test_reference_setup_keywords_are_accepted_and_checked(tmp_path):
    text = open(os.path.join(GOLDEN, "examples", "ex01_cubic.mcin")).read()
    r = _run(tmp_path, text + "\n length_unit = parsec\n")
    assert r.returncode != 0 and "length_unit not recognised" in r.stdout


def _obs_input(**kw):
    text = open(os.path.join(GOLDEN, "examples", "ex01_cubic.mcin")).read()
    return text + "\n" + "\n".join(f" {k} = {v}" for k, v in kw.items()) + "\n"


def test_observable_keyword_checks(tmp_path):
    """The reference's checks of the observable keywords (parameters.F90:440-592)."""
    r = _run(tmp_path, _obs_input(order_parameter=".true."))
    assert r.returncode != 0 and "No order_parameter_axes_frac or order_parameter_axes_cart" in r.stdout
    r = _run(tmp_path, _obs_input(binning_error=".true.", binning_level="1-40"))
    assert r.returncode != 0 and "binning_level contain non-valid numbers" in r.stdout
    r = _run(tmp_path, _obs_input(sfactor=".true.", sfactor_steps_measure=0))
    assert r.returncode != 0 and "sfactor_steps_measure must be greater than zero" in r.stdout
    r = _run(tmp_path, _obs_input(sfactor=".true.", sfactor_2dqmesh="0 4"))
    assert r.returncode != 0 and "sfactor_2dqmesh must be greater than zero" in r.stdout
    r = _run(tmp_path, _obs_input(sfactor_q1="1 0 0", sfactor_q2="2 0 0"))
    assert r.returncode != 0 and "not linearly independent" in r.stdout
    lines = _obs_input(spin_glass=".true.").splitlines()
    k = next(i for i, l in enumerate(lines) if "jij=" in l.replace(" ", "").lower())
    lines[k] = lines[k].split("!")[0].rstrip() + ":sig=-0.001"
    r = _run(tmp_path, "\n".join(lines) + "\n")
    assert r.returncode != 0 and "sig in jij_parameters block must be positive" in r.stdout


def _echo_section(lines):
    i = next(k for k, l in enumerate(lines) if l.strip() == "SYSTEM") - 2
    j = next(k for k, l in enumerate(lines) if l.startswith(" Time to read parameters"))
    return lines[i:j]


@pytest.mark.parametrize("name", ["ex01_cubic", "ex02_cubic", "ex03_fef3", "ex04_bmno_pt", "ex05_snwo"])
def test_parameter_echo_matches_published_output(tmp_path, name):
    """The .mcout parameter echo (param_write, parameters.F90:1074-1444) of every published
    example, line for line: lattice, reciprocal vectors, atoms, Monte Carlo and Hamiltonian
    settings, the J/B/D tables, DM vectors, order-parameter axes and spin-correlation shells
    (fixtures made by tests/golden/make_echo_golden.py from the examples' .mcout)."""
    seed = name.split("_", 1)[1]
    (tmp_path / f"{seed}.mcin").write_text(open(os.path.join(GOLDEN, "examples", f"{name}.mcin")).read())
    r = subprocess.run([EXE, "-dryrun", seed], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    mine = _echo_section((tmp_path / f"{seed}.mcout").read_text().splitlines())
    want = open(os.path.join(GOLDEN, "examples", f"{name}_echo.txt")).read().splitlines()
    assert len(mine) == len(want)
    for a, b in zip(mine, want):
        assert a == b


def test_parameter_echo_units_and_cartesian_rows(tmp_path):
    """parameter_unit = ryd prints couplings times 2/hart with (Ryd) headers; length_unit = bohr
    scales lengths by 1/bohr; coordinate = cart prints the rows' sites in Cartesian
    coordinates, the first atom scaled by 1/bohr and the second not (parameters.F90:1206-1213)."""
    text = open(os.path.join(GOLDEN, "examples", "ex03_fef3.mcin")).read()
    text += "\n parameter_unit = ryd\n length_unit = bohr\n coordinate = cart\n"
    (tmp_path / "u.mcin").write_text(text)
    r = subprocess.run([EXE, "-dryrun", "u"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    out = (tmp_path / "u.mcout").read_text()
    bohr, hart = 0.5291772108, 27.21138602
    assert "Lattice Vectors (Bohr)" in out and "(Bohr^3)" in out and "J(Ryd)" in out and "D(Ryd)" in out
    a1 = [float(x) for x in re.search(r"a_1\s+(\S+)\s+(\S+)\s+(\S+)", out).groups()]
    assert a1 == pytest.approx([5.1625 / bohr, 0.0, 5.1625 / bohr], abs=1e-6)
    assert "In Cartesian Coordinate(Bohr):" in out
    rows = re.findall(r"^ \|\s*(\S+)\s+(\S+)\s+(\S+) \|\s*(\S+)\s+(\S+)\s+(\S+) \|\s*(\S+) \|$", out, re.M)
    j = [r_ for r_ in rows if float(r_[6]) < 0]           # the exchange rows (J = -0.0327 eV)
    assert len(j) == 24
    assert float(j[0][6]) == pytest.approx(-0.0327 * 2 / hart, abs=1e-8)
    # second row: atom 1 at the origin, atom 2 at frac (0, 0.5, 0) -> Cartesian (0, 2.58125, 2.58125) in Ang
    assert [float(x) for x in j[1][3:6]] == pytest.approx([0.0, 2.58125, 2.58125], abs=1e-6)
    # seventh row: atom 1 at frac (0, 0, 0.5) -> (2.58125, 2.58125, 0) Ang, printed in Bohr
    assert [float(x) for x in j[6][0:3]] == pytest.approx([2.58125 / bohr, 2.58125 / bohr, 0.0], abs=1e-5)


def test_parameter_echo_spin_glass_sigma_table(tmp_path):
    """With spin_glass the echo prints the Sigma table (one row per Jij row, sig= in eV) in
    place of the shell table (parameters.F90:1217-1247)."""
    text = open(os.path.join(GOLDEN, "examples", "ex01_cubic.mcin")).read()
    rows = []
    for l in text.splitlines():
        if "jij=" in l.replace(" ", "").lower() and l.strip().lower().startswith(("f1", "c1")):
            l = l.split("!")[0].rstrip() + ":sig=0.003"
        rows.append(l)
    text = "\n".join(rows) + "\n spin_glass = .true.\n spin_glass_seed = 11\n"
    (tmp_path / "sg.mcin").write_text(text)
    r = subprocess.run([EXE, "-dryrun", "sg"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    out = (tmp_path / "sg.mcout").read_text()
    assert re.search(r"Using random exchange parameters\s*:\s*T", out)
    sigma = out[out.index("-- Sigma --"):]
    assert "Sigma(eV)" in sigma.splitlines()[3]
    vals = re.findall(r"\|\s*0\.00300000 \|$", sigma, re.M)
    nj = len(re.findall(r"\|\s*-?\d+\.\d{8} \|$", out[out.index("-- Exchange --"):out.index("-- Sigma --")], re.M))
    assert nj > 0 and len(vals) == nj
