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
