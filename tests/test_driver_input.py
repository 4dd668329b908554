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
