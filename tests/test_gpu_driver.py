"""The Fortran driver (espins_amd/driver/mc_gpu.x) end to end on the reference's example inputs:
same .mcin in, same _mc.dat / .mcout / _sconfig.dat layout out."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "espins_amd", "driver", "mc_gpu.x")


def _prepare(tmp_path, src, seed, edits):
    text = open(os.path.join(GOLDEN, "examples", src)).read()
    for key, val in edits.items():
        pat = re.compile(rf"^\s*!?!?\s*{key}\s*[=:].*$", re.I | re.M)
        if pat.search(text):
            text = pat.sub(f" {key} = {val}", text, count=1)
        else:
            text = f" {key} = {val}\n" + text
    p = tmp_path / f"{seed}.mcin"
    p.write_text(text)
    return p


def _run(tmp_path, seed):
    assert os.path.exists(EXE), "driver not built"
    r = subprocess.run([EXE, seed], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    return (tmp_path / f"{seed}.mcout").read_text()


def test_example01_plain_mc(tmp_path):
    _prepare(tmp_path, "ex01_cubic.mcin", "cubic",
             {"steps_warmup": 4000, "steps_mc": 20000, "seeds": 12345})
    out = _run(tmp_path, "cubic")
    # the published initial (ferro) energy, printed in the same E15.7 format
    # (examples/example01/reference/cubic.mcout:110)
    assert out.count("-0.6684205E+04  <-- ENERGY") == 20
    adds = [float(x) for x in re.findall(r"adding differences\s*=\s*(\S+)", out)]
    direct = [float(x) for x in re.findall(r"direct calculations\s*=\s*(\S+)", out)]
    assert len(adds) == 20 and np.allclose(adds, direct, rtol=1e-6)
    mine = np.loadtxt(tmp_path / "cubic_mc.dat", comments="#")
    ref = np.loadtxt(os.path.join(GOLDEN, "examples", "ex01_cubic_mc.dat"), comments="#")
    assert mine.shape == ref.shape == (20, 7)
    assert np.allclose(mine[:, 0], ref[:, 0], atol=1e-4)
    # energy per site within 2 % of the published run at every temperature
    assert np.allclose(mine[:, 2], ref[:, 2], rtol=0.02)
    # _sconfig.dat: header, N, T, then unit spins (moment 1)
    lines = (tmp_path / "cubic_sconfig.dat").read_text().splitlines()
    assert lines[1].split() == ["4", "4", "4"] and lines[2].split() == ["64", "20"]
    s = np.array([[float(v) for v in l.split()] for l in lines[4:68]])
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=2e-6)


def test_sconfig_restart(tmp_path):
    _prepare(tmp_path, "ex01_cubic.mcin", "cubic",
             {"steps_warmup": 200, "steps_mc": 400, "seeds": 7})
    _run(tmp_path, "cubic")
    final = [float(x) for x in re.findall(r"direct calculations\s*=\s*(\S+)",
                                          (tmp_path / "cubic.mcout").read_text())]
    _prepare(tmp_path, "ex01_cubic.mcin", "cubic",
             {"steps_warmup": 0, "steps_mc": 10, "seeds": 8, "initial_sconfig": "file"})
    out = _run(tmp_path, "cubic")
    init = [float(x) for x in re.findall(r"Initial Energy\s*=\s*(\S+)", out)]
    # restart energies agree with the energies the first run ended with (3F12.6 spins)
    assert np.allclose(init, final, rtol=2e-4)


def test_example02_parallel_tempering(tmp_path):
    _prepare(tmp_path, "ex02_cubic.mcin", "cubic",
             {"steps_warmup": 400, "steps_mc": 1000, "supercell_size": "6 6 6", "seeds": 3,
              "energy_write": ".false."})
    out = _run(tmp_path, "cubic")
    mine = np.loadtxt(tmp_path / "cubic_mc.dat", comments="#")
    assert mine.shape == (32, 7)
    assert np.all(np.diff(mine[:, 0]) > 0)
    m = re.findall(r"Final configuration =\s*([\d\s]+)", out)
    perm = [int(x) for x in m[-1].split()]
    assert sorted(perm) == list(range(1, 33))
    assert "swaps" in out


def test_example01_learned_proposal(tmp_path):
    """mcarlo_mode = learned: warm-up with 'random', training on the device, measured phase with
    the learned proposal; observables as in the published plain run."""
    _prepare(tmp_path, "ex01_cubic.mcin", "cubic",
             {"steps_warmup": 2000, "steps_mc": 10000, "seeds": 99, "mcarlo_mode": "learned",
              "learned_records": 65536, "learned_epochs": 2})
    out = _run(tmp_path, "cubic")
    assert "Learned proposal: training on" in out and out.count("mean -log q") == 2
    adds = [float(x) for x in re.findall(r"adding differences\s*=\s*(\S+)", out)]
    direct = [float(x) for x in re.findall(r"direct calculations\s*=\s*(\S+)", out)]
    assert len(adds) == 20 and np.allclose(adds, direct, rtol=1e-6)
    mine = np.loadtxt(tmp_path / "cubic_mc.dat", comments="#")
    ref = np.loadtxt(os.path.join(GOLDEN, "examples", "ex01_cubic_mc.dat"), comments="#")
    assert np.allclose(mine[:, 2], ref[:, 2], rtol=0.02)


def test_energy_series_files(tmp_path):
    """energy_write = .true.: one seedname_energy-TNNN.dat per temperature with the running
    energy per site of every measured step; their means are the <E>/N column of _mc.dat."""
    _prepare(tmp_path, "ex02_cubic.mcin", "cubic",
             {"steps_warmup": 200, "steps_mc": 500, "supercell_size": "4 4 4", "seeds": 5,
              "energy_write": ".true."})
    _run(tmp_path, "cubic")
    mine = np.loadtxt(tmp_path / "cubic_mc.dat", comments="#")
    for t in range(mine.shape[0]):
        lines = (tmp_path / f"cubic_energy-T{t + 1:03d}.dat").read_text().splitlines()
        assert float(lines[1].split("=")[1]) == pytest.approx(mine[t, 0], abs=1e-4)
        nsteps = int(lines[2])
        data = np.array([[float(v) for v in l.split()] for l in lines[3:]])
        assert data.shape == (nsteps, 2)
        assert np.array_equal(data[:, 0], np.arange(1, nsteps + 1))
        assert data[:, 1].mean() == pytest.approx(mine[t, 2], rel=1e-6, abs=1e-7)


def _rand_config(m, seed):
    """The reference's 'rand' start (mcarlo.F90:275-286): two draws of its MT19937 stream per
    site, S = m (sin(pi r1) cos(2 pi r2), sin(pi r1) sin(2 pi r2), cos(pi r1)); the oracle's
    MT19937 is pinned to the reference's own mtprng.f90."""
    from oracle import oracle as O
    import math
    st = O.mt_state(seed)
    s = np.zeros((m.N, 3))
    for i in range(m.N):
        r1 = O.lib.orc_mt_real2(st)
        r2 = O.lib.orc_mt_real2(st)
        s[i] = m.mag[i] * np.array([math.sin(r1 * math.pi) * math.cos(r2 * 2 * math.pi),
                                    math.sin(r1 * math.pi) * math.sin(r2 * 2 * math.pi),
                                    math.cos(r1 * math.pi)])
    return s


def test_rand_start_is_the_reference_stream(tmp_path):
    """initial_sconfig = rand with seeds = s: every temperature starts from the configuration
    the reference draws from its MT19937 stream seeded with s -- the printed initial energies
    equal the oracle energy of that configuration -- and a second run reproduces the first."""
    from espins_amd import mcin
    from oracle import oracle as O
    edits = {"steps_warmup": 100, "steps_mc": 200, "seeds": 4357, "initial_sconfig": "rand"}
    _prepare(tmp_path, "ex01_cubic.mcin", "cubic", edits)
    out = _run(tmp_path, "cubic")
    init = [float(x) for x in re.findall(r"Initial Energy\s*=\s*(\S+)", out)]
    m = mcin.to_model(mcin.read(str(tmp_path / "cubic.mcin")), builder=O.build_table_fn)
    e_ref = O.Oracle(m).energy(_rand_config(m, 4357))
    assert len(init) == 20
    assert np.allclose(init, e_ref, rtol=1e-6, atol=1e-6), (init[:2], e_ref)
    first = (tmp_path / "cubic_mc.dat").read_text().splitlines()[1:]
    _run(tmp_path, "cubic")
    assert (tmp_path / "cubic_mc.dat").read_text().splitlines()[1:] == first


def test_pt_warmup_outputs_and_map_prints(tmp_path):
    """Parallel tempering: the WARMUP1 / WARMUP2 tables, _sconfig_warmup.dat after warmup1, and
    the replica map printed at step 1, every pt_print_swap steps and at the last step of each
    phase (pt.F90:557-713, 773-781)."""
    _prepare(tmp_path, "ex02_cubic.mcin", "cubic",
             {"steps_warmup": 400, "steps_mc": 600, "supercell_size": "4 4 4", "seeds": 11,
              "energy_write": ".false.", "pt_print_swap": 150})
    out = _run(tmp_path, "cubic")
    for tag in ("<-- WARMUP1", "<-- ENERGY1", "<-- WARMUP2", "<-- ENERGY2", "<-- MCARLO"):
        assert out.count(tag) == 32, tag
    assert out.count("<-- ENERGY3") == 3 * 32          # initial, by adding, direct
    # warm-up halves: 200 plain + 200 with exchanges (pt.F90:529-537); main 600 steps
    prints = [int(l.split("|")[0]) for l in out.splitlines() if re.match(r"^\s*\d+\s+\|", l)]
    assert prints == [1, 150, 200] + [1, 150, 300, 450, 600]
    lines = (tmp_path / "cubic_sconfig_warmup.dat").read_text().splitlines()
    assert lines[1].split() == ["64", "32"]
    s = np.array([[float(v) for v in l.split()] for l in lines[3:67]])
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=2e-6)
    # running and direct energies agree at the end of both exchange phases
    for tag in ("ENERGY2",):
        rows = [l.split() for l in out.splitlines() if l.rstrip().endswith("<-- " + tag)]
        e = np.array([[float(r[2]), float(r[3])] for r in rows])
        assert np.allclose(e[:, 0], e[:, 1], rtol=1e-6)
