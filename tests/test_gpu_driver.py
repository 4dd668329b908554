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
    # closing lines as mc_prog.F90:199-208 prints them
    assert re.search(r"^ Total Execution Time\s+\d+\.\d{3} \(sec\)$", out, re.M)
    assert re.search(r"^ \|mcarlo_main\s+:\s+1\s+\d+\.\d{3}\s+\|$", out, re.M)
    assert out.rstrip().endswith("All done: mcarlo exiting")
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
    assert re.search(r"^ \|pt_main\s+:\s+1\s+\d+\.\d{3}\s+\|$", out, re.M)
    assert out.rstrip().endswith("All done: ptmc exiting")
    # per temperature slot: running energy = direct energy, and the direct energy is the
    # oracle's energy of the configuration written for that slot (3F12.6 spins, E15.7 print)
    adds = [float(x) for x in re.findall(r"adding differences\s*=\s*(\S+)\s+<-- ENERGY3", out)]
    direct = [float(x) for x in re.findall(r"direct calculations\s*=\s*(\S+)\s+<-- ENERGY3", out)]
    assert len(adds) == len(direct) == 32
    assert np.allclose(adds, direct, rtol=1e-6)
    from espins_amd import mcin
    from oracle import oracle as O
    model = mcin.to_model(mcin.read(str(tmp_path / "cubic.mcin")), builder=O.build_table_fn)
    orc = O.Oracle(model)
    lines = (tmp_path / "cubic_sconfig.dat").read_text().splitlines()
    N, R = (int(v) for v in lines[2].split())
    assert (N, R) == (model.N, 32)
    p = 3
    for t in range(R):
        assert float(lines[p]) == pytest.approx(mine[t, 0], abs=1e-4)
        spins = np.array([[float(v) for v in lines[p + 1 + i].split()] for i in range(N)])
        p += N + 1
        assert orc.energy(spins) == pytest.approx(direct[t], rel=1e-4)


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


def _two_kind_bcc(tmp_path, seed, extra):
    """BCC-type cell with two kinds (A at the corner, B at the centre), cell edge 2 A, J1 between
    kinds (8 bonds, sh = 1), J2 within a kind along the axes (6 bonds, sh = 2)."""
    import itertools
    rows = []
    for a, fa in (("a", (0.0, 0.0, 0.0)), ("b", (0.5, 0.5, 0.5))):
        other = (0.5, 0.5, 0.5) if a == "a" else (0.0, 0.0, 0.0)
        for sx, sy, sz in itertools.product((-0.5, 0.5), repeat=3):
            f2 = (fa[0] + sx, fa[1] + sy, fa[2] + sz)
            rows.append(f"f1={fa[0]},{fa[1]},{fa[2]}:f2={f2[0]},{f2[1]},{f2[2]}:jij=-0.002:sh=1")
        for ax in range(3):
            for sg in (-1.0, 1.0):
                f2 = list(fa)
                f2[ax] += sg
                rows.append(f"f1={fa[0]},{fa[1]},{fa[2]}:f2={f2[0]},{f2[1]},{f2[2]}:jij=0.0005:sh=2")
    text = f"""
 Begin Unit_Cell_Cart
   2.0 0.0 0.0
   0.0 2.0 0.0
   0.0 0.0 2.0
 End Unit_Cell_Cart
 Begin Atoms_Frac
  Fe 0.0 0.0 0.0 1.0
  Co 0.5 0.5 0.5 1.0
 End Atoms_Frac
 tems_mode = lin
 tem_start = 10
 tem_end = 40
 tems_num = 2
 supercell_size = 4 4 4
 initial_sconfig = rand
 seeds = {seed}
 Begin Jij_parameters
{chr(10).join(rows)}
 End Jij_parameters
{extra}
"""
    (tmp_path / "bcc2.mcin").write_text(text)


def _sconfig_spins(path, N, R):
    lines = path.read_text().splitlines()
    out, k = [], 3
    for t in range(R):
        k += 1                                     # temperature line
        out.append(np.array([[float(v) for v in l.split()] for l in lines[k:k + N]]))
        k += N
    return out


def test_observable_outputs_one_measurement(tmp_path):
    """One measured step (steps_mc = 1): every observable file holds the observable of the final
    configuration, which _sconfig.dat prints (3F12.6), compared with the oracle on those spins."""
    from espins_amd import lattice as Lt
    from oracle import obs_oracle as B
    from oracle import oracle as O
    extra = """
 steps_warmup = 30
 steps_mc = 1
 steps_measure = 1
 order_parameter = .true.
 Begin order_parameter_axes_cart
  Fe 0 0 1
  Co 0 0 -1
 End order_parameter_axes_cart
 staggered_m = .true.
 staggered_m_coeff = 1 -1
 spin_correlation = .true.
 sfactor = .true.
 sfactor_2dqmesh = 4 4
 sfactor_steps_measure = 1
 sfactor_polar = 0 0 1
"""
    _two_kind_bcc(tmp_path, 17, extra)
    _run(tmp_path, "bcc2")
    N, R = 128, 2
    spins = _sconfig_spins(tmp_path / "bcc2_sconfig.dat", N, R)
    axes = [(0, 0, 1), (0, 0, -1)]
    coeff = [1.0, -1.0]
    op = np.loadtxt(tmp_path / "bcc2_op.dat", comments="#")
    sm = np.loadtxt(tmp_path / "bcc2_staggered.dat", comments="#")
    # 3F12.6 spins: observables agree to ~1e-5
    for t in range(R):
        assert op[t, 1] == pytest.approx(abs(B.orderparam(spins[t], 2, axes)), abs=2e-5)
        assert sm[t, 1] == pytest.approx(B.staggered(spins[t], 2, coeff), abs=2e-5)
    # P(m): one measurement, every cell counted once; sum P dm = 1 per column
    pm_lines = (tmp_path / "bcc2_pm.dat").read_text().splitlines()
    blocks = [i for i, l in enumerate(pm_lines) if l.strip().startswith("#T=")]
    assert len(blocks) == R
    for t, b in enumerate(blocks):
        rows = np.array([[float(v) for v in l.split()] for l in pm_lines[b + 1:b + 402]])
        assert rows.shape == (401, 4)
        assert np.allclose(rows[:, 1:].sum(axis=0) * 0.01, 1.0, rtol=1e-7)
        h = B.pm_counts(spins[t], 2, [1.0, 1.0], axes=axes, coeff=coeff)
        got = np.rint(rows[:, 1:].T * 0.01 * (N // 2)).astype(int)
        # bins agree except where a 3F12.6-rounded value moved across a bin edge
        assert np.abs(got - h).sum() <= 4
    # spin correlations: shells 1 (A-B) and 2 (A-A, B-B), kinds 1, 2
    sc_rows = [l.split() for l in (tmp_path / "bcc2_spincorr.dat").read_text().splitlines()
               if l.strip() and not l.strip().startswith("#")]
    assert len(sc_rows) == R * 2 * 2 * 2
    assert {r[2] for r in sc_rows} == {"Fe", "Co"}
    sc = np.array([[float(r[k]) for k in (0, 1, 3, 5, 6, 7, 8)] for r in sc_rows])
    sc = np.insert(np.insert(sc, 2, 0.0, axis=1), 4, 0.0, axis=1)   # columns as in the file
    kinds = np.arange(N) % 2
    c1, n1 = Lt.shell_tables("bcc", 4, 2, builder=O.build_table_fn)
    for t in range(R):
        tmp, tab = B.spincorr(spins[t], kinds, c1, n1, np.ones(N), 2)
        rows = sc[t * 8:(t + 1) * 8]
        for row in rows:
            l, a, b = int(row[1]) - 1, int(row[3]) - 1, int(row[5]) - 1
            assert row[6] == pytest.approx(tmp[l, a, b] / N, abs=5e-5)
            assert row[7] == pytest.approx(tab[l, a, b] / N, abs=5e-5)
            assert row[8] == pytest.approx(abs(tmp[l, a, b]) / N, abs=5e-5)
    # structure factor: 25 q-points per temperature, 5 columns with the polarisation
    pos = Lt.site_positions("bcc", 4) * 2.0
    tp = 2 * np.pi / 2.0
    qs = [(lx / 4 * tp, 0.0, ly / 4 * tp) for lx in range(5) for ly in range(5)]
    for t in range(R):
        d = np.loadtxt(tmp_path / f"bcc2_sf-T{t + 1:03d}.dat")
        assert d.shape == (25, 5)
        want = B.sfactor(spins[t], pos, qs, polar=(0, 0, tp)) / N
        assert np.allclose(d[:, 2:].T, want, rtol=1e-3, atol=2e-4 * N)


def test_binning_error_file(tmp_path):
    """binerror of E against the oracle's binning of the _energy-T series (E/N per measured step,
    E16.8) -- the same series the engine's trace holds."""
    from oracle import obs_oracle as B
    extra = """
 steps_warmup = 50
 steps_mc = 64
 steps_measure = 1
 binning_error = .true.
 binning_level = 1-3
 energy_write = .true.
"""
    _two_kind_bcc(tmp_path, 5, extra)
    _run(tmp_path, "bcc2")
    lines = (tmp_path / "bcc2_binerror.dat").read_text().splitlines()
    blocks = [i for i, l in enumerate(lines) if l.strip().startswith("#T=")]
    assert len(blocks) == 2
    for t, b in enumerate(blocks):
        got = np.array([[float(v) for v in l.split()] for l in lines[b + 1:b + 4]])
        assert got[:, 0].tolist() == [1, 2, 3]
        e = np.loadtxt(tmp_path / f"bcc2_energy-T{t + 1:03d}.dat", skiprows=3)[:, 1]
        assert len(e) == 64
        var = B.binning(e[:, None], [1, 2, 3])
        assert np.allclose(got[:, 1], np.sqrt(var[0] / 128), rtol=1e-4, atol=1e-12)


def test_spin_glass_initial_energy(tmp_path):
    """spin_glass = true with sig= on every J row: the driver's initial (ferro) energy equals the
    oracle's energy of the same state on the disordered table mcin.to_model builds (same rows,
    same spin_glass_seed, esp_spin_glass on both sides), and differs from the clean one."""
    from espins_amd import lattice as Lt
    from espins_amd import mcin
    from oracle import oracle as O
    text = open(os.path.join(GOLDEN, "examples", "ex01_cubic.mcin")).read()
    out = []
    for l in text.splitlines():
        if "jij=" in l.replace(" ", "").lower() and l.strip().lower().startswith(("f1", "c1")):
            l = l.split("!")[0].rstrip() + ":sig=0.003"
        out.append(l)
    text = "\n".join(out) + "\n spin_glass = .true.\n spin_glass_seed = 11\n"
    (tmp_path / "sg.mcin").write_text(text)
    p = tmp_path / "sg.mcin"
    t = p.read_text()
    for key, val in {"steps_warmup": 10, "steps_mc": 10, "seeds": 3}.items():
        pat = re.compile(rf"^\s*!?!?\s*{key}\s*[=:].*$", re.I | re.M)
        t = pat.sub(f" {key} = {val}", t, count=1) if pat.search(t) else f" {key} = {val}\n" + t
    p.write_text(t)
    res = _run(tmp_path, "sg")
    e0 = [float(x) for x in re.findall(r"Initial Energy\s*=\s*(\S+)", res)]
    assert re.search(r"Used seed for normal random generator\s*:\s*11\s*\|", res)
    inp = mcin.read(str(p))
    m = mcin.to_model(inp, builder=O.build_table_fn)
    want = O.Oracle(m).energy(Lt.ferro_spins(m.N, 1, m.mag)[0])
    assert len(e0) == 20
    assert np.allclose(e0, want, rtol=2e-7)
    assert abs(want - (-0.6684205e4)) > 1.0


def test_pt_learned_proposal(tmp_path):
    """Parallel tempering with mcarlo_mode = learned: training after warmup2, the learned sweep
    captured in the PT rounds; the final replica map is a permutation and the running energies
    equal the direct ones at the end (ENERGY3 rows)."""
    _prepare(tmp_path, "ex02_cubic.mcin", "cubic",
             {"steps_warmup": 400, "steps_mc": 600, "supercell_size": "6 6 6", "seeds": 21,
              "energy_write": ".false.", "mcarlo_mode": "learned", "learned_records": 32768,
              "learned_epochs": 2})
    out = _run(tmp_path, "cubic")
    assert "Learned proposal: training on" in out
    tail = out[out.rindex("Final configuration =") + len("Final configuration ="):].split()
    perm = [int(v) for v in tail[:32]]           # list-directed output wraps over lines
    assert sorted(perm) == list(range(1, 33))
    rows = [l.split() for l in out.splitlines() if "<-- ENERGY3" in l]
    adds = [float(r[-3]) for r in rows if "adding" in " ".join(r)]
    direct = [float(r[-3]) for r in rows if "direct" in " ".join(r)]
    assert len(adds) == len(direct) == 32
    assert np.allclose(adds, direct, rtol=1e-6)
    mine = np.loadtxt(tmp_path / "cubic_mc.dat", comments="#")
    assert mine.shape == (32, 7) and np.isfinite(mine).all()
