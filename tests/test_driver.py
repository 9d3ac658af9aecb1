"""The `tensum` command-line driver: param.in grammar (CPU) and a full run (GPU).

The param.in fixtures tests/golden/param_isentropic_vortex.in and
param_uq_shocktube.in are the reference's own example inputs
(examples/Euler/isentropic_vortex/param.in, examples/UQ/shocktube/param.in).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_binding as O
import tensum_amd as T
from cases import VORTEX_PARAM, rel_l2, vortex, vortex_mesh, write_part_file

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tensum_amd", "tensum")
GOLD = os.path.join(ROOT, "tests", "golden")


def check(path):
    r = subprocess.run([EXE, "--check", "-i", path], capture_output=True, text=True, timeout=60)
    return r.returncode, dict(l.split(" ", 1) for l in r.stdout.splitlines() if " " in l), r


def test_reads_reference_vortex_example():
    rc, kv, _ = check(os.path.join(GOLD, "param_isentropic_vortex.in"))
    assert rc == 0
    assert kv["cell_type"] == "1" and kv["time_scheme"] == "rk4" and kv["flux"] == "2"
    assert float(kv["cfl"]) == 0.4 and float(kv["final_time"]) == 20.0
    assert kv["reconstruct"] == "1" and kv["time_stamps"] == "10" and kv["n_bc"] == "4"
    # initial state at (0.25, -0.5): the example's own expressions, in numpy
    gam, beta, M, alpha, x, y = 1.4, 5.0, 1.0, 0.0, 0.25, -0.5
    r2 = x * x + y * y
    Tm = 1.0 - (gam - 1) * (beta ** 2 / (8 * gam * np.pi ** 2)) * np.exp(1 - r2)
    u = M * np.cos(alpha * np.pi / 180) - beta / (2 * np.pi) * y * np.exp(0.5 * (1 - r2))
    v = M * np.sin(alpha * np.pi / 180) + beta / (2 * np.pi) * x * np.exp(0.5 * (1 - r2))
    p = Tm ** (gam / (gam - 1))
    lines = subprocess.run([EXE, "--check", "-i", os.path.join(GOLD, "param_isentropic_vortex.in")],
                           capture_output=True, text=True).stdout.splitlines()
    ic = {l.split()[1]: float(l.split()[2]) for l in lines if l.startswith("ic ")}
    np.testing.assert_allclose([ic["temperature"], ic["xvelocity"], ic["yvelocity"],
                                ic["pressure"]], [Tm, u, v, p], rtol=1e-15)


def test_reads_reference_uq_example_with_pert():
    lines = subprocess.run([EXE, "--check", "-i", os.path.join(GOLD, "param_uq_shocktube.in")],
                           capture_output=True, text=True).stdout.splitlines()
    kv = dict(l.split(" ", 1) for l in lines)
    assert kv["n_samples"] == "11" and kv["rnd_per_sample"] == "6" and kv["flux"] == "0"
    assert kv["reconstruct"] == "4" and kv["nd_active"] == "1"
    ic = {l.split()[1]: float(l.split()[2]) for l in lines if l.startswith("ic ")}
    # pert(4,0,0,k) returns 0.1 k in the check mode; x = 0.25 < xc*(1+dd*(0.1-0.5))
    xc, dd, Tl, pl = 0.3, 0.1, 1.0, 1.0
    assert 0.25 < xc * (1 + dd * (0.1 - 0.5))
    assert ic["temperature"] == pytest.approx(Tl * (1 + dd * (0.2 - 0.5)), rel=1e-15)
    assert ic["pressure"] == pytest.approx(pl * (1 + dd * (0.5 - 0.5)), rel=1e-15)


@pytest.mark.parametrize("mutation,msg", [
    (("kepes_tecno_roe", "kepes_magic"), "unknown flux scheme"),
    (("cfl          0.8", "cfl          0.0"), "Either time_step or cfl"),
    (("method   minmod", "method   superbee"), "unknown reconstruction"),
])
def test_rejects_bad_parameters(tmp_path, mutation, msg):
    text = VORTEX_PARAM.format(final_time=1.0, stamps=2).replace(*mutation)
    f = tmp_path / "param.in"
    f.write_text(text)
    rc, _, r = check(str(f))
    assert rc == 1 and msg in r.stderr


@pytest.mark.gpu
def test_driver_run_matches_oracle(tmp_path):
    mesh = vortex_mesh(17)
    write_part_file(mesh, str(tmp_path), periodic_tags=(1, 2, 3, 4))
    (tmp_path / "param.in").write_text(VORTEX_PARAM.format(final_time=1.0, stamps=2))
    r = subprocess.run([EXE, "-i", "param.in", "-npart", "1", "-L", "."], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    sd = tmp_path / "SAMPLE_0"
    res = np.loadtxt(sd / "residual.dat")
    o = O.Oracle(mesh, T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8,
                                final_time=1.0, time_stamps=2))
    # the param.in expressions evaluated as written (no periodic wrap of x, y)
    o.set_state(vortex(mesh.x, mesh.y, box=None))
    n, log = o.run(10 ** 6, want_log=True)
    assert res.shape[0] == n
    np.testing.assert_allclose(res[:, 1], log[:, 1], rtol=1e-11)    # dt column
    np.testing.assert_allclose(res[:, 2], log[:, 2], rtol=1e-11)    # time column
    for k in (0, 1, 2):
        assert (sd / f"sol000{k}_0.vtk").exists()
    master = (sd / "master_file.visit").read_text().split()
    assert master[:2] == ["!NBLOCKS", "1"] and master[2:] == [f"sol000{k}_0.vtk" for k in range(3)]
    err = (sd / "error_norms.dat").read_text().splitlines()
    dens = [float(v) for v in err[6].split(":")[1].split()]
    ref = o.error_norms(vortex(mesh.x, mesh.y, 1.0, box=None))[5]
    np.testing.assert_allclose(dens, ref, rtol=1e-6)
    glob = np.loadtxt(sd / "global.dat")
    assert glob.shape[0] == n
    head = int((sd / "LAST_RESTART_HEAD.txt").read_text())
    rst = np.loadtxt(sd / f"restart_{head}_0.dat", max_rows=mesh.n_vertex)
    assert rst.shape == (mesh.n_vertex, 5)


@pytest.mark.gpu
def test_driver_steady_run_matches_oracle(tmp_path):
    """time_mode steady: local time steps, the 8-column steady residual log
    (src/fv.cc:1944-1957), sol0001 rewritten every `frequency` steps and listed
    once, at the end (src/fv.cc:2241-2245, :2318-2323, :2371-2377)."""
    mesh = vortex_mesh(17)
    write_part_file(mesh, str(tmp_path), periodic_tags=(1, 2, 3, 4))
    text = VORTEX_PARAM.format(final_time=1.0, stamps=2)
    for a, b in (("time_mode    unsteady", "time_mode    steady"), ("max_iter     5000", "max_iter     30"),
                 ("frequency             100", "frequency             10"),
                 ("find_error            yes", "find_error            no")):
        assert a in text
        text = text.replace(a, b)
    (tmp_path / "param.in").write_text(text)
    r = subprocess.run([EXE, "-i", "param.in", "-npart", "1", "-L", "."], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    sd = tmp_path / "SAMPLE_0"
    res = np.loadtxt(sd / "residual.dat")
    o = O.Oracle(mesh, T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8,
                                time_mode="steady", max_iter=30))
    o.set_state(vortex(mesh.x, mesh.y, box=None))
    n, log = o.run(10 ** 6, want_log=True)
    assert n == 30 and res.shape == (30, 8)
    assert np.array_equal(res[:, 0], log[:, 0])
    np.testing.assert_allclose(res[:, 1], log[:, 1], rtol=6e-5)   # dt_global, 5 digits
    np.testing.assert_allclose(res[:, 2], log[:, 3], rtol=6e-5)   # normalised residual
    comp = res[:, 3:]
    assert np.all(comp[:, 3] == 0.0)                              # z momentum
    tot = np.sqrt((comp[:, [0, 1, 2, 4]] ** 2).sum(1))
    np.testing.assert_allclose(tot / tot[0], log[:, 3], rtol=3e-4)
    master = (sd / "master_file.visit").read_text().split()
    assert master == ["!NBLOCKS", "1", "sol0000_0.vtk", "sol0001_0.vtk"]
    assert not (sd / "sol0002_0.vtk").exists()
    text = (sd / "sol0001_0.vtk").read_text()
    assert "TIME" not in text                                     # steady files carry no time
    lines = text.splitlines()
    i = lines.index("SCALARS pressure double 1")
    p = np.array([float(v) for v in lines[i + 2:i + 2 + mesh.n_vertex]])
    np.testing.assert_allclose(p, o.get_state()[:, 3], rtol=1e-10)
