"""The multi-rank CPU baseline's mesh split: a gmsh-2.2 file in x-strips with
ghost tags (cases.write_msh_strips), partitioned by the reference's own
grid_part and run by the reference on several MPI ranks, gives the one-rank
run's residual log.  Runs where the reference's binaries are built
(oracle/_ref, from its sources) and MPICH's mpirun is present; skipped
elsewhere."""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))

from cases import riemann_mesh, write_msh_strips, write_part_file  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "tensum")
PART = os.path.join(ROOT, "oracle", "_ref", "grid_part")
MPIRUN = "/opt/conda/bin/mpirun"

pytestmark = pytest.mark.skipif(
    not all(os.access(p, os.X_OK) for p in (EXE, PART, MPIRUN)),
    reason="the reference's binaries (oracle/ref.mk) or MPICH's mpirun are absent")


def _param(d, t_final):
    from ref_cases import param_text
    text = param_text("riemann33", t_final, 10 ** 8).replace(
        "   restart               yes", "   restart               no")
    with open(os.path.join(d, "param.in"), "w") as f:
        f.write(text)


def _log(d):
    return np.loadtxt(os.path.join(d, "SAMPLE_0", "residual.dat"))


def test_strips_have_ghosts_and_owned_boundary_lines(tmp_path):
    m = riemann_mesh(17)
    owner = write_msh_strips(m, str(tmp_path / "mesh.msh"), 3)
    assert sorted(set(owner.tolist())) == [0, 1, 2]
    text = open(tmp_path / "mesh.msh").read().split("$Elements\n")[1].splitlines()
    n = int(text[0])
    rows = [list(map(int, r.split())) for r in text[1:1 + n]]
    lines = [r for r in rows if r[1] == 1]
    tris = [r for r in rows if r[1] == 2]
    assert len(lines) == m.bface_tag.size and len(tris) == m.cell_v.size // 3
    assert all(r[2] == 4 and r[5] == 1 and 1 <= r[6] <= 3 for r in lines)
    # every triangle: owner first, ghosts negative; strips 0 and 2 never touch
    ghosts = 0
    for r in tris:
        ntags, tags = r[2], r[3:3 + r[2]]
        assert tags[2] == ntags - 3 and tags[3] > 0
        g = [-t for t in tags[4:]]
        assert all(t > 0 for t in g) and all(abs(t - tags[3]) == 1 for t in g)
        ghosts += len(g)
    assert ghosts > 0


@pytest.mark.parametrize("ranks", [2, 3])
def test_reference_on_strips_matches_one_rank(tmp_path, ranks):
    m = riemann_mesh(33)
    t_final = 0.02
    one = tmp_path / "one"
    many = tmp_path / "many"
    one.mkdir()
    many.mkdir()
    write_part_file(m, str(one))
    _param(str(one), t_final)
    r = subprocess.run([EXE, "-i", "param.in", "-npart", "1", "-L", "."], cwd=one,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    write_msh_strips(m, str(many / "mesh.msh"), ranks)
    r = subprocess.run([PART, "-D", "2", "-I", str(many / "mesh.msh"), "-P", str(ranks), "-L",
                        str(many)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    assert sorted(os.listdir(many / "PARTITION")) == [f"part_file_{p}.dat" for p in range(ranks)]
    _param(str(many), t_final)
    r = subprocess.run([MPIRUN, "-np", str(ranks), EXE, "-i", "param.in", "-npart", str(ranks),
                        "-L", "."], cwd=many, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    a, b = _log(str(one)), _log(str(many))
    assert a.shape == b.shape and a.shape[0] > 10
    # same steps, dt and time; the residual norm to its printed digits (the
    # partitioner re-prints coordinates at 15-16 digits and the ranks sum the
    # shared vertices in another order)
    np.testing.assert_array_equal(a[:, 0], b[:, 0])
    np.testing.assert_allclose(a[:, 1:3], b[:, 1:3], rtol=1e-10)
    np.testing.assert_allclose(a[:, 3], b[:, 3], rtol=1e-8)
