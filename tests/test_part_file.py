"""read_part_file (src/read_part_mesh_file.cc:43-214): a mesh written in the
reference's NPART = 1 part-file layout reads back bit for bit, periodic sets
included; the reference's own example input parses; bad files are refused."""
import numpy as np
import pytest

import tensum_amd as T
from cases import vortex_mesh, write_part_file


def _same(a, b):
    for f in ("x", "y", "bface_tag", "bface_v", "cell_v", "periodic_ptr", "periodic_v"):
        assert np.array_equal(np.asarray(getattr(a, f)), np.asarray(getattr(b, f))), f


@pytest.mark.parametrize("periodic", [False, True])
def test_round_trip(tmp_path, periodic):
    m = vortex_mesh(9) if periodic else T.jittered_mesh(11, seed=4)
    write_part_file(m, str(tmp_path), periodic_tags=(1, 2, 3, 4) if periodic else ())
    r = T.read_part_file(str(tmp_path / "PARTITION" / "part_file_0.dat"), m.bc)
    _same(r, m)
    assert r.bc == m.bc


def test_bad_files_are_refused(tmp_path):
    m = T.jittered_mesh(5, seed=1)
    write_part_file(m, str(tmp_path))
    p = tmp_path / "PARTITION" / "part_file_0.dat"
    text = p.read_text()
    for bad in (text.replace("PARTITIONS 1", "PARTITIONS 2"),
                text.replace("DIMENSION 2", "DIMENSION 3"),
                text[: len(text) // 2],
                text.rsplit(" 0\n", 1)[0] + " 1\n"):   # the last cell a ghost
        p.write_text(bad)
        with pytest.raises(T.TsmError):
            T.read_part_file(str(p), m.bc)
