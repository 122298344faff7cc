"""The Tecplot writer (cnd_write_plt, host only): the layout of the
reference's techead + write_output_file (src/mo_io.f90:345-562) read back
with an independent reader of the binary format, and the byte offsets the
reference's header counter lh gives for a known case."""
import numpy as np

from canard_amd.solver import write_plt
from plt_reader import read_plt


def test_two_zone_solution_file_round_trips(tmp_path):
    rng = np.random.default_rng(3)
    dims = [(5, 4, 3), (2, 6, 2)]
    data = [rng.standard_normal((5, int(np.prod(d)))).astype(np.float32) for d in dims]
    path = tmp_path / "n007.plt"
    names = ["rho", "u", "v", "w", "p"]
    write_plt(path, "n007", names, list(zip(dims, data)), time=1.25, file_type=2)
    f = read_plt(path)
    assert f["title"] == "n007" and f["file_type"] == 2 and f["names"] == names
    assert [z["name"] for z in f["zones"]] == ["z000", "z001"]
    assert [z["strand"] for z in f["zones"]] == [1, 2]
    assert all(z["parent"] == -1 and z["time"] == 1.25 for z in f["zones"])
    for z, d, a in zip(f["zones"], dims, data):
        assert z["dims"] == d
        assert np.array_equal(z["data"], a)
        assert np.array_equal(z["minmax"][:, 0], a.min(axis=1).astype(np.float64))
        assert np.array_equal(z["minmax"][:, 1], a.max(axis=1).astype(np.float64))


def test_grid_file_header_size_follows_the_reference_counter(tmp_path):
    """techead advances lh (4-byte units) by: 2 (magic) + 1 + 1 + title
    chars + 1 + 1 (nvar) + names; per zone 1 (marker) + name + 1 (parent) +
    1 (strand) + 2 (time) + 1 (-1) + 4 (zone type, location, neighbours,
    connections) + 3 (I, J, K) + 1 (auxiliary); the end marker 1; then per
    zone 1 + nvar + 3 + 4 nvar.  For a grid file of one 3x2x2 zone the data
    therefore start at byte 4 * 55."""
    dims = (3, 2, 2)
    xyz = np.arange(3 * 12, dtype=np.float32).reshape(3, 12)
    path = tmp_path / "grid.plt"
    write_plt(path, "grid", ["x", "y", "z"], [(dims, xyz)], time=0.0, file_type=1)
    b = open(path, "rb").read()
    lh = 2 + 1 + 1 + (4 + 1) + 1 + 3 * 2          # magic .. names
    lh += 1 + (4 + 1) + 1 + 1 + 2 + 1 + 4 + 3 + 1  # zone header
    lh += 1                                        # end of header
    lh += 1 + 3 + 3 + 4 * 3                        # zone data record
    assert lh == 55
    assert len(b) == 4 * lh + 4 * xyz.size
    assert np.array_equal(np.frombuffer(b[4 * lh:], dtype="<f4"), xyz.ravel())
    f = read_plt(path)
    assert f["file_type"] == 1 and f["title"] == "grid"
