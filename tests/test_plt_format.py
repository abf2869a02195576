"""The independent Tecplot reader/writer used by the restart tests: a
round trip of a two-zone file with the reference's variable layout."""
import numpy as np

import plt_format


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    names = ["x", "y", "z", "density", "u", "mach"]
    zones = [((4, 3, 1), {n: rng.standard_normal(12) for n in names}),
             ((2, 2, 2), {n: rng.standard_normal(8) for n in names})]
    p = str(tmp_path / "f.plt")
    plt_format.write(p, names, zones, time=1.25e-3)
    f = plt_format.read(p)
    assert f["names"] == names and f["title"] == "Solution file" and f["order"] == 1
    for (dims, data), z in zip(zones, f["zones"]):
        assert z["dims"] == dims and z["time"] == 1.25e-3 and set(z["formats"]) == {2}
        for q, n in enumerate(names):
            np.testing.assert_array_equal(z["data"][n], data[n])
            assert z["minmax"][q].tolist() == [data[n].min(), data[n].max()]
