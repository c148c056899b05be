"""The reference's network files (Read_Network.cpp readers, Printer.cpp writers).

* tests/golden/netio_csv holds files in the reference's text layout, with the spacing
  variants its stringstream parsing accepts ("a, b", "a,b", leading blanks, exponents);
* the binary layout is written here independently with struct.pack, field by field as
  Printer::Print_cnt_points_and_structure_binary / Print_gnp_data_binary write it;
* round trips through the writers: text exact to the printed digits, binary bit for bit;
* Printer::Recover_angle's quirk: an angle of 0 is written as 2 pi;
* short, malformed and inconsistent files are rejected."""
import os
import struct

import numpy as np
import pytest

from cicy_geom_amd import netio

GOLD = os.path.join(os.path.dirname(__file__), "golden", "netio_csv")


def test_golden_text_files():
    corner, lengths = netio.read_sample_geometry(GOLD)
    assert corner.tolist() == [0.0, -1.5, 2.0] and lengths.tolist() == [3.0, 4.5, 6.0]
    t = netio.read_cnts(GOLD, "csv")
    assert t["npts"].tolist() == [2, 3, 1]
    assert t["radii"].tolist() == [0.005, 0.0075, 0.01]
    assert t["xyz"].tolist() == [[0.1, 0.2, 0.3], [0.4, 0.5, 0.6], [1e-3, -0.25, 7.0],
                                 [1.0, 1.0, 1.0], [2.0, 3.0, 4.0], [-0.5, 0.25, 0.125]]
    assert [s.tolist() for s in t["structure"]] == [[0, 1], [2, 3, 4], [5]]
    g = netio.read_gnps(GOLD, "csv")
    assert g["geo"].tolist() == [[1, 0.03, 1.5708, 0.785398, 1, 2, 3],
                                 [0.5, 0.01, 6.28319, 3.14159, -1, 0, 0.5]]
    for (l, t_, th, ph, *_), m in zip(g["geo"], g["rot"]):
        np.testing.assert_allclose(m, netio.rotation(th, ph), rtol=0, atol=1e-15)
        assert m[2, 1] == 0.0
        np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-14)  # a rotation


def pack_files(d, npts, radii, xyz, plates):
    with open(os.path.join(d, "cnt_struct.dat"), "wb") as f:
        f.write(struct.pack("<i", len(npts)))
        for p, r in zip(npts, radii):
            f.write(struct.pack("<id", p, r))
    with open(os.path.join(d, "cnt_coordinates.dat"), "wb") as f:
        for p in xyz:
            f.write(struct.pack("<3d", *p))
    with open(os.path.join(d, "gnp_data.dat"), "wb") as f:
        f.write(struct.pack("<i", len(plates)))
        for row in plates:
            f.write(struct.pack("<7d", *row))


def test_binary_layout(tmp_path):
    rng = np.random.default_rng(3)
    npts = [5, 1, 0, 7]
    radii = rng.uniform(0.005, 0.01, 4)
    xyz = rng.standard_normal((13, 3))
    plates = rng.uniform(0, 1, (3, 7))
    pack_files(tmp_path, npts, radii, xyz, plates)
    t = netio.read_cnts(tmp_path, "dat")
    assert t["npts"].tolist() == npts
    assert np.array_equal(t["radii"], radii) and np.array_equal(t["xyz"], xyz)
    g = netio.read_gnps(tmp_path, "dat")
    assert np.array_equal(g["geo"], plates)


def random_network(seed, n_cnts=40, n_gnps=6):
    rng = np.random.default_rng(seed)
    npts = rng.integers(1, 30, n_cnts)
    xyz = rng.uniform(-1, 5, (int(npts.sum()), 3))
    radii = rng.uniform(0.005, 0.01, n_cnts)
    th = rng.uniform(0.1, 3.0, n_gnps)
    ph = rng.uniform(0.1, 6.2, n_gnps)
    rot = np.stack([netio.rotation(a, b) for a, b in zip(th, ph)])
    return dict(npts=npts, xyz=xyz, radii=radii, side=rng.uniform(0.5, 2, n_gnps),
                thick=rng.uniform(0.01, 0.05, n_gnps), rot=rot,
                centre=rng.uniform(0, 4, (n_gnps, 3)), theta=th, phi=ph)


@pytest.mark.parametrize("prec", [4, 10])
def test_text_round_trip(tmp_path, prec):
    n = random_network(1)
    netio.write_network(tmp_path, "csv", [0.5, 0.25, 0.0], [4.0, 4.0, 3.5], n["npts"],
                        n["radii"], n["xyz"], n["side"], n["thick"], n["rot"], n["centre"],
                        prec_cnt=prec, prec_gnp=prec)
    t = netio.read_cnts(tmp_path, "csv")
    assert np.array_equal(t["npts"], n["npts"])
    fixed = np.vectorize(lambda x: float(f"{x:.{prec}f}"))
    sig6 = np.vectorize(lambda x: float(f"{x:g}"))
    assert np.array_equal(t["xyz"], fixed(n["xyz"]))
    assert np.array_equal(t["radii"], sig6(n["radii"]))
    corner, lengths = netio.read_sample_geometry(tmp_path)
    assert corner.tolist() == [0.5, 0.25, 0.0] and lengths.tolist() == [4.0, 4.0, 3.5]
    g = netio.read_gnps(tmp_path, "csv")
    assert np.array_equal(g["geo"][:, 0], sig6(n["side"]))
    assert np.array_equal(g["geo"][:, 4:], fixed(n["centre"]))
    # the angles come back to 6 significant digits from the rotation (Recover_angle)
    np.testing.assert_allclose(g["geo"][:, 2], n["theta"], rtol=1e-5)
    np.testing.assert_allclose(g["geo"][:, 3], n["phi"], rtol=1e-5)


def test_binary_round_trip(tmp_path):
    n = random_network(2)
    netio.write_network(tmp_path, "dat", [0, 0, 0], [4, 4, 4], n["npts"], n["radii"], n["xyz"],
                        n["side"], n["thick"], n["rot"], n["centre"])
    t = netio.read_cnts(tmp_path, "dat")
    assert np.array_equal(t["xyz"], n["xyz"]) and np.array_equal(t["radii"], n["radii"])
    g = netio.read_gnps(tmp_path, "dat")
    assert np.array_equal(g["geo"][:, 0], n["side"]) and np.array_equal(g["geo"][:, 1],
                                                                         n["thick"])
    assert np.array_equal(g["geo"][:, 4:], n["centre"])
    np.testing.assert_allclose(g["geo"][:, 2], n["theta"], rtol=1e-12)
    np.testing.assert_allclose(g["geo"][:, 3], n["phi"], rtol=1e-12)
    np.testing.assert_allclose(g["rot"], n["rot"], atol=1e-14)


def test_zero_angle_is_written_as_two_pi(tmp_path):
    """Printer::Recover_angle: cos > Zero and sin <= Zero gives asin(sin) + 2 pi."""
    netio.write_network(tmp_path, "dat", [0, 0, 0], [1, 1, 1], side=[1.0], thick=[0.01],
                        rot=netio.rotation(0.0, 0.0)[None], centre=np.zeros((1, 3)))
    g = netio.read_gnps(tmp_path, "dat")
    assert g["geo"][0, 2] == 2 * 3.1415926535897932384626433832795
    assert g["geo"][0, 3] == 2 * 3.1415926535897932384626433832795


def test_bad_files_are_rejected(tmp_path):
    def write(name, text):
        with open(os.path.join(tmp_path, name), "w") as f:
            f.write(text)
    with pytest.raises(RuntimeError, match="failed to open"):
        netio.read_cnts(tmp_path, "csv")
    write("cnt_struct.csv", "2, 3\n1, 0.005\n1, 0.005\n")
    with pytest.raises(RuntimeError, match="different numbers"):
        netio.read_cnts(tmp_path, "csv")
    write("cnt_struct.csv", "2, 2\n1, 0.005\n")
    with pytest.raises(RuntimeError, match="ends before"):
        netio.read_cnts(tmp_path, "csv")
    write("cnt_struct.csv", "2, 2\n1, 0.005\n2, 0.005\n")
    write("cnt_coordinates.csv", "0, 0, 0\n1, 1, 1\n")
    with pytest.raises(RuntimeError, match="ends before"):
        netio.read_cnts(tmp_path, "csv")
    write("cnt_coordinates.csv", "0, 0, 0\n1, x, 1\n2, 2, 2\n")
    with pytest.raises(RuntimeError, match="malformed line 2"):
        netio.read_cnts(tmp_path, "csv")
    with open(os.path.join(tmp_path, "cnt_struct.dat"), "wb") as f:
        f.write(struct.pack("<iid", 2, 3, 0.005))
    with pytest.raises(RuntimeError, match="ends before"):
        netio.read_cnts(tmp_path, "dat")
    with pytest.raises(ValueError):
        netio.read_cnts(tmp_path, "txt")
