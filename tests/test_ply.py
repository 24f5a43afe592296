"""Binary PLY writer (save_ply, i3d/meshing.py:316-384): layout and round trip."""
import numpy as np

from mplicits.ply import read_ply, save_ply


def test_header_and_round_trip(tmp_path):
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0.5]], dtype=np.float64)
    faces = np.array([[0, 1, 2]])
    fn = tmp_path / "t.ply"
    save_ply(verts, faces, str(fn))
    raw = fn.read_bytes()
    header = raw[:raw.index(b"end_header\n") + 11].decode()
    assert header == ("ply\nformat binary_little_endian 1.0\nelement vertex 3\n"
                      "property float x\nproperty float y\nproperty float z\n"
                      "element face 1\nproperty list uchar int vertex_indices\nend_header\n")
    assert len(raw) == len(header) + 3 * 12 + 1 * 13
    d = read_ply(str(fn))
    np.testing.assert_array_equal(np.stack([d["vertex"]["x"], d["vertex"]["y"], d["vertex"]["z"]], 1),
                                  verts.astype(np.float32))
    np.testing.assert_array_equal(d["face"], faces)


def test_vertex_attributes_and_empty(tmp_path):
    verts = np.random.default_rng(0).random((5, 6)).astype(np.float32)
    faces = np.array([[0, 1, 2], [2, 3, 4]])
    attrs = [("nx", "f4"), ("ny", "f4"), ("nz", "f4")]
    fn = tmp_path / "n.ply"
    save_ply(verts, faces, str(fn), vertex_attributes=attrs)
    d = read_ply(str(fn))
    np.testing.assert_array_equal(d["vertex"]["nz"], verts[:, 5])
    fn2 = tmp_path / "e.ply"
    save_ply(np.zeros((0, 3)), np.zeros((0, 3)), str(fn2))
    d = read_ply(str(fn2))
    assert len(d["vertex"]) == 0 and len(d["face"]) == 0
