"""Renderer formats (.bin pairs and the .exp descriptor): round trip of a
three-level stack, byte layout, descriptor parsing and its errors."""
import numpy as np
import pytest
import torch

from mplicits import renderer_io
from mplicits.synthetic import seeded_siren


def _levels():
    return [seeded_siren([64] * 3, 30.0, 1), seeded_siren([128] * 2, 60.0, 2, last_scale=0.1),
            seeded_siren([32] * 4, 90.0, 3, last_scale=0.1)]


def test_round_trip_is_bit_exact(tmp_path):
    levels = _levels()
    desc = renderer_io.export_experiment(levels, str(tmp_path / "data" / "released"), "bumpy", data_root="data")
    text = open(desc).read()
    assert "lod0_layers = 2" in text and "lod0_hidden = 64" in text
    assert "lod1_layers = 1" in text and "lod2_layers = 3" in text and "lod2_hidden = 32" in text
    assert "lod1_weights = data/released/bumpy_medium_weights.bin" in text
    loaded, e = renderer_io.load_experiment(desc, data_root=str(tmp_path))
    assert e["surface_w0"] == 1 and e["is3d"] and not e["flip_y"]
    for a, b in zip(levels, loaded):
        sa, sb = a.state_dict(), b.state_dict()
        assert list(sa) == list(sb)
        for k in sa:
            assert torch.equal(sa[k], sb[k]), k


def test_bin_layout_is_row_major_weights_then_biases(tmp_path):
    m = _levels()[0]
    renderer_io.export_experiment([m, m, m], str(tmp_path / "d"), "x", data_root="data")
    w = np.fromfile(tmp_path / "d" / "x_coarse_weights.bin", dtype="<f4")
    b = np.fromfile(tmp_path / "d" / "x_coarse_biases.bin", dtype="<f4")
    lin = [seq[0] for seq in m.net]
    assert w.size == sum(l.weight.numel() for l in lin) and b.size == sum(l.bias.numel() for l in lin)
    assert np.array_equal(w[:3 * 64].reshape(64, 3), lin[0].weight.detach().numpy())
    assert np.array_equal(w[3 * 64:3 * 64 + 64 * 64].reshape(64, 64), lin[1].weight.detach().numpy())
    assert np.array_equal(b[-1:], lin[-1].bias.detach().numpy())


def test_descriptor_errors(tmp_path, capsys):
    with pytest.raises(FileNotFoundError):
        renderer_io.parse_experiment(str(tmp_path / "missing.exp"))
    p = tmp_path / "bad.exp"
    p.write_text("# comment\nlod0_weights = a.bin\nlod1_weights = b.bin\nfancy_key = 3\n")
    with pytest.raises(ValueError):
        renderer_io.parse_experiment(str(p))
    assert "unknown key 'fancy_key' ignored" in capsys.readouterr().err
    levels = _levels()
    desc = renderer_io.export_experiment(levels, str(tmp_path / "r"), "m", data_root="data")
    with pytest.raises(FileNotFoundError):
        renderer_io.load_experiment(desc, data_root=str(tmp_path / "nowhere"))
