"""Host logic of LevelStack (no GPU): what a re-pack copies from the modules."""
import torch

from mplicits.model import SIREN
from mplicits.stack import LevelStack


def _host_stack(models):
    # the packing inputs only (no library handle): _copy_weights fills the
    # flat FP32 weight and bias arrays the device packer reads
    st = LevelStack.__new__(LevelStack)
    st.models = list(models)
    st._flat_w = [torch.empty(sum(p.numel() for n, p in m.named_parameters() if n.endswith("weight")))
                  for m in models]
    st._flat_b = [torch.empty(sum(p.numel() for n, p in m.named_parameters() if n.endswith("bias")))
                  for m in models]
    st._copy_weights()
    return st


def test_repack_after_update_omegas_keeps_the_function():
    # i3d checkpoints are saved after update_omegas(1) (the frequency baked into
    # the weights); a stack built before that call must re-pack the baked
    # weights with the new frequency 1, not with the old one.
    torch.manual_seed(5)
    m = SIREN(3, 1, [32, 32], w0=30, ww=30)
    st = _host_stack([m])
    w_before, b_before = st._flat_w[0].clone(), st._flat_b[0].clone()
    v_before = st._current_versions()
    m.update_omegas(1)
    assert st._current_versions() != v_before  # refresh() would re-pack
    st._copy_weights()
    assert torch.equal(st._flat_w[0], w_before)
    assert torch.equal(st._flat_b[0], b_before)
    assert st._current_versions() == st._versions


def test_frequency_change_alone_triggers_a_repack():
    torch.manual_seed(6)
    m = SIREN(3, 1, [16], w0=30, ww=30)
    st = _host_stack([m])
    v = st._current_versions()
    m.net[0][1].w0 = 20  # no parameter changed
    assert st._current_versions() != v
    st._copy_weights()
    w0 = m.net[0][0].weight.detach().reshape(-1)
    assert torch.equal(st._flat_w[0][:w0.numel()], w0 * 20)
