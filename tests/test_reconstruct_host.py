"""reconstruct.py on a host without a GPU: the drop-in has no CPU path and says
so (the reference would fall back to torch on the CPU, reconstruct.py:86-89)."""
import importlib.util
import os.path as osp

import pytest
import torch

from conftest import ROOT


@pytest.mark.skipif(torch.cuda.device_count() > 0, reason="checks the behaviour without a GPU")
def test_cli_refuses_to_run_without_a_gpu(tmp_path):
    spec = importlib.util.spec_from_file_location("reconstruct_cli", osp.join(ROOT, "m-plicits_amd", "reconstruct.py"))
    cli = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cli)
    with pytest.raises(RuntimeError, match="GPU only"):
        cli.main([str(tmp_path / "missing.pth"), str(tmp_path / "out.ply"), "--device", "cuda"])


def test_cli_arguments_are_the_reference_ones():
    spec = importlib.util.spec_from_file_location("reconstruct_cli", osp.join(ROOT, "m-plicits_amd", "reconstruct.py"))
    cli = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cli)
    src = open(osp.join(ROOT, "m-plicits_amd", "reconstruct.py")).read()
    for opt in ("model_path", "output_path", "--coarse_path", "--medium_path", "--w0", "--resolution", "-r",
                "--device", "--multistage", "--input", "--deltas"):
        assert f'"{opt}"' in src, opt
    assert callable(cli.adaptive_deltas)
