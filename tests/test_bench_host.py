"""Host logic of bench.py (no GPU): the training-path keys and the choice of
the fastest path per workload within the accuracy bounds."""
import bench


def test_train_precisions():
    assert bench.train_precisions("fp16") == ("fp16", "fp16")
    assert bench.train_precisions("fp16/bf16x3") == ("fp16", "bf16x3")
    assert bench.train_precisions("fp32") == ("fp32", "fp32")


def _trace(ms, err):
    return {"ms": ms, "rays_per_s": 1e3 / ms, "vs_fp64": {"eval_err_max": err}}


def _train(ms, err):
    return {"ms_per_step": ms, "vs_fp64": {"param_grad_rel_err_max_tensor": err}}


def test_fastest_within_bounds():
    results = {"bf16": {"ms": 36.0}, "fp16": {"ms": 38.0}, "bf16x3f": {"ms": 79.0}, "fp32": {"ms": 437.0}}
    errs = {"bf16": {"max": 2e-3}, "fp16": {"max": 2.2e-4}, "bf16x3f": {"max": 5e-5}, "fp32": {"max": 6e-7}}
    # FP16 is within the bound at its hits but not over the grid: the tracing pick skips it
    tr = {"bf16": _trace(40.0, 7e-4), "fp16": _trace(42.0, 9e-5), "bf16x3f": _trace(95.0, 4e-5),
          "fp32": _trace(590.0, 2e-7), "stack3_bf16": _trace(160.0, 7e-4), "stack3_bf16x3f": _trace(250.0, 4e-5),
          "stack3_fp32": _trace(1700.0, 2e-7)}
    tn = {"bf16": _train(7.6, 2e-2), "fp16": _train(7.8, 2.8e-3), "fp16/bf16x3": _train(8.9, 7e-4),
          "fp32": _train(39.6, 6e-6)}
    p = bench.fastest_within_bounds(results, "bf16x3f", errs, tr, tn)
    assert p["extraction"] == {"path": "bf16x3f", "ms": 79.0, "max_sdf_err": 5e-5, "vs_fp32": round(437 / 79, 2)}
    assert p["trace_16x1080p_coarse"]["path"] == "bf16x3f"
    assert p["trace_16x1080p_coarse"]["vs_fp32"] == round(590 / 95, 2)
    assert p["trace_16x1080p_stack3"]["path"] == "bf16x3f"
    assert p["trace_16x1080p_stack3"]["vs_fp32"] == round(1700 / 250, 2)
    assert p["train_step_2e18"]["path"] == "fp16/bf16x3"
    assert p["train_step_2e18"]["vs_fp32"] == round(39.6 / 8.9, 2)
    # without the checker leg or the side figures only the extraction entry remains
    p = bench.fastest_within_bounds(results, "bf16x3f", None, None, None)
    assert set(p) == {"extraction"} and p["extraction"]["max_sdf_err"] is None
