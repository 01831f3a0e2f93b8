"""Every launch of the engine's forward, saving forward and adjoint stays inside its buffers, checked on the CPU
(meta tensors, a model of each kernel's largest offset; tests/launch_bounds.py) at the shapes the GPU runs: the
benchmark batch, serving tiles, odd small shapes, and the whole 504 x 504 x 256 volume at the checkpoint stride its
adjoint uses."""
import pytest

from launch_bounds import check_shape


def _model(width=64, layers=8, rank=8):
    from ffnoml_amd import synthetic
    from ffnoml_amd.atoms import species
    return synthetic.random_model(0, in_channels=6, out_channels=12, width=width, n_layers=layers,
                                  hidden_spectral=128, hidden_vertical=128, freq_mix_rank=rank,
                                  z_reference_length=40.0, z_min_wavelength=1.28, z_feature_wavelength=0.3,
                                  species=species(["H", "CA"]))


@pytest.mark.parametrize("B,D,H,W,stride", [
    (8, 128, 64, 64, 1),          # the benchmark step
    (1, 256, 504, 504, 2),        # the whole volume
    (1, 256, 504, 504, 1),
    (4, 192, 256, 256, 1),        # largest serving tile, batched
    (3, 96, 32, 32, 3),
    (2, 7, 9, 10, 3),             # odd sizes: partial tiles everywhere
])
def test_engine_launches_stay_in_bounds(B, D, H, W, stride):
    n = check_shape(_model(), B, D, H, W, stride=stride)
    assert n > 100
    if stride == 1:
        assert check_shape(_model(layers=2), B, D, H, W, stride=1, two_pass=False) > 20
        assert check_shape(_model(layers=2), B, D, H, W, stride=0) > 20          # every intermediate kept


@pytest.mark.parametrize("fuse_chunk", [None, 96, 1000, False])
def test_one_sample_chunked_fusion_adjoint_stays_in_bounds(fuse_chunk):
    """The one-sample fusion adjoint without a whole block's gate buffer (stride 2 here): chunks of 96 and 1000
    voxels (a ragged last chunk), the default chunk, and the one-pass kernel."""
    assert check_shape(_model(layers=3), 1, 13, 11, 9, stride=2, fuse_chunk=fuse_chunk) > 50


@pytest.mark.parametrize("B,D,H,W,width", [(2, 20, 12, 14, 64), (1, 33, 9, 7, 96), (3, 16, 8, 8, 32)])
def test_per_column_grids_stay_in_bounds(B, D, H, W, width):
    assert check_shape(_model(width=width, layers=2), B, D, H, W, stride=1, percol=True) > 50


@pytest.mark.parametrize("width,rank", [(96, 8), (32, 0), (16, 4)])
def test_other_widths_stay_in_bounds(width, rank):
    assert check_shape(_model(width=width, layers=3, rank=rank), 2, 20, 12, 14, stride=2) > 50


def test_checker_catches_a_short_buffer_and_covers_every_entry_point():
    """A planted mistake -- the half-spectrum scratch one float short -- is reported, and the runs above exercise
    every compute entry point the header declares."""
    import re
    from pathlib import Path
    import torch
    from ffnoml_amd import engine as E
    import launch_bounds as LB
    real = E.Engine._scratch

    def short(self, plan, dev, adjoint=False):
        scr = real(self, plan, dev, adjoint)
        scr["s1"] = torch.empty(scr["s1"].numel() - 1, device=dev)
        return scr

    E.Engine._scratch = short
    try:
        with pytest.raises(AssertionError, match="s1|tensor"):
            check_shape(_model(layers=1), 1, 8, 8, 8)
    finally:
        E.Engine._scratch = real
    eng, chk, restore = LB.dry_engine(_model(layers=2))
    restore()
    header = (Path(__file__).resolve().parents[1] / "include" / "ffno_hip.h").read_text()
    declared = set(re.findall(r"^int (ffno_\w+)\(", header, re.M)) - {"ffno_version"}
    calls = set()
    orig = LB.Checker.__call__

    def record(self, name, *a):
        calls.add(name)
        return orig(self, name, *a)

    LB.Checker.__call__ = record
    try:
        check_shape(_model(layers=2), 1, 12, 8, 8, stride=2)
        check_shape(_model(layers=2), 1, 12, 8, 8, stride=2, fuse_chunk=False)
        check_shape(_model(layers=2), 1, 12, 8, 8, stride=1, two_pass=True)
        check_shape(_model(layers=2), 1, 12, 8, 8, stride=1, percol=True)
    finally:
        LB.Checker.__call__ = orig
    assert declared <= calls, f"entry points never checked: {sorted(declared - calls)}"


@pytest.mark.parametrize("percol", [False, True])
def test_no_float64_arithmetic_on_the_device(percol, monkeypatch):
    """Every tensor op of the engine's plans, forward, adjoint (with the height gradient) and tangent that touches
    the device (the meta device of the dry run) is float32 or integer, and so is every tensor handed to the library:
    float64 runs on the host only. (The host's forward-mode derivatives of the Filon weights run torch.func, whose
    own shape propagation uses meta tensors; the watch pauses inside them.)"""
    import torch
    from torch.utils._python_dispatch import TorchDispatchMode
    from torch.utils._pytree import tree_leaves
    from ffnoml_amd import engine as E, transforms as T
    seen, paused = [], [0]

    class Watch(TorchDispatchMode):
        def __torch_dispatch__(self, func, types, args=(), kwargs=None):
            out = func(*args, **(kwargs or {}))
            if not paused[0]:
                for t in tree_leaves(out):
                    if torch.is_tensor(t) and t.device.type == "meta" and t.dtype == torch.float64:
                        seen.append(str(func))
            return out

    real_bands = T.filon_height_bands_t

    def host_bands(*a, **k):
        paused[0] += 1
        try:
            return real_bands(*a, **k)
        finally:
            paused[0] -= 1

    monkeypatch.setattr(T, "filon_height_bands_t", host_bands)
    passed = []
    import launch_bounds as LB
    real_dry = LB.dry_engine

    def dry(model):
        eng, chk, restore = real_dry(model)
        inner = E._p
        E._p = lambda t: (passed.append(t.dtype) if t is not None else None) or inner(t)
        return eng, chk, restore

    monkeypatch.setattr(LB, "dry_engine", dry)
    with Watch():
        LB.check_shape(_model(layers=2), 2, 12, 8, 8, stride=1, percol=percol)
    assert not seen, f"float64 on the device: {sorted(set(seen))}"
    assert passed and torch.float64 not in passed


def test_footprint_counts_the_vertical_modes_of_tall_columns():
    """Admission's footprint of a VJP covers the adjoint's own count (Plan.vjp_bytes at stride 1, two-pass fusion;
    Plan.full_bytes for the adjoint that keeps every intermediate, less the plan's cached tables) also when twice the
    column span exceeds the reference length and the box holds more vertical modes."""
    import numpy as np
    import launch_bounds as LB
    from ffnoml_amd import engine as E
    model = _model(layers=2)
    eng, chk, restore = LB.dry_engine(model)
    try:
        for span in (15.0, 30.0):                        # 2 x 30 > 40: a longer box, more modes
            D, H, W = 20, 8, 8
            col = np.linspace(span, 0.0, D)
            plan = E.Plan(eng, 1, D, H, W, col, span, 48e3, 48e3)
            need = eng.footprint_bytes(1, D, H, W, vjp=True, span=span)
            assert plan.modes == eng.vertical_modes(span)
            assert need >= plan.vjp_bytes(1, two_pass=True), (span, need, plan.vjp_bytes(1, two_pass=True))
            # the adjoint that keeps every intermediate, which admission grants when it fits
            full = eng.footprint_bytes(1, D, H, W, vjp=True, span=span, full=True)
            assert eng.footprint_bytes(1, D, H, W, vjp=True, span=span, keep=2) == full     # 2 layers: all kept
            assert need < eng.footprint_bytes(1, D, H, W, vjp=True, span=span, keep=1) < full
            tables = 2 * plan.uncached_table_bytes()
            assert full >= plan.full_bytes(want_z=True) - tables, (span, full, plan.full_bytes(want_z=True))
        assert eng.vertical_modes(30.0) > eng.vertical_modes(15.0)
    finally:
        restore()
