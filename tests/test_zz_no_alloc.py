"""The allocation-free entry point ffno_post_gn_noalloc (its partial GroupNorm statistics in the library's static
device scratch) against the entry points it stands for, ffno_post with a workspace then ffno_gn_finalize, through the
C ABI: y, the statistics and the affine bit for bit, on two calls in a row, the second smaller (the scratch keeps the
first call's contents), and a size beyond the scratch rejected before any launch."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _case(B, voxels, C, nwg, seed):
    from ffnoml_amd import _lib
    from ffnoml_amd.engine import pack_mlp_weight
    g = torch.Generator().manual_seed(seed)
    u = torch.randn(B * voxels, C, generator=g).to(DEV)
    w1 = pack_mlp_weight(0.2 * torch.randn(2 * C, C, generator=g), parts=3).to(DEV)
    w2 = pack_mlp_weight(0.2 * torch.randn(C, 2 * C, generator=g), parts=3).to(DEV)
    depth_w = (0.5 + torch.rand(B, voxels, generator=g)).to(DEV)
    gamma = (1 + 0.1 * torch.randn(C, generator=g)).to(DEV)
    beta = (0.1 * torch.randn(C, generator=g)).to(DEV)
    rows = _lib.FfnoRows(B, voxels, C, nwg)
    return rows, u, w1, w2, depth_w, gamma, beta


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _outputs(B, voxels, C):
    return (torch.full((B * voxels, C), float("nan"), device=DEV), torch.full((B, 8, 3), float("nan"), device=DEV),
            torch.full((B, C), float("nan"), device=DEV), torch.full((B, C), float("nan"), device=DEV))


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _existing(rows, u, w1, w2, depth_w, gamma, beta):
    from ffnoml_amd import _lib
    y, stats, a, b = _outputs(rows.B, rows.voxels, rows.C)
    partials = torch.empty(rows.B, rows.nwg, 8, 3, device=DEV)
    _lib.call("ffno_post", rows, _p(u), _p(w1), _p(w2), _p(y), _lib.FfnoStatsOut(_p(depth_w), _p(partials)), _stream())
    _lib.call("ffno_gn_finalize", rows.B, rows.nwg, rows.C, _p(partials), _p(gamma), _p(beta), _p(stats), _p(a), _p(b),
              _stream())
    return y, stats, a, b


def _noalloc(rows, u, w1, w2, depth_w, gamma, beta):
    from ffnoml_amd import _lib
    y, stats, a, b = _outputs(rows.B, rows.voxels, rows.C)
    _lib.call("ffno_post_gn_noalloc", rows, _p(u), _p(w1), _p(w2), _p(y), _p(depth_w), _p(gamma), _p(beta), _p(stats),
              _p(a), _p(b), _stream())
    return y, stats, a, b


def test_noalloc_equals_the_workspace_entry_points_on_two_calls():
    for B, voxels, C, nwg, seed in ((3, 16 * 20 * 20, 64, 200, 0), (1, 8 * 8 * 9, 32, 7, 1)):
        print(f"[no_alloc] B={B} voxels={voxels} C={C} nwg={nwg}", flush=True)
        case = _case(B, voxels, C, nwg, seed)
        want = _existing(*case)
        got = _noalloc(*case)
        torch.cuda.synchronize()
        for name, w, g in zip(("y", "stats", "aff_a", "aff_b"), want, got):
            assert bool(torch.isfinite(w).all()), name
            assert torch.equal(w, g), name


def test_noalloc_rejects_more_slots_than_the_scratch_holds():
    from ffnoml_amd import _lib
    rows, u, w1, w2, depth_w, gamma, beta = _case(2, 4096, 16, 2049, 2)        # 4098 slots > 4096
    y, stats, a, b = _outputs(2, 4096, 16)
    with pytest.raises(_lib.FfnoError, match="static scratch"):
        _lib.call("ffno_post_gn_noalloc", rows, _p(u), _p(w1), _p(w2), _p(y), _p(depth_w), _p(gamma), _p(beta),
                  _p(stats), _p(a), _p(b), _stream())
