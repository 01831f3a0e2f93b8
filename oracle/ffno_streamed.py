"""A streamed float64 forward of FFNO3D for volumes whose float64 intermediates do not all fit host RAM at once.

TEST INFRASTRUCTURE ONLY, like oracle/ffno_oracle.py, whose functions it composes: it is the accuracy checker of the
whole-volume shapes (tests/golden/make_workload_fixtures.py), and nothing in the product imports it.

The reference streams a whole cube through its layers the same way (ffno_batched_backend.py:159-199, :253-289): every
operation of FFNO3D is either local or factorises, so it runs over chunks without changing the operator --
  1x1 layers, gates, the fusion MLP and the head   any chunk of voxels (here: depth planes)
  the horizontal spectral convolution                depth planes (each (b, d) plane is transformed on its own)
  the vertical spectral convolution                  rows of columns (each column is transformed on its own; the
                                                     periodic box is the whole volume's)
  the depth-weighted GroupNorm                       statistics over the volume, accumulated over chunks, then
                                                     applied chunk by chunk
Unlike the reference's single-pass sums, the GroupNorm here keeps the model's two-pass form (weighted mean, then the
weighted mean square about it; models/ffno_model.py:77-95), so that it equals ffno_oracle.weighted_group_norm up to the
order of the float64 sums. At most four full-size hidden tensors are alive at a time.
"""
import torch

from . import ffno_oracle as O


def _planes(D, chunk):
    return [(d0, min(D, d0 + chunk)) for d0 in range(0, D, chunk)]


def group_norm_(t, w, weight, bias, groups, planes, act=None, eps=1e-5):
    """In place: t [B, C, D, H, W] <- act(weighted GroupNorm(t)) with statistics over the whole volume
    (models/ffno_model.py:77-95), accumulated in two passes over depth chunks."""
    B, C, D = t.shape[:3]
    cpg = C // groups

    def grouped(a, b):
        return t[:, :, a:b].reshape(B, groups, cpg, b - a, *t.shape[3:])

    num = torch.zeros(B, groups, dtype=t.dtype)
    total = torch.zeros(B, groups, dtype=t.dtype)
    for a, b in planes:
        g = grouped(a, b)
        wg = w[:, :, a:b].unsqueeze(1)                         # [B, 1, 1, d, H, W]
        num += (wg * g).sum(dim=(2, 3, 4, 5))
        total += wg.expand_as(g).sum(dim=(2, 3, 4, 5))
    mean = num / total
    sq = torch.zeros(B, groups, dtype=t.dtype)
    for a, b in planes:
        g = grouped(a, b)
        wg = w[:, :, a:b].unsqueeze(1)
        sq += (wg * (g - mean[:, :, None, None, None, None]) ** 2).sum(dim=(2, 3, 4, 5))
    rstd = torch.rsqrt(sq / total + eps)
    shape = (1, C, 1, 1, 1)
    for a, b in planes:
        g = grouped(a, b)
        y = ((g - mean[:, :, None, None, None, None]) * rstd[:, :, None, None, None, None]).reshape(
            B, C, b - a, *t.shape[3:])
        y = y * weight.view(shape) + bias.view(shape)
        t[:, :, a:b] = act(y) if act is not None else y
    return t


def block(sd, i, x, z_scale, dx, dy, cfg, planes, rows, span):
    """O.block over chunks; x [B, C, D, H, W] is replaced by the block's output (returned)."""
    p = f"blocks.{i}."
    G = cfg["groups"]
    w = O.depth_weights(z_scale)
    z5 = z_scale.unsqueeze(1)
    H = x.shape[3]

    def norm_(name, t, act=None):
        return group_norm_(t, w, sd[p + name + ".weight"], sd[p + name + ".bias"], G, planes, act)

    spec = torch.empty_like(x)
    for a, b in planes:
        spec[:, :, a:b] = O.horizontal_branch(sd, p + "spec.", x[:, :, a:b], dx, dy, cfg)
    norm_("norm_spec", spec, O.gelu)
    vert = torch.empty_like(x)
    for r0 in range(0, H, rows):
        r1 = min(H, r0 + rows)
        vert[:, :, :, r0:r1] = O.vertical_branch(sd, p + "vertical.", x[:, :, :, r0:r1], z5[:, :, :, r0:r1], cfg,
                                                 span=span)
    norm_("norm_vert", vert, O.gelu)
    fused = torch.empty_like(x)
    for a, b in planes:
        fused[:, :, a:b] = O.fuse_branches(sd, p, x[:, :, a:b], spec[:, :, a:b], vert[:, :, a:b])
    del spec, vert
    norm_("norm_fuse", fused, O.gelu)
    for a, b in planes:                                          # x <- x1 = x + res_fused * fused
        x[:, :, a:b] += sd[p + "res_fused"] * fused[:, :, a:b]
    for a, b in planes:                                          # fused <- pw(x1)
        fused[:, :, a:b] = O.pw_mlp(sd, p, x[:, :, a:b])
    norm_("norm_pw", fused)
    for a, b in planes:                                          # x <- x2 = x1 + res_pw * GN(pw)
        x[:, :, a:b] += sd[p + "res_pw"] * fused[:, :, a:b]
    for a, b in planes:                                          # fused <- mlp(x2)
        fused[:, :, a:b] = O.point_mlp(sd, p, x[:, :, a:b])
    norm_("norm_mlp", fused)
    for a, b in planes:                                          # x <- x2 + res_mlp * GN(mlp)
        x[:, :, a:b] += sd[p + "res_mlp"] * fused[:, :, a:b]
    return x


@torch.no_grad()
def forward(sd, cfg, x, z_scale, dx, dy, plane_chunk=8, row_chunk=16, log=None):
    """O.forward (models/ffno_model.py:800-832) with bounded memory: the same normalised log10 n [B, Cout, D, H, W]."""
    dx, dy = float(dx), float(dy)
    B, _, D, H, W = x.shape
    planes = _planes(D, plane_chunk)
    span = float((z_scale.amax(dim=-3) - z_scale.amin(dim=-3)).max())
    w = O.depth_weights(z_scale)
    h = torch.empty(B, cfg["width"], D, H, W, dtype=x.dtype)
    for a, b in planes:
        h[:, :, a:b] = O.conv1x1(x[:, :, a:b], sd["lift.0.weight"])
    group_norm_(h, w, sd["lift.1.weight"], sd["lift.1.bias"], cfg["groups"], planes, O.gelu)
    for i in range(cfg["n_layers"]):
        h = block(sd, i, h, z_scale, dx, dy, cfg, planes, row_chunk, span)
        if log:
            log(f"block {i + 1}/{cfg['n_layers']} done")
    y = torch.empty(B, cfg["out_channels"], D, H, W, dtype=x.dtype)
    for a, b in planes:
        head = O.conv1x1(O.gelu(O.conv1x1(h[:, :, a:b], sd["proj1.weight"])), sd["proj2.weight"], sd["proj2.bias"])
        y[:, :, a:b] = O.head_output(sd, head, x[:, :, a:b], cfg)
    return y
