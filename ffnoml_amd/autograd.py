"""torch.autograd.Function around the HIP engine: forward runs the fused kernels; backward is the hand-written
adjoint (per-block recomputation from the saved block inputs), returning x_bar and z_bar (models/ffno_model.py:751-780
is the reference's checkpointed backward). z_bar is the per-column gradient of every column's own heights; on a shared
grid the reference's autograd instead sends the vertical transform's height sensitivity of all columns to the first
column's heights (models/ffno_model.py:270-276), so the two agree where the reference takes its per-column path."""
import torch


class FFNOFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, z, module, dx, dy):
        engine = module.engine()
        need = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        if not need:
            return engine.forward(x.detach(), z.detach(), dx, dy)
        y, saved = engine.forward_saved(x.detach(), z.detach(), dx, dy, want_z=ctx.needs_input_grad[1])
        ctx.save_for_backward(x, z)
        ctx.saved = saved
        ctx.module, ctx.dx, ctx.dy = module, dx, dy
        return y

    @staticmethod
    def backward(ctx, ybar):
        x, z = ctx.saved_tensors
        engine = ctx.module.engine()
        xbar, zbar = engine.vjp(x, z, ctx.dx, ctx.dy, ybar.contiguous(), ctx.saved, want_z=ctx.needs_input_grad[1])
        ctx.saved = None
        return xbar if ctx.needs_input_grad[0] else None, zbar, None, None, None


def _spacing(v):
    """A spacing as a Python float (models/ffno_model.py:11-18), read once on the host."""
    if isinstance(v, (float, int)):
        return float(v)
    if torch.is_tensor(v):
        if v.numel() != 1:
            raise ValueError("spacing must be scalar.")
        return float(v.detach().item())
    raise TypeError(f"Unsupported spacing type: {type(v)}")


def ffno_apply(module, x, z_scale, dx, dy):
    """FFNO3D.forward(x, z_scale, dx, dy) through the HIP engine."""
    if not x.is_cuda:
        raise RuntimeError("ffnoml_amd.FFNO3D runs on the GPU through libffno_hip.so; it has no CPU path")
    if z_scale.dim() == 5:
        z_scale = z_scale[:, 0]
    return FFNOFunction.apply(x, z_scale, module, _spacing(dx), _spacing(dy))
