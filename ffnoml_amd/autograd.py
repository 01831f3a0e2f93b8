"""torch.autograd.Function around the HIP engine: forward runs the fused kernels; backward is the hand-written
adjoint (per-block recomputation from the saved block inputs), returning x_bar and z_bar (models/ffno_model.py:751-780
is the reference's checkpointed backward). z_bar is the per-column gradient of every column's own heights; on a shared
grid the reference's autograd instead sends the vertical transform's height sensitivity of all columns to the first
column's heights (models/ffno_model.py:270-276), so the two agree where the reference takes its per-column path."""
import contextlib
import threading

import torch

_STATE = threading.local()     # per calling thread: whether the call in progress will be differentiated


def _plain(t):
    """The tensor under torch.func's wrappers (a transform hands a custom Function's jvp wrapped tangents)."""
    f = getattr(torch._C, "_functorch", None)
    while f is not None and t is not None and f.is_functorch_wrapped_tensor(t):
        t = f.get_unwrapped(t)
    return t


def _engine_mode():
    """A context in which the engine's own tensor work is not lifted into torch.func's wrappers: inside a custom
    Function's jvp under torch.func.jvp, even fresh tensors come out wrapped and have no data pointer to pass to the
    HIP library."""
    disable = getattr(torch._C, "_DisableFuncTorch", None)
    return disable() if disable is not None else contextlib.nullcontext()


class _Info:
    """What forward hands to setup_context: the inputs as forward saw them and the adjoint's saved state. A plain
    object, not a dict or tuple, so that torch.func does not wrap the tensors inside as outputs of the Function."""

    def __init__(self, x, z):
        self.x, self.z, self.saved = x, z, None


class FFNOFunction(torch.autograd.Function):
    """forward returns (y, info): y the normalised log10 n; info (_Info) the inputs as forward saw them and what the
    adjoint needs (None when no input requires a gradient). setup_context keeps info, so the same Function serves
    autograd (backward), torch.func transforms and forward-mode duals (jvp)."""

    @staticmethod
    def forward(x, z, module, dx, dy):
        engine = module.engine()
        x, z = _plain(x).detach(), _plain(z).detach()
        info = _Info(x, z)
        need = getattr(_STATE, "needs_grad", False)
        with _engine_mode():
            if not need:
                return engine.forward(x, z, dx, dy), info
            y, info.saved = engine.forward_saved(x, z, dx, dy, want_z=need == "z")
        return y, info

    @staticmethod
    def setup_context(ctx, inputs, output):
        _, _, module, dx, dy = inputs
        ctx.info = output[1]
        ctx.module, ctx.dx, ctx.dy = module, dx, dy
        ctx.set_materialize_grads(False)

    @staticmethod
    def jvp(ctx, x_dot, z_dot, *_):
        """Forward-mode derivative along x_dot (the hand-written tangent kernels); a tangent of the heights is not
        supported (the reference's jvp_local perturbs the features only, ffno_fsdp_runtime.py:350-372)."""
        z_dot = _plain(z_dot)
        if z_dot is not None and bool(torch.any(z_dot != 0)):
            raise NotImplementedError("FFNO3D JVP along the heights is not implemented; give a feature tangent")
        x, z = ctx.info.x, ctx.info.z
        if x_dot is None:
            return torch.zeros(x.shape[0], ctx.module.out_channels, *x.shape[2:], device=x.device), None
        with _engine_mode():
            _, y_dot = ctx.module.engine().jvp(x, z, ctx.dx, ctx.dy, _plain(x_dot))
        return y_dot, None

    @staticmethod
    def backward(ctx, ybar, _info_bar=None):
        info = ctx.info
        engine = ctx.module.engine()
        with _engine_mode():
            if info.saved is None:          # a transform asked for the adjoint of a forward that kept nothing
                _, info.saved = engine.forward_saved(info.x, info.z, ctx.dx, ctx.dy, want_z=ctx.needs_input_grad[1])
            xbar, zbar = engine.vjp(info.x, info.z, ctx.dx, ctx.dy, _plain(ybar).contiguous(), info.saved,
                                    want_z=ctx.needs_input_grad[1])
        info.saved = None
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
    # whether the adjoint will be asked for (forward then keeps what it needs): decided here, where the autograd state
    # of the caller's tensors is visible
    grad = torch.is_grad_enabled() and (x.requires_grad or z_scale.requires_grad)
    _STATE.needs_grad = ("z" if z_scale.requires_grad else "x") if grad else False
    try:
        return FFNOFunction.apply(x, z_scale, module, _spacing(dx), _spacing(dy))[0]
    finally:
        _STATE.needs_grad = False


class HeadOutputFunction(torch.autograd.Function):
    """FFNO3D.head_output through the HIP library: forward ffno_head_output, backward ffno_head_output_bwd."""

    @staticmethod
    def forward(ctx, head, x_in, module):
        head, x_in = head.detach().contiguous().float(), x_in.detach().contiguous().float()
        ctx.save_for_backward(head, x_in)
        ctx.module = module
        return module.engine().head_output(head, x_in)

    @staticmethod
    def backward(ctx, ybar):
        head, x_in = ctx.saved_tensors
        head_bar, x_bar = ctx.module.engine().head_output_bwd(head, x_in, ybar.contiguous().float())
        return (head_bar if ctx.needs_input_grad[0] else None), (x_bar if ctx.needs_input_grad[1] else None), None


def head_output_apply(module, head, x_in):
    if not (head.is_cuda and x_in.is_cuda):
        raise RuntimeError("ffnoml_amd.FFNO3D.head_output runs on the GPU through libffno_hip.so; it has no CPU path")
    return HeadOutputFunction.apply(head, x_in, module)
