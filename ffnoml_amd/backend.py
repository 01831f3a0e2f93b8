"""Single-GPU backend with the reference's predict/VJP contract (ffno_fsdp_runtime.py:81-372), running the whole
volume on one MI355X through the HIP engine instead of H slabs over FSDP ranks.

Inputs are SI, as the Julia client sends them: features (6, nz, nx, ny) = (T [K], vx, vy, vz [m/s], log10 ne,
log10 rho), z (nz, nx, ny) [m], dx, dy [m]; outputs are linear populations (nz, nx, ny, levels) [m^-3] clamped to
10^(+-37). Numpy in, fresh numpy out, synchronous. Errors: ValueError for bad requests, RuntimeError for non-finite
or non-positive results, as the reference raises them.
"""
import hashlib
import importlib
import math
from pathlib import Path

import numpy as np
import torch

from .factory import get_checkpoint_io_metadata, get_checkpoint_normalization, load_checkpoint

INPUT_CHANNELS = ("temperature", "vx", "vy", "vz", "log10_ne", "log10_rho")
Z_METRES_PER_MODEL_UNIT = 1.0e6                 # ffno_fsdp_runtime.py:43-45
# metres -> Mm as one float32 multiply, which rounds alike on the host and on the device (the device's division by a
# scalar multiplies by its reciprocal, the host's divides): a plan built from heights in host memory then sees the
# very heights the device computes with
MM_PER_METRE = 1.0 / Z_METRES_PER_MODEL_UNIT
VELOCITY_M_S_PER_MODEL_UNIT = 1.0e3
LOG10_POPULATION_LIMIT = 37.0
LN10 = math.log(10.0)


class GPUBackend:
    """One model resident on one GPU (ffno_fsdp_runtime.py:81-221 for the construction it replaces)."""

    def __init__(self, checkpoint_path, factory_module="ffnoml_amd.factory", factory_name="create_ffno3d",
                 level_names=(), factory_kwargs=None, device="cuda:0", name="model", model=None, normalization=None):
        self.device = torch.device(device)
        self.name = name
        if model is None:
            self.checkpoint_path = str(Path(checkpoint_path).resolve())
            ckpt = load_checkpoint(self.checkpoint_path, map_location="cpu")
            normalization = get_checkpoint_normalization(ckpt)
            io = get_checkpoint_io_metadata(ckpt)
            if normalization is None or io is None:
                raise RuntimeError("checkpoint must contain normalization_stats and io_metadata")
            if io["Cin"] != len(INPUT_CHANNELS):
                raise RuntimeError(f"checkpoint Cin={io['Cin']} does not match {INPUT_CHANNELS}")
            self.level_names = tuple(str(v) for v in level_names)
            if io["Cout"] != len(self.level_names):
                raise RuntimeError("checkpoint Cout differs from configured level_names")
            factory = getattr(importlib.import_module(factory_module), factory_name)
            model = factory(Cin=io["Cin"], Cout=io["Cout"], device="cpu", **dict(factory_kwargs or {}))
            model.load_state_dict(ckpt["model_state"], strict=True)
            self.checkpoint_hash = hashlib.sha256(Path(self.checkpoint_path).read_bytes()).hexdigest()
        else:
            self.checkpoint_path = None
            self.checkpoint_hash = None
            self.level_names = tuple(str(v) for v in level_names) or tuple(f"level{i}" for i in
                                                                           range(model.out_channels))
            if normalization is None:
                normalization = {k: getattr(model, k).detach().cpu().numpy() for k in
                                 ("mean_X", "std_X", "mean_Y", "std_Y")}
        self.model = model.eval().to(self.device)
        self.parameter_count = sum(p.numel() for p in self.model.parameters())

        def norm(key):
            return torch.as_tensor(np.asarray(normalization[key]), dtype=torch.float32,
                                   device=self.device)[None, :, None, None, None]

        self.mean_x, self.std_x = norm("mean_X"), norm("std_X")
        self.mean_y, self.std_y = norm("mean_Y"), norm("std_Y")
        if torch.any(self.std_x <= 0) or torch.any(self.std_y <= 0):
            raise RuntimeError("checkpoint normalization standard deviations must be positive")
        self.call_count = 0

    def describe(self):
        return {"backend": "single_gpu_whole_volume", "input_channels": INPUT_CHANNELS,
                "level_names": self.level_names, "checkpoint_hash": self.checkpoint_hash, "device": str(self.device),
                "world_size": 1, "parameter_count": self.parameter_count,
                "output_representation": "linear_population_m3"}

    # ---------------------------------------------------------------------------------------------- requests
    def inputs(self, features, z_scale):
        """Validated device tensors [1, 6, nz, nx, ny] and [1, nz, nx, ny] (ffno_fsdp_runtime.py:239-263)."""
        features = np.asarray(features, dtype=np.float32)
        z_scale = np.asarray(z_scale, dtype=np.float32)
        if features.ndim != 4 or features.shape[0] != len(INPUT_CHANNELS):
            raise ValueError("features must have shape (6,nz,nx,ny)")
        if tuple(features.shape[1:]) != tuple(z_scale.shape):
            raise ValueError("z_scale must have shape (nz,nx,ny)")
        if not np.isfinite(features).all() or not np.isfinite(z_scale).all():
            raise ValueError("FFNO request contains NaN or Inf")
        if not (features[0] > 0).all():
            raise ValueError("FFNO request temperature must be positive")
        x = torch.from_numpy(np.ascontiguousarray(features)).to(self.device)[None]
        z = torch.from_numpy(np.ascontiguousarray(z_scale)).to(self.device)[None]
        return x, z

    def model_inputs(self, x_si, z_si):
        """SI -> normalised training features and z in Mm (ffno_fsdp_runtime.py:48-58)."""
        f = torch.cat((torch.log10(x_si[:, 0:1]), x_si[:, 1:4] / VELOCITY_M_S_PER_MODEL_UNIT, x_si[:, 4:6]), dim=1)
        return (f - self.mean_x) / self.std_x, z_si * MM_PER_METRE

    def log_populations(self, y):
        return y * self.std_y + self.mean_y

    @staticmethod
    def linear(log10_n):
        """10^log10 n within float32 range (ffno_fsdp_runtime.py:76-78)."""
        return torch.pow(10.0, log10_n.clamp(-LOG10_POPULATION_LIMIT, LOG10_POPULATION_LIMIT))

    @staticmethod
    def canonical(populations):
        """[1, L, nz, nx, ny] -> numpy (nz, nx, ny, L) (ffno_fsdp_runtime.py:283-292)."""
        return populations[0].permute(1, 2, 3, 0).float().detach().cpu().numpy()

    def predict_tensor(self, x_si, z_si, dx, dy, plan=None):
        """Linear populations [B, L, nz, nx, ny] on the device. plan: a plan built beforehand (for graph capture,
        where nothing may synchronise with the host)."""
        xm, zm = self.model_inputs(x_si, z_si)
        with torch.no_grad():
            y = self.model.engine().forward(xm, zm, float(dx), float(dy), plan=plan)
            return self.linear(self.log_populations(y))

    def plan_for(self, z_si, dx, dy):
        """The engine's plan for SI heights z_si [B, nz, nx, ny] (m). Heights in host memory build it without
        waiting for the device; on the device they are read back first."""
        B, D, H, W = z_si.shape
        eng = self.model.engine()
        eng.sync(self.device)
        return eng.plan(B, D, H, W, z_si * MM_PER_METRE, float(dx), float(dy))

    def predict_local(self, features, z_scale, dx, dy):
        x, z = self.inputs(features, z_scale)
        populations = self.predict_tensor(x, z, dx, dy)
        if not torch.isfinite(populations).all() or torch.any(populations <= 0):
            raise RuntimeError("FFNO returned non-positive or non-finite populations")
        torch.cuda.synchronize(self.device)
        self.call_count += 1
        return self.canonical(populations)

    def vjp_tensor(self, x_si, z_si, dx, dy, cotangent, stride=None, plan=None, keep=0):
        """(populations, feature_bar [B,6,...] SI, z_bar [B,...] SI) of <cotangent, populations(x_si, z_si)>;
        cotangent [B, L, nz, nx, ny]. The SI chain rule around the engine's adjoint is written out here: through the
        clamp and 10^, the output normalisation, the input normalisation and the log10 T / km/s / Mm conversions.
        stride: the adjoint's checkpoint stride (None: the engine's choice for the HBM free now); keep: with stride
        1, the last `keep` blocks kept whole; plan: one built beforehand (plan_for), so that nothing here waits for
        the device."""
        xm, zm = self.model_inputs(x_si, z_si)
        engine = self.model.engine()
        y, saved = engine.forward_saved(xm, zm, float(dx), float(dy), plan=plan, stride=stride, keep=keep)
        log_n = self.log_populations(y)
        n = self.linear(log_n)
        inside = (log_n >= -LOG10_POPULATION_LIMIT) & (log_n <= LOG10_POPULATION_LIMIT)
        ybar = cotangent * n * LN10 * inside * self.std_y
        del y, log_n, inside
        xbar, zbar = engine.vjp(xm, zm, float(dx), float(dy), ybar, saved, want_z=True, plan=plan)
        del saved
        fbar = xbar / self.std_x
        feature_bar = torch.cat((fbar[:, 0:1] / (x_si[:, 0:1] * LN10), fbar[:, 1:4] / VELOCITY_M_S_PER_MODEL_UNIT,
                                 fbar[:, 4:6]), dim=1)
        return n, feature_bar, zbar / Z_METRES_PER_MODEL_UNIT

    def vjp_local(self, features, z_scale, dx, dy, population_cotangent):
        """ffno_fsdp_runtime.py:306-348: {"features": (6,nz,nx,ny), "z_scale": (nz,nx,ny)}."""
        x, z = self.inputs(features, z_scale)
        ct = np.asarray(population_cotangent, dtype=np.float32)
        expected = tuple(np.asarray(z_scale).shape) + (len(self.level_names),)
        if tuple(ct.shape) != expected:
            raise ValueError(f"population_cotangent has shape {ct.shape}; expected {expected}")
        if not np.isfinite(ct).all():
            raise ValueError("FFNO VJP cotangent contains NaN or Inf")
        cot = torch.from_numpy(np.ascontiguousarray(ct)).to(self.device).permute(3, 0, 1, 2)[None]
        _, fb, zb = self.vjp_tensor(x, z, dx, dy, cot)
        if not torch.isfinite(fb).all() or not torch.isfinite(zb).all():
            raise RuntimeError("FFNO VJP returned NaN or Inf")
        torch.cuda.synchronize(self.device)
        self.call_count += 1
        return {"features": fb[0].float().cpu().numpy(), "z_scale": zb[0].float().cpu().numpy()}

    def jvp_tensor(self, x_si, z_si, dx, dy, tangent_si, plan=None):
        """(populations, populations_dot) [B, L, nz, nx, ny] along a tangent of the SI features [B, 6, nz, nx, ny]:
        the SI chain written out around the engine's tangent pass -- through log10 T (T_dot / (T ln 10)), km/s, the
        input normalisation, then the output normalisation, 10^ and its clamp (zero slope outside +-37)."""
        xm, zm = self.model_inputs(x_si, z_si)
        fd = torch.cat((tangent_si[:, 0:1] / (x_si[:, 0:1] * LN10), tangent_si[:, 1:4] / VELOCITY_M_S_PER_MODEL_UNIT,
                        tangent_si[:, 4:6]), dim=1) / self.std_x
        y, yd = self.model.engine().jvp(xm, zm, float(dx), float(dy), fd, plan=plan)
        log_n = self.log_populations(y)
        n = self.linear(log_n)
        inside = (log_n >= -LOG10_POPULATION_LIMIT) & (log_n <= LOG10_POPULATION_LIMIT)
        return n, n * LN10 * inside * self.std_y * yd

    def jvp_local(self, features, z_scale, dx, dy, feature_tangent):
        """Forward-mode derivative of the linear populations along feature_tangent (6, nz, nx, ny), SI units
        (ffno_fsdp_runtime.py:350-372): (nz, nx, ny, L)."""
        x, z = self.inputs(features, z_scale)
        td = np.asarray(feature_tangent, dtype=np.float32)
        if tuple(td.shape) != tuple(np.asarray(features).shape):
            raise ValueError(f"feature_tangent has shape {td.shape}; expected {np.asarray(features).shape}")
        if not np.isfinite(td).all():
            raise ValueError("FFNO JVP tangent contains NaN or Inf")
        tangent = torch.from_numpy(np.ascontiguousarray(td)).to(self.device)[None]
        _, nd = self.jvp_tensor(x, z, dx, dy, tangent)
        if not torch.isfinite(nd).all():
            raise RuntimeError("FFNO JVP returned NaN or Inf")
        torch.cuda.synchronize(self.device)
        self.call_count += 1
        return self.canonical(nd)
