"""Batched not-a-knot cubic splines on the GPU (the reference's
``utils/interpolate.py``, ``CubicSplineInterpolant`` :42-257)."""

from __future__ import annotations

from .arrays import to_dev
from .parallelbase import BBHxParallelModule


class CubicSplineInterpolant(BBHxParallelModule):
    """Not-a-knot cubic splines of ``y_all`` (num_interp_params, num_bin_all,
    num_modes, length) over ``x`` (num_bin_all, num_modes, length), or their
    flattened forms with the dimensions given. The coefficients satisfy
    y(x) = y_i + c1_i dx + c2_i dx^2 + c3_i dx^3 on [x_i, x_{i+1}]."""

    def __init__(self, x, y_all, num_interp_params=None, num_bin_all=None, num_modes=None, length=None,
                 force_backend=None):
        super().__init__(force_backend=force_backend)
        torch = self.xp
        x = to_dev(x, self.device)
        y_all = to_dev(y_all, self.device)
        num_interp_params, num_bin_all, num_modes, length = self.spline_dims(
            x.ndim, x.numel(), tuple(x.shape), y_all.ndim, y_all.numel(), tuple(y_all.shape), num_interp_params,
            num_bin_all, num_modes, length)
        x = x.flatten()
        y_all = y_all.flatten()
        ninterps = num_modes * num_interp_params * num_bin_all
        self.degree = 3
        self.length = length
        self.reshape_shape = (num_interp_params, num_bin_all, num_modes, length)
        self.x_reshape_shape = (num_bin_all, num_modes, length)
        B = torch.zeros(ninterps * length, dtype=torch.float64, device=self.device)
        self.c1 = upper_diag = torch.zeros_like(B)
        self.c2 = diag = torch.zeros_like(B)
        self.c3 = lower_diag = torch.zeros_like(B)
        self.y = y_all
        self.B = B
        self.interpolate_arrays(x, y_all, B, upper_diag, diag, lower_diag, length, num_interp_params, num_modes,
                                num_bin_all)
        self.x = x.clone()

    @staticmethod
    def spline_dims(x_ndim, x_numel, x_shape, y_ndim, y_numel, y_shape, num_interp_params=None, num_bin_all=None,
                    num_modes=None, length=None):
        """(num_interp_params, num_bin_all, num_modes, length) of the arrays, with
        the reference's shape rules (interpolate.py:93-126): both flat with the
        four sizes given, or x (B, M, L) and y_all (P, B, M, L); at least 3 knots."""
        if x_ndim == 1 or y_ndim == 1:
            if x_ndim != 1 or y_ndim != 1:
                raise ValueError("If providing flattened x and y_all, need to both be flattened.")
            if length is None or num_modes is None or num_bin_all is None or num_interp_params is None:
                raise ValueError("If providing flattened arrays, need to provide dimensional information: "
                                 "length, num_modes, num_bin_all, num_interp_params.")
            if x_numel != length * num_modes * num_bin_all:
                raise ValueError(f"Length of the x array is not correct. It is supposed to be "
                                 f"{length * num_modes * num_bin_all}. It is currently {x_numel}.")
            if y_numel != length * num_modes * num_bin_all * num_interp_params:
                raise ValueError(f"Length of the y_all array is not correct. It is supposed to be "
                                 f"{length * num_modes * num_bin_all * num_interp_params}. It is currently {y_numel}.")
        else:
            if x_ndim != 3:
                raise ValueError("If providing shaped array, needs to be 3D with shape (num_bin_all, num_modes, length).")
            if y_ndim != 4:
                raise ValueError("If providing shaped array, needs to be 4D with shape "
                                 "(num_interp_params, num_bin_all, num_modes, length).")
            if tuple(y_shape[1:]) != tuple(x_shape):
                raise ValueError(f"y_all shape {tuple(y_shape)} does not extend x shape {tuple(x_shape)}.")
            num_interp_params, num_bin_all, num_modes, length = (int(v) for v in y_shape)
        if length < 3:
            raise ValueError("splines need at least 3 knots.")
        return int(num_interp_params), int(num_bin_all), int(num_modes), int(length)

    @property
    def interpolate_arrays(self):
        return self.backend.interpolate_wrap

    @property
    def x_shaped(self):
        return self.x.reshape(self.x_reshape_shape)

    @property
    def y_shaped(self):
        return self.y.reshape(self.reshape_shape)

    @property
    def c1_shaped(self):
        return self.c1.reshape(self.reshape_shape)

    @property
    def c2_shaped(self):
        return self.c2.reshape(self.reshape_shape)

    @property
    def c3_shaped(self):
        return self.c3.reshape(self.reshape_shape)

    @property
    def container(self):
        return [self.x, self.y, self.c1, self.c2, self.c3]

    def __call__(self, x_new):
        """Evaluate all splines at x_new (num_bin_all, num_modes, n) (interpolate.py:213-257)."""
        torch = self.xp
        x_new = to_dev(x_new, self.device)
        if x_new.ndim == 3 and tuple(x_new.shape[:2]) != self.x_reshape_shape[:2]:
            raise ValueError("x_new must have 3D shape of (num_bin_all, num_modes, # of new points)")
        xs = self.x_shaped
        assert bool(torch.all(x_new <= xs.max(dim=-1).values[:, :, None])) and bool(
            torch.all(x_new >= xs.min(dim=-1).values[:, :, None]))
        seg = torch.searchsorted(xs.reshape(-1, self.length).contiguous(),
                                 x_new.reshape(-1, x_new.shape[-1]).contiguous(), right=True) - 1
        seg = seg.reshape(x_new.shape)
        seg = torch.where(seg == self.length - 1, torch.full_like(seg, self.length - 2), seg)
        x0 = torch.take_along_dim(xs, seg, dim=-1)
        idx = seg[None].expand(self.reshape_shape[0], -1, -1, -1)
        y0 = torch.take_along_dim(self.y_shaped, idx, dim=-1)
        c1 = torch.take_along_dim(self.c1_shaped, idx, dim=-1)
        c2 = torch.take_along_dim(self.c2_shaped, idx, dim=-1)
        c3 = torch.take_along_dim(self.c3_shaped, idx, dim=-1)
        dx = x_new - x0
        return y0 + c1 * dx + c2 * dx**2 + c3 * dx**3
