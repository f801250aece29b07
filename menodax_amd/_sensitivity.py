"""Forward sensitivities: which directions a joint solve carries, and the
joint system's layout.

Restates modax/_sensitivity.py:64-135 (``SensitivitySpec``) and 258-300
(``augmented_y0``, ``augmented_error_weights``, ``split_augmented``). Writing
``S(t) = dy(t)/dtheta`` for a block of directions ``theta``, the variational
equation ``S' = J_y S + J_p`` is integrated with the state as one joint system
of ``n_vars * (1 + n_sens)`` components::

    z[j]                        = y[j]
    z[n_vars + k * n_vars + r]  = S[r, k]

directions running over the initial-state block first (``n_vars`` of them,
when ``y0`` is differentiated) and then over the chosen parameter columns.
``S(t0)`` is the identity for the initial-state block and zero for the
parameter block. The kernel integrates it with the Jacobian and LU of the state
block alone; the sensitivity rows' right-hand side, their df/dt and their stage
coupling come from forward-mode sweeps of the callback (``_codegen``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class SensitivitySpec:
    """Which sensitivity blocks a joint solve carries (hashable: part of the
    kernel key). ``error_control=False`` gives the sensitivity rows zero error
    weight and divides the norm by ``n_vars``, so the joint solve takes exactly
    the plain solve's steps."""

    n_vars: int
    n_params: int
    wrt_y0: bool
    wrt_params: bool
    error_control: bool = True
    param_columns: tuple | None = None

    def __post_init__(self):
        if not (self.wrt_y0 or self.wrt_params):
            raise ValueError("a sensitivity spec must carry at least one direction block")
        if self.param_columns is not None:
            cols = tuple(int(c) for c in self.param_columns)
            object.__setattr__(self, "param_columns", cols)
            if len(set(cols)) != len(cols):
                raise ValueError("param_columns must not repeat a column")
            bad = [c for c in cols if not 0 <= c < self.n_params]
            if bad:
                raise ValueError(
                    f"param_columns {bad} lie outside the {self.n_params} parameter columns")

    @property
    def n_y0_dirs(self) -> int:
        return self.n_vars if self.wrt_y0 else 0

    @property
    def n_param_dirs(self) -> int:
        if not self.wrt_params:
            return 0
        return self.n_params if self.param_columns is None else len(self.param_columns)

    @property
    def param_seed_columns(self) -> tuple:
        return tuple(range(self.n_params)) if self.param_columns is None else self.param_columns

    @property
    def n_sens(self) -> int:
        return self.n_y0_dirs + self.n_param_dirs

    @property
    def n_aug(self) -> int:
        return self.n_vars * (1 + self.n_sens)

    @property
    def n_error(self) -> int:
        return self.n_aug if self.error_control else self.n_vars

    @property
    def direction_columns(self) -> tuple:
        """Per direction, the parameter it seeds, or -1 for an initial-state one."""
        cols = (-1,) * self.n_y0_dirs
        if self.wrt_params:
            cols += self.param_seed_columns
        return cols


def augmented_y0(y0, spec: SensitivitySpec):
    """``(N, n_vars)`` -> ``(N, n_aug)``: identity block for the initial-state
    directions, zeros for the parameter directions (modax/_sensitivity.py:258-272)."""
    import torch

    n = y0.shape[0]
    blocks = [y0]
    if spec.wrt_y0:
        eye = torch.eye(spec.n_vars, dtype=y0.dtype, device=y0.device).reshape(1, -1)
        blocks.append(eye.expand(n, -1))
    if spec.wrt_params:
        blocks.append(torch.zeros((n, spec.n_vars * spec.n_param_dirs), dtype=y0.dtype,
                                  device=y0.device))
    return torch.cat(blocks, dim=1)


def augmented_error_weights(weights, spec: SensitivitySpec):
    """``(N, n_vars)`` -> ``(N, n_aug)``: ones for the sensitivity rows, or zeros
    when they are kept out of step-size control (modax/_sensitivity.py:275-293)."""
    import torch

    fill = torch.ones if spec.error_control else torch.zeros
    tail = fill((weights.shape[0], spec.n_aug - spec.n_vars), dtype=weights.dtype,
                device=weights.device)
    return torch.cat([weights, tail], dim=1)


def split_augmented(hist, spec: SensitivitySpec):
    """``(N, n_save, n_aug)`` -> state ``(N, n_save, n_vars)`` and
    ``S`` ``(N, n_save, n_vars, n_sens)`` (modax/_sensitivity.py:296-300)."""
    n, n_save, _ = hist.shape
    state = hist[:, :, : spec.n_vars]
    sens = hist[:, :, spec.n_vars:].reshape(n, n_save, spec.n_sens, spec.n_vars)
    if isinstance(sens, np.ndarray):
        return state, np.swapaxes(sens, 2, 3)
    return state, sens.transpose(2, 3)
