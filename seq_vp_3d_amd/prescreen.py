"""T75 floating-point Bernstein prescreen on the device (certificate_prescreen.py:125-153).

prescreen_lower_bounds keeps the reference signature (u, grid, coefficients, cells) and returns the
same per-cell values bit for bit; prescreen_flags applies the certificate's proof rule
(feasibility.py:665-695) to every cell at once and counts the cells left for the authoritative
interval certificate on the CPU. The GPU never declares a cell infeasible.
"""
import numpy as np
import torch

from . import _lib
from .problem import InputValidationError


def _table(coefficients, device):
    if isinstance(coefficients, torch.Tensor):          # a device table already checked by device_table()
        return coefficients, True
    D = np.ascontiguousarray(np.asarray(coefficients, dtype=np.float64))
    if D.shape != (3, 27, 8):
        raise InputValidationError("prescreen: closed27 coefficient table (3, 27, 8) required")
    nonzero = np.abs(D[D != 0])
    usable = np.all(np.isfinite(D)) and not (nonzero.size and nonzero.min() < np.finfo(np.float64).tiny)
    return torch.tensor(D, device=device), usable


def prescreen_lower_bounds(u, grid, coefficients, cells=None):
    """min beta_fl - B for each listed cell (all cells when cells is None); NaN where not finite,
    and NaN everywhere when the table has a subnormal or nonfinite entry (no proofs then)."""
    if not isinstance(u, torch.Tensor) or u.dtype != torch.float64 or tuple(u.shape) != tuple(grid.shape) + (3,):
        raise InputValidationError("prescreen: u must be a float64 device (nz, ny, nx, 3) field")
    nz, ny, nx = grid.shape
    D, usable = _table(coefficients, u.device)
    if cells is None:
        ncells = grid.n_cells
        cell_t = None
    else:
        cell_t = torch.as_tensor(np.asarray(cells, dtype=np.int64) if not isinstance(cells, torch.Tensor) else cells,
                                 dtype=torch.int64, device=u.device).contiguous()
        ncells = cell_t.numel()
    out = torch.full((ncells,), float("nan"), dtype=torch.float64, device=u.device)
    uc = u.contiguous()     # held until the launch is enqueued behind every allocation it reads
    if usable and ncells:
        _lib.call("seqvp_prescreen", _lib.ptr(uc), _lib.ptr(D), _lib.ptr(cell_t), ncells, nx, ny, nz,
                  _lib.ptr(out), _lib.stream_handle(u.device))
    return out


def prescreen_flags(u, grid, coefficients, delta, flags=None, count=None):
    """(flags uint8 per cell: 1 = proven J > delta on the whole cell, count of unproven cells)."""
    nz, ny, nx = grid.shape
    D, usable = _table(coefficients, u.device)
    flags = torch.empty(grid.n_cells, dtype=torch.uint8, device=u.device) if flags is None else flags
    count = torch.zeros(1, dtype=torch.int64, device=u.device) if count is None else count
    if not usable:
        flags.zero_()
        count.fill_(grid.n_cells)
        return flags, count
    uc = u.contiguous()
    _lib.call("seqvp_prescreen_flags", _lib.ptr(uc), _lib.ptr(D), float(delta), nx, ny, nz,
              _lib.ptr(flags), _lib.ptr(count), _lib.stream_handle(u.device))
    return flags, count


def device_table(coefficients, device):
    """The closed27 table uploaded once (None when it has a subnormal or nonfinite entry: no proofs)."""
    D, usable = _table(coefficients, device)
    return D if usable else None
