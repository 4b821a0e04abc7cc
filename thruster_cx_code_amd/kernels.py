"""Drop-in kernel functions with the reference's Python signatures.

``push_particles_boris_hip``, ``accumulate_rho_hip`` and
``thermal_conduction_hip`` take exactly the arguments of
``push_particles_boris_cpu`` / ``accumulate_rho_cpu`` /
``thermal_conduction_taichi`` (Python/physics_engine.py:188-190, :178,
:147-156) and mutate their array arguments in place.  Arrays may be NumPy (copied to the device and back, as the Taichi path
does at :1215-1232) or torch tensors already on the GPU (no copies).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check
from .device import ptr, stream_ptr, require_gpu


def _dev_view(a, n, dtype):
    """(device tensor, writeback) for the first n entries of a."""
    if isinstance(a, torch.Tensor):
        if not a.is_cuda or a.dtype != dtype or not a.is_contiguous():
            raise TypeError("device tensors must be contiguous CUDA tensors of dtype %s" % dtype)
        return a, None
    arr = np.asarray(a)
    t = torch.from_numpy(np.ascontiguousarray(arr.reshape(-1)[:n], dtype=np.dtype(str(dtype).split('.')[-1]))).cuda()
    return t, arr


def _field(F):
    if isinstance(F, torch.Tensor):
        if not F.is_cuda or F.dtype != torch.float32:
            raise TypeError("device fields must be float32 CUDA tensors")
        return F.contiguous()
    return torch.from_numpy(np.ascontiguousarray(F, dtype=np.float32)).cuda()


def _is_zero_field(F):
    if F is None:
        return True
    if isinstance(F, torch.Tensor):
        return not bool(torch.any(F != 0))
    return not np.any(F)


def push_particles_boris_hip(x, y, vx, vy, vz, Ex, Ey, Bx, By, Bz, num_p, dx, dy, nx, ny, dt, q_m):
    """Gather + Boris + drift on the GPU, bit-identical to push_particles_boris_cpu."""
    n = int(num_p)
    if n == 0:
        return
    require_gpu()
    lib = _lib.load()
    views = [_dev_view(a, n, torch.float32) for a in (x, y, vx, vy, vz)]
    ex, ey = _field(Ex), _field(Ey)
    if _is_zero_field(Bx) and _is_zero_field(By) and _is_zero_field(Bz):
        bx = by = bz = None
    else:
        bx, by, bz = _field(Bx), _field(By), _field(Bz)
    check(lib.bemcs_push_boris(*[ptr(v[0]) for v in views], n, ptr(ex), ptr(ey), ptr(bx), ptr(by),
                               ptr(bz), int(nx), int(ny), float(dx), float(dy), float(dt), float(q_m),
                               1, stream_ptr()), "push_boris")
    for t, host in views:
        if host is not None:
            host.reshape(-1)[:n] = t.cpu().numpy()


def accumulate_rho_hip(x, y, rho, num_p, dx, dy, nx, ny, charge_density):
    """NGP deposit on the GPU, bit-identical to accumulate_rho_cpu (np.add.at)."""
    n = int(num_p)
    if n == 0:
        return
    require_gpu()
    lib = _lib.load()
    xt, _ = _dev_view(x, n, torch.float32)
    yt, _ = _dev_view(y, n, torch.float32)
    if isinstance(rho, torch.Tensor):
        rt, host = rho, None
    else:
        host = rho
        rt = torch.from_numpy(np.ascontiguousarray(rho, dtype=np.float32).reshape(-1)).cuda()
    counts = torch.zeros(int(nx) * int(ny), dtype=torch.int32, device=rt.device)
    s = stream_ptr()
    check(lib.bemcs_deposit_ngp_counts(ptr(xt), ptr(yt), n, float(dx), float(dy), int(nx), int(ny),
                                       ptr(counts), s), "deposit")
    check(lib.bemcs_counts_to_rho(ptr(counts), int(nx) * int(ny), float(charge_density), ptr(rt), s),
          "counts_to_rho")
    if host is not None:
        host[...] = rt.cpu().numpy().reshape(host.shape)


def accumulate_rho_cic_hip(x, y, rho, num_p, dx, dy, nx, ny, charge_density):
    """Bilinear (cloud-in-cell) deposit of the model manual
    (Python/docs/PY-BEMCS_Manual.tex:215-230) with accumulate_rho_cpu's
    signature (Python/physics_engine.py:178): rho[node] += charge_density *
    (sum of the node's bilinear weights), the weights summed exactly in
    2^-32 fixed point and added to rho once per node in float64 (one
    rounding to float32)."""
    n = int(num_p)
    if n == 0:
        return
    require_gpu()
    lib = _lib.load()
    xt, _ = _dev_view(x, n, torch.float32)
    yt, _ = _dev_view(y, n, torch.float32)
    if isinstance(rho, torch.Tensor):
        rt, host = rho, None
    else:
        host = rho
        rt = torch.from_numpy(np.ascontiguousarray(rho, dtype=np.float32).reshape(-1)).cuda()
    w = torch.zeros(int(nx) * int(ny), dtype=torch.int64, device=rt.device)
    s = stream_ptr()
    check(lib.bemcs_deposit_cic(ptr(xt), ptr(yt), n, None, float(dx), float(dy), int(nx), int(ny), ptr(w), s),
          "deposit_cic")
    check(lib.bemcs_weights_to_rho(ptr(w), int(nx) * int(ny), float(charge_density), ptr(rt), s),
          "weights_to_rho")
    if host is not None:
        host[...] = rt.cpu().numpy().reshape(host.shape)


def thermal_conduction_hip(T, T_new, mask, nx, ny, Fo_x, Fo_y):
    """One conduction sweep with thermal_conduction_taichi's signature
    (Python/physics_engine.py:147-156): writes T_new from T.  Semantics and
    float32 arithmetic of the reference NumPy loop (:1598-1613), which the
    call site at :1586-1597 uses when Taichi is off: only interior electrode
    nodes (mask == 1) change, to max(T + dT, 300), insulated towards
    non-electrode neighbours; every other node copies T.  Bit-identical to
    that loop.  Arrays may be NumPy (copied to the device and back) or float32
    / int32 CUDA tensors of shape (ny, nx)."""
    require_gpu()
    lib = _lib.load()
    nx, ny = int(nx), int(ny)
    n = nx * ny
    tin, _ = _dev_view(T, n, torch.float32)
    if isinstance(T_new, torch.Tensor):
        tout, host = _dev_view(T_new, n, torch.float32)[0], None
    else:
        host = T_new
        tout = torch.empty(n, dtype=torch.float32, device=tin.device)
    if isinstance(mask, torch.Tensor):
        mt, _ = _dev_view(mask, n, torch.int32)
    else:
        mt = torch.from_numpy(np.ascontiguousarray(np.asarray(mask).reshape(-1)[:n], dtype=np.int32)).cuda()
    check(lib.bemcs_thermal_conduction(ptr(tin), ptr(tout), ptr(mt), nx, ny, float(np.float32(Fo_x)),
                                       float(np.float32(Fo_y)), stream_ptr()), "thermal_conduction")
    if host is not None:
        host.reshape(-1)[:n] = tout.cpu().numpy()
