"""bemcs_thermal_noalloc (the thermal update with its scratch in a
library-held device array, for callers that cannot allocate) gives the same
bits as bemcs_thermal with a caller workspace: on two calls in a row, the
second on a smaller mesh, and it rejects a mesh beyond its scratch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

F32 = np.float32


def _inputs(nx, ny, seed):
    from test_gpu_surface import _thermal_case
    return _thermal_case(nx, ny, seed)


def _params(lib, T, B, heat, masks, ws):
    from thruster_cx_code_amd import _lib
    ny, nx = T.shape
    bits = np.zeros((ny, nx), np.uint8)
    for g, m in enumerate(masks):
        bits |= (m.astype(np.uint8) << g)
    keep = dict(Ti=torch.from_numpy(T.ravel().copy()).cuda(),
                To=torch.zeros(nx * ny, dtype=torch.float32, device="cuda"),
                hd=torch.from_numpy(heat.ravel().view(np.int64).copy()).cuda(),
                clear=torch.full((nx * ny,), 5, dtype=torch.int64, device="cuda"),
                bd=torch.from_numpy(B.astype(np.uint8).ravel()).cuda(),
                gd=torch.from_numpy(bits.ravel()).cuda(),
                tg=torch.full((len(masks),), -1.0, dtype=torch.float32, device="cuda"))
    p = _lib.Thermal()
    v = lambda t: C.c_void_p(t.data_ptr())
    p.T_in, p.T_out, p.heat_acc, p.heat_clear = v(keep["Ti"]), v(keep["To"]), v(keep["hd"]), v(keep["clear"])
    p.is_bound, p.grid_bits, p.T_grids = v(keep["bd"]), v(keep["gd"]), v(keep["tg"])
    p.workspace = v(ws) if ws is not None else None
    p.cool32, p.t300_4_32 = float(F32(3e-12)), float(F32(300.0 ** 4))
    p.fo_x32, p.fo_y32 = float(F32(0.2)), float(F32(0.137))
    p.nx, p.ny, p.n_grids, p.sweeps = nx, ny, len(masks), 10
    return p, keep


def test_noalloc_thermal_matches_workspace_version(hip_lib):
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for nx, ny, seed in ((301, 157, 21), (160, 72, 22)):
        print(f"[gpu] thermal noalloc {nx}x{ny}", flush=True)
        T, B, heat, masks = _inputs(nx, ny, seed)
        ws = torch.zeros(int(hip_lib.bemcs_thermal_workspace_bytes(nx, ny, len(masks))), dtype=torch.uint8,
                         device="cuda")
        pa, ka = _params(hip_lib, T, B, heat, masks, ws)
        pb, kb = _params(hip_lib, T, B, heat, masks, None)
        assert hip_lib.bemcs_thermal(C.byref(pa), s) == 0
        assert hip_lib.bemcs_thermal_noalloc(C.byref(pb), s) == 0
        torch.cuda.synchronize()
        assert torch.equal(ka["To"].view(torch.int32), kb["To"].view(torch.int32))
        assert torch.equal(ka["tg"].view(torch.int32), kb["tg"].view(torch.int32))
        assert torch.equal(ka["clear"], kb["clear"])
        assert float(kb["tg"].min().item()) >= 300.0


def test_noalloc_thermal_rejects_mesh_beyond_scratch(hip_lib):
    from thruster_cx_code_amd import _lib
    nx, ny = 8192, 8192
    assert hip_lib.bemcs_thermal_workspace_bytes(nx, ny, 8) > (1 << 20)
    p = _lib.Thermal()
    dummy = torch.zeros(4, dtype=torch.float32, device="cuda")
    p.T_in = p.T_out = p.is_bound = p.grid_bits = p.T_grids = C.c_void_p(dummy.data_ptr())
    p.nx, p.ny, p.n_grids, p.sweeps = nx, ny, 8, 10
    rc = hip_lib.bemcs_thermal_noalloc(C.byref(p), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == _lib.EINVAL
    assert b"1 MiB" in hip_lib.bemcs_last_error()
