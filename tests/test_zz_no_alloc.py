"""fmlh_potrs_noalloc (the solve with its scratch in a static device array)
against fmlh_potrs with a workspace, through the C ABI: bitwise equal on two
calls in a row, the second smaller, for one and several right-hand sides;
a size beyond the static array is refused before any device work."""
import ctypes

import numpy as np
import pytest

import fml_amd._native as nat


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n))
    return B @ B.T + n * np.eye(n)


def test_noalloc_refuses_oversize():
    lib = nat.load()
    cap = lib.fmlh_potrs_noalloc_max_bytes()
    assert cap == 1 << 20
    fake = ctypes.c_void_p(256)  # never dereferenced: the size check comes first
    n = 100000
    rc = lib.fmlh_potrs_noalloc(fake, n, n, fake, n, 64, None)
    assert rc == -3
    assert b"scratch" in lib.fmlh_last_error()


@pytest.mark.gpu
def test_noalloc_matches_workspace_solve(gpu):
    import torch
    from fml_amd.math.cho_solve import potrf_inplace
    lib = nat.load()
    for n, nrhs in ((3000, 1), (1100, 1), (2000, 5), (700, 3)):
        A = torch.from_numpy(_spd(n, n + nrhs)).cuda()
        assert potrf_inplace(A) == 0
        rng = np.random.default_rng(nrhs)
        b = torch.from_numpy(np.asfortranarray(rng.standard_normal((n, nrhs)))).cuda()
        # Fortran (n, nrhs) storage: column j at j * n, as a row-major (nrhs, n) tensor
        xb = b.t().contiguous()
        x_ws, x_na = xb.clone(), xb.clone()
        wsb = lib.fmlh_potrs_ws_bytes(n, nrhs)
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device="cuda")
        st = nat.stream_ptr()
        nat.call("fmlh_potrs", nat.ptr(A), n, n, nat.ptr(x_ws), n, nrhs, nat.ptr(ws), wsb, st)
        nat.call("fmlh_potrs_noalloc", nat.ptr(A), n, n, nat.ptr(x_na), n, nrhs, st)
        torch.cuda.synchronize()
        assert torch.equal(x_ws, x_na), (n, nrhs)
        # and the solve is right: A x = b from the factor's own product
        r = np.linalg.norm(_spd(n, n + nrhs) @ x_na.cpu().numpy().T - b.cpu().numpy())
        assert r <= 1e-10 * np.linalg.norm(b.cpu().numpy()), (n, nrhs, r)
