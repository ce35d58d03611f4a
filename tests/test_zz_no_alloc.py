"""tm_bem_apply_noalloc (the boundary product that allocates nothing: its
partial products live in a static device array) against tm_bem_apply of a
context, through the C ABI, on two calls in a row, the second with a smaller
matrix; and its size check."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from tetmag_amd import _native as N
    L = N.lib()
    f = L.tm_bem_apply_noalloc
    f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return N, f


def test_noalloc_matches_context_product(gpu_available):
    import torch
    from tetmag_amd import problem as P
    from tetmag_amd.context import FieldContext
    N, noalloc = _lib()
    probs = [P.cuboid_20x20x4(), P.small_problem(nx=7, ny=6, nz=3, h=4.0, jitter=0.1, seed=11)]
    assert probs[1].bnd.bnx < probs[0].bnd.bnx
    for k, prob in enumerate(probs):  # two calls in a row, the second smaller
        print(f"\n[case] noalloc call {k} bnx={prob.bnd.bnx}", flush=True)
        ctx = FieldContext(prob, tol=1e-12)
        try:
            b = prob.bnd.bnx
            Bt = ctx.bem_matrix_device()  # (b, b) view; the storage is column-major with ld = b
            Bs = Bt.t()
            assert Bs.is_contiguous()
            ld = b if b % 2 == 0 else b + 1
            with torch.cuda.stream(ctx.stream):
                Bp = torch.zeros((b, ld), dtype=torch.float64, device=ctx.device)  # column j at Bp[j, :]
                Bp[:, :b] = Bs
                x = torch.rand(b, dtype=torch.float64, device=ctx.device) - 0.5
                y = torch.full((b,), np.nan, dtype=torch.float64, device=ctx.device)
            ctx.stream.synchronize()
            st = noalloc(b, ld, Bp.data_ptr(), x.data_ptr(), y.data_ptr(), C.c_void_p(ctx.stream.cuda_stream))
            assert st == N.TM_OK
            ctx.stream.synchronize()
            ref = ctx.bem_apply(x.cpu().numpy())
            got = y.cpu().numpy()
            assert np.all(np.isfinite(got))
            assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
        finally:
            ctx.close()


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu_available):
    import torch
    N, noalloc = _lib()
    buf = torch.zeros(16, dtype=torch.float64, device="cuda:0")
    p = buf.data_ptr()
    # 1,048,576 columns need more partial products than the 32 MiB scratch
    # holds: rejected before any device work
    assert noalloc(1 << 20, 1 << 20, p, p, p, None) == N.TM_ERR_BAD_ARG
    assert noalloc(64, 63, p, p, p, None) == N.TM_ERR_BAD_ARG  # ld < bnx
