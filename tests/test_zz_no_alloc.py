"""smg_pe_forward_noalloc, the entry point that allocates nothing (no
counterpart in the reference; for callers that cannot allocate): through
the C ABI into a caller's device buffer, against smg_pe_forward on two calls
in a row (the second on a smaller geometry), and its size limit."""
import ctypes as C

import numpy as np
import pytest

from conftest import small_geometry

pytestmark = pytest.mark.gpu

CAPACITY = 2


def _noalloc(smg_mod, ctx, prec):
    import torch
    n = ctx.dims.n_det * ctx.dims.n_vox
    buf = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()  # the fill is on torch's stream, the library uses its own
    st = smg_mod.Stats()
    s = smg_mod.lib().smg_pe_forward_noalloc(ctx._h, 0, smg_mod.PREC[prec], C.c_void_p(buf.data_ptr()),
                                             C.byref(st))
    assert s == 0, smg_mod.lib().smg_last_error(ctx._h)
    return buf.cpu().numpy().reshape(ctx.dims.n_det, ctx.dims.n_vox), st.as_dict()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_noalloc_equals_forward(smg_mod, params_mod, prec):
    ctx = smg_mod.Context(0)
    try:
        cases = [(params_mod.config(1), (4, 4, 4)), (small_geometry(params_mod, beta=0.07), (3, 4, 2))]
        sizes = []
        for p, sub in cases:
            ctx.set_geometry(p)
            ctx.set_pe_subsamples(*sub)
            got, st = _noalloc(smg_mod, ctx, prec)
            ref, _ = ctx.pe_forward(0, prec)
            print(f"noalloc {prec}: {got.shape}, kernel {st['kernel_ms']:.3f} ms", flush=True)
            assert np.array_equal(got, ref)
            sizes.append(got.size)
        assert sizes[1] < sizes[0]
    finally:
        ctx.close()


def test_noalloc_refuses_beyond_scratch(smg_mod, params_mod):
    """2048 bars at the reference's 8 x 16 x 8 sub-cells need a 285 MB
    sub-cell table: beyond the 32 MiB scratch, refused with a message."""
    import torch
    p = params_mod.config(2)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(8, 16, 8)
        buf = torch.empty(16, dtype=torch.float32, device="cuda")
        s = smg_mod.lib().smg_pe_forward_noalloc(ctx._h, 0, 0, C.c_void_p(buf.data_ptr()), None)
        assert s == CAPACITY
        assert b"scratch" in smg_mod.lib().smg_last_error(ctx._h)
