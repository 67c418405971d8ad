"""cr_mlp_train_step_static, the allocation-free training step: through the
C ABI it gives the parameters and loss of cr_mlp_train_step bit for bit on
two calls in a row, the second with fewer rows (the fixed partials array then
still holds the first call's larger contents), and it refuses more rows than
its array holds."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_train_step_static_matches_train_step(cv, gpu_ctx):
    import torch
    from gpu_common import log
    lib = cv.lib()
    params = cv.FarFieldMLP.init_params(3)
    a = cv.FarFieldMLP(gpu_ctx, params)
    b = cv.FarFieldMLP(gpu_ctx, params)
    gen = torch.Generator(device="cpu").manual_seed(5)
    stream = torch.cuda.current_stream()
    try:
        for step, n in enumerate((cv._abi.CR_MLP_STATIC_ROWS, 5000)):
            log(f"[no-alloc] train step {step}: {n} rows, cr_mlp_train_step vs cr_mlp_train_step_static")
            x = torch.randn((n, 17), generator=gen).cuda()
            t = torch.rand((n, 5), generator=gen).cuda()
            a.train_step(x, t, lr=1e-3)
            rc = lib.cr_mlp_train_step_static(b.handle, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()), n,
                                              C.c_float(1e-3), C.c_void_p(stream.cuda_stream))
            assert rc == cv._abi.CR_OK, lib.cr_last_error(gpu_ctx.handle)
            torch.cuda.synchronize()
            pa, pb = a.get_params(), b.get_params()
            assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), f"parameters differ after step {step}"
            la, lb = a.last_loss(), b.last_loss()
            assert np.float32(la).view(np.uint32) == np.float32(lb).view(np.uint32), (la, lb)
            assert np.isfinite(la) and la > 0
        # more rows than the fixed array holds: refused, nothing launched
        n = cv._abi.CR_MLP_STATIC_ROWS + 1
        x = torch.zeros((n, 17), device="cuda")
        t = torch.zeros((n, 5), device="cuda")
        with pytest.raises(cv.CloudVoxError) as e:
            b.train_step_static(x, t, lr=1e-3)
        assert e.value.code == cv._abi.CR_EINVAL and "CR_MLP_STATIC_ROWS" in str(e.value)
        assert np.array_equal(b.get_params().view(np.uint32), pb.view(np.uint32))
        log("[no-alloc] both steps bit-identical; oversize call refused")
    finally:
        a.close()
        b.close()
