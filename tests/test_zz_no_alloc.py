"""sdg_proxy_run_static (no device allocation, face traces in the library's
per-GPU scratch array) against sdg_proxy_run with a caller-owned trace buffer:
bitwise equal DOFs, on two calls in a row, the second with a smaller layer
(GPU tests)."""
import ctypes as C

import pytest

from seissol_amd._capi import check, lib
from seissol_amd.kernels import Engine
from seissol_amd.proxy import DeviceLayer, kernel_id, make_setup, run_proxy

pytestmark = pytest.mark.gpu


def _static(engine, layer, kernel, steps, dt):
    import torch
    s = torch.cuda.current_stream().cuda_stream
    return lib().sdg_proxy_run_static(engine.handle, kernel_id(kernel), steps, float(dt), C.byref(layer.view), s)


@pytest.mark.parametrize("form,order", [("f16x3", 6), ("bf16x3", 4)])
def test_static_traces_match_caller_traces(gpu, form, order):
    import torch
    eng = Engine(order, form)
    for n, kernel, steps in ((6000, "all", 3), (1500, "all", 2), (1500, "local", 2)):
        print(f"no-alloc proxy {form} order={order} n={n} kernel={kernel}", flush=True)
        setup = make_setup(order, n=n, mesh="proxy", seed=100 + n)
        ref = DeviceLayer(eng, setup)
        run_proxy(eng, ref, kernel, steps, setup.dt)
        got = DeviceLayer(eng, setup)
        got.view.face_traces = None  # the entry point brings its own traces
        check(_static(eng, got, kernel, steps, setup.dt), "sdg_proxy_run_static")
        torch.cuda.synchronize()
        assert torch.isfinite(got.dofs).all()
        assert torch.equal(got.dofs, ref.dofs), (form, order, n, kernel)


def test_static_rejects_what_it_cannot_hold(gpu):
    eng = Engine(6, "f16x3")
    big = DeviceLayer(eng, make_setup(6, n=20_000, mesh="proxy", seed=3))
    assert _static(eng, big, "all", 1, 1e-4) != 0
    assert b"at most 19418 elements" in lib().sdg_last_error()
    small = DeviceLayer(eng, make_setup(6, n=100, mesh="proxy", seed=4))
    assert _static(eng, small, "neigh", 1, 1e-4) != 0
