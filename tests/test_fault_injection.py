"""Error handling without a failing device: the library's fault-injection
flag (hvr_debug_inject_fault) makes an entry point fail before it does any
device work, and the Python layers surface that failure as RuntimeError with
the library's code and message -- the reference raises RuntimeError for
failures of its GL device layer (pyvr/moderngl_renderer/manager.py:190-191,
224-225).  Runs on the CPU: the injected entry points return before any HIP
call."""

import os

import pytest

from pyvr_amd.hip_renderer import _lib

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libhipvr.so not built")


@pytest.fixture(autouse=True)
def disarm():
    yield
    _lib.inject_fault(None)


def test_injected_fault_fires_once_on_the_chosen_call():
    L = _lib.lib()
    # hvr_set_params(NULL, NULL) fails validation with HVR_EINVAL ...
    assert L.hvr_set_params(None, None) == -1
    _lib.inject_fault("hvr_set_params", code=-4, after=1)
    assert L.hvr_set_params(None, None) == -1            # call 0: not yet
    assert L.hvr_set_params(None, None) == -4            # call 1: injected
    assert L.hvr_last_error() == b"injected fault in hvr_set_params"
    assert L.hvr_set_params(None, None) == -1            # once only
    # other entry points are not affected
    _lib.inject_fault("hvr_bind", code=-2, after=0)
    assert L.hvr_set_params(None, None) == -1
    assert L.hvr_bind(None, None, 0, None, 0) == -2


def test_bad_injection_arguments_are_rejected():
    with pytest.raises(RuntimeError, match="negative code"):
        _lib.inject_fault("hvr_create", code=1)
    with pytest.raises(RuntimeError, match="negative code"):
        _lib.inject_fault("hvr_create", code=-1, after=-1)


@pytest.mark.parametrize("code", [-1, -2, -3, -4])
def test_renderer_construction_failure_is_a_runtime_error(code):
    from pyvr_amd.hip_renderer import VolumeRenderer

    _lib.inject_fault("hvr_create", code=code)
    with pytest.raises(RuntimeError, match=rf"hvr_create failed \({code}\): injected fault in hvr_create"):
        VolumeRenderer(64, 64)


def test_batch_and_plain_renderers_surface_failures():
    from pyvr_amd.hip_renderer import BatchRenderer
    from pyvr_amd.hip_renderer.plain import PlainRenderer

    _lib.inject_fault("hvr_create", code=-3)
    with pytest.raises(RuntimeError, match="injected fault in hvr_create"):
        BatchRenderer()
    _lib.inject_fault("hvr_create", code=-2)
    with pytest.raises(RuntimeError, match="injected fault in hvr_create"):
        PlainRenderer()


def test_server_construction_failure_is_a_runtime_error():
    from pyvr_amd.hip_renderer.serve import RenderServer

    _lib.inject_fault("hvr_create", code=-3)
    with pytest.raises(RuntimeError, match="injected fault in hvr_create"):
        RenderServer(memory_budget=1 << 30)
