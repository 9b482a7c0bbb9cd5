"""API surface that needs no GPU: the Python names and semantics of the reference
(test/test_bindings.py:10-71, activemodelbplus/integrator/__init__.py:39-161) and
the C ABI's exported symbols (include/amb.h)."""
import ctypes
import os
import re

import numpy as np
import pytest

from activemodelbplus.integrator import (CudaException, Ensemble, Integrator, Model, Stencil, device_count,
                                         step_iterator, steps_from_time, tile_shapes)

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_model():
    """test/test_bindings.py:10-31."""
    params = (1, 2, 3, 4, 5, 6, 7)
    model = Model(*params)
    copy = Model(model)
    assert params == model.as_tuple()
    assert (model.a, model.b, model.c, model.kappa, model.lamb, model.zeta, model.T) == params
    assert copy is not model
    assert copy == model
    assert copy.as_tuple() == model.as_tuple()
    assert copy.as_dict() == model.as_dict()
    copy.a += 1
    assert copy != model
    assert copy.as_tuple() != model.as_tuple()
    assert copy.as_dict() != model.as_dict()


def test_model_dict_keys_and_repr():
    m = Model(a=-0.25, b=0, c=0.25, kappa=1, lamb=-1, zeta=-4, T=0.1)
    assert list(m.as_dict()) == ["a", "b", "c", "kappa", "lamb", "zeta", "temperature"]
    assert repr(m).startswith("<Model {") and repr(m).endswith("}>")
    assert str(m) == str(m.as_dict())
    assert Model().as_tuple() == (0.0,) * 7


def test_stencil():
    """test/test_bindings.py:33-50."""
    params = (1, 2, 3)
    stencil = Stencil(*params)
    copy = Stencil(stencil)
    assert params == stencil.as_tuple()
    assert (stencil.dt, stencil.dx, stencil.dy) == params
    assert copy is not stencil and copy == stencil
    assert copy.as_dict() == stencil.as_dict() == {"dt": 1.0, "dx": 2.0, "dy": 3.0}
    copy.dx += 1
    assert copy.as_tuple() != stencil.as_tuple()
    assert repr(Stencil(dt=1e-2, dx=1, dy=1)).startswith("<Stencil ")


def test_steps_from_time():
    assert steps_from_time(1e-2, 1.0) == 100
    assert steps_from_time(5e-3, 1e4) == 2_000_000
    with pytest.raises(ValueError):
        steps_from_time(1e-2, 1.005)
    assert list(steps_from_time(0.1, np.array([0.5, 1.0]))) == [5, 10]


def test_step_iterator():
    # chunks: linspace(0, nsteps, nupdates+1), nupdates = clamp(nsteps // min_steps, 1, max)
    s = step_iterator(1e-2, 10.0, max_updates=4, min_update_time=1.0)
    assert list(s) == [250, 250, 250, 250]
    s = step_iterator(1e-2, 10.0, max_updates=100, min_update_time=3.0)
    assert list(s) == [333, 333, 334]
    s = step_iterator(1e-2, 1.0, max_updates=100, min_update_time=10.0)
    assert list(s) == [100]
    s = step_iterator(1e-2, 10.0, max_updates=2, min_update_time=1.0, t_interrupt=[1.0, 2.5, 10.0])
    assert list(s) == [100, 150, 250, 500] and s.sum() == 1000


def test_tile_tables():
    for precision in (32, 64):
        shapes = tile_shapes(precision)
        assert len(shapes) >= 3
        assert all(tw % 4 == 0 and (tw * th) % nt == 0 and nt % 64 == 0 for tw, th, nt in shapes)


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(REPO, "include", "amb.h")).read()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(amb_\w+)\s*\(", header, re.M))
    assert len(declared) >= 25
    lib = ctypes.CDLL(os.path.join(REPO, "active-model-bplus_amd", "lib", "libamb.so"))
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    # calls that need no device
    n = ctypes.c_int(-1)
    assert lib.amb_device_count(ctypes.byref(n)) == 0 and n.value >= 0
    assert lib.amb_tile_count(32, ctypes.byref(n)) == 0 and n.value >= 3
    assert lib.amb_tile_count(16, ctypes.byref(n)) == 1          # AMB_E_ARG
    lib.amb_last_error.restype = ctypes.c_char_p
    assert b"precision" in lib.amb_last_error()


@pytest.mark.skipif(device_count() > 0, reason="checks the no-device error path")
def test_no_device_fails_loudly():
    with pytest.raises(CudaException, match="no HIP device"):
        Integrator(np.zeros((16, 16)), Stencil(1e-2, 1, 1), Model())
    with pytest.raises(CudaException, match="no HIP device"):
        Ensemble([np.zeros((16, 16))], Stencil(1e-2, 1, 1), Model(), [0])


def test_argument_errors():
    with pytest.raises(ValueError):
        Integrator(np.zeros(16), Stencil(1e-2, 1, 1), Model())
    with pytest.raises(TypeError):
        Integrator(np.zeros((4, 4)), Stencil(1e-2, 1, 1))
    with pytest.raises(TypeError):
        Model(1, 2)


def test_integrator_is_subclassable():
    from activemodelbplus.integrator import integrator as core
    assert issubclass(Integrator, core.Integrator)
    assert hasattr(Integrator, "run_for_time") and hasattr(Integrator, "from_checkpoint")


def test_surrogate_argument_errors():
    """Shape and batch limits are checked before any device work (no GPU needed)."""
    from activemodelbplus.integrator import Surrogate
    with pytest.raises(ValueError):
        Surrogate(100, 128)           # sides must be multiples of 64
    with pytest.raises(ValueError):
        Surrogate(256, 256, max_batch=0)
    with pytest.raises(ValueError):
        Surrogate(4096, 4096, max_batch=8)  # 2^27 pixels
