import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def _oracle_backend():
    from oracle.oracle_backend import OracleBackend
    return OracleBackend.make_state


def _gpu_backend():
    from qip_amd import GpuBackend
    return GpuBackend.make_state


def _gpu64_backend():
    import numpy
    from qip_amd import GpuBackend

    def make(n, groups, feeds, statetype=None, **kw):
        return GpuBackend.make_state(n, groups, feeds, statetype=numpy.complex64, **kw)
    return make


# Graph-level tests run twice: on the CPU oracle (the reference's algorithm, no GPU) and on the
# HIP engine (marked gpu).
BACKENDS = [
    pytest.param("oracle", id="oracle"),
    pytest.param("gpu", id="gpu", marks=pytest.mark.gpu),
]


@pytest.fixture
def backend_constructor(request):
    name = request.param
    return {"oracle": _oracle_backend, "gpu": _gpu_backend, "gpu64": _gpu64_backend}[name]()
