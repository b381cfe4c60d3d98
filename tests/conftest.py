import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on the GPU)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def O():
    from oracle import oracle as _O
    return _O


def load_package():
    """Import caustic-tracer_amd/ (a hyphenated directory) as caustic_tracer_amd."""
    import importlib.util

    name = "caustic_tracer_amd"
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(ROOT, "caustic-tracer_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(
        name, path, submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ct():
    return load_package()


def oracle_packed(O, ps):
    """The oracle's view of a product PackedScene (same arrays)."""
    return O.Packed(ps.prims, ps.nodes, ps.nodeCount, ps.mats, ps.lights, ps.primObject, ps.order,
                    ps.sky, getattr(ps, "leafWords", None))
