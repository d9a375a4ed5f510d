import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def workdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("colore"))


_RUNS = {}


def make_run(name, n_grid, workdir, **kw):
    """Param file + parsed run for config `name` at grid n_grid (cached)."""
    from colore_amd import inputs, lib
    key = (name, n_grid, tuple(sorted(kw.items())))
    if key not in _RUNS:
        d = os.path.join(workdir, f"{name}_{n_grid}_{len(_RUNS)}")
        path = inputs.write_config(name, d, n_grid=n_grid, **kw)
        _RUNS[key] = (path, lib.Run(path))
    return _RUNS[key]
