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


def report(name, obj):
    """Writes a JSON report into $CLR_REPORT_DIR (default gpurun_out/ under the
    repository root when that folder exists) and prints it."""
    import json
    d = os.environ.get("CLR_REPORT_DIR")
    if not d:
        d = os.path.join(ROOT, "gpurun_out") if os.path.isdir(os.path.join(ROOT, "gpurun_out")) else None
    text = json.dumps(obj, indent=1)
    print(text)
    if d:
        with open(os.path.join(d, name), "w") as f:
            f.write(text + "\n")
