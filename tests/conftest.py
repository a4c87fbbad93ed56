import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session", autouse=True)
def torch_gpu_first(request):
    """When GPU tests run, torch's HIP runtime is initialised before the
    library's first HIP call: torch's wheel carries its own copy of the runtime,
    which finds no device when it starts after the system one."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except ImportError:
            pass
    yield


@pytest.fixture(scope="session")
def oracle():
    import oracle_ffi
    oracle_ffi.lib()
    return oracle_ffi


STUB_GPMETIS = """#!/bin/sh
# stand-in for gpmetis: checks the arguments, writes <graph>.part.<k> with
# line i holding (3*i + 1) mod k, and logs its call
[ "$1" = "-ptype=rb" ] || exit 3
n=$(head -1 "$2" | cut -d' ' -f1)
i=0
: > "$2.part.$3"
while [ $i -lt $n ]; do echo $(( (3 * i + 1) % $3 )) >> "$2.part.$3"; i=$((i + 1)); done
echo "$2 $3" >> "$STUB_GPMETIS_LOG"
"""


@pytest.fixture
def stub_gpmetis(tmp_path, monkeypatch):
    """Puts a stand-in `gpmetis` first on PATH; returns the file that logs
    each call as "<graph path> <k>"."""
    bindir = tmp_path / "stub_bin"
    bindir.mkdir()
    stub = bindir / "gpmetis"
    stub.write_text(STUB_GPMETIS)
    stub.chmod(0o755)
    monkeypatch.setenv("PATH", f"{bindir}:{os.environ['PATH']}")
    log = tmp_path / "gpmetis_calls.log"
    monkeypatch.setenv("STUB_GPMETIS_LOG", str(log))
    return log
