import importlib.util
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]


def _load_package():
    if "acoustic_agent_amd" in sys.modules:
        return sys.modules["acoustic_agent_amd"]
    pkg_dir = REPO / "acoustic-agent_amd"
    spec = importlib.util.spec_from_file_location("acoustic_agent_amd", pkg_dir / "__init__.py",
                                                  submodule_search_locations=[str(pkg_dir)])
    module = importlib.util.module_from_spec(spec)
    sys.modules["acoustic_agent_amd"] = module
    spec.loader.exec_module(module)
    return module


_load_package()
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long CPU test")
