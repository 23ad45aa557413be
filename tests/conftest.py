import importlib.util
import os
import sys

import pytest
# torch first: its bundled HIP runtime then serves libvolprim.so too (both
# carry the soname libamdhip64.so.7).  Loading libvolprim.so first would map
# the system runtime, and torch's own copy, loaded beside it, finds no GPU.
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "cuda-volprim_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")


def _load_pkg():
    if "volprim_amd" in sys.modules:
        return sys.modules["volprim_amd"]
    spec = importlib.util.spec_from_file_location("volprim_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["volprim_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load_pkg()
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def ctx():
    from volprim_amd import volprim
    c = volprim.Context(0)
    name, cus, mhz = c.device_info()
    print(f"\n[device] {name} CUs={cus} clock={mhz:.0f} MHz", flush=True)
    yield c
    c.close()
