"""Test configuration: the ``gpu`` marker and the loader of the package (its directory name,
algebraic-multilevel-preconditioner_amd, is not a Python identifier; it is registered as ``amp_amd``)."""
import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def load_package():
    if "amp_amd" in sys.modules:
        return sys.modules["amp_amd"]
    pkg = os.path.join(REPO, "algebraic-multilevel-preconditioner_amd")
    spec = importlib.util.spec_from_file_location("amp_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["amp_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
