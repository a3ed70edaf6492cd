"""Loads the product's Python handle (dac_g-kway_amd/__init__.py) under the
module name `gkway_amd` (the directory name is not an identifier)."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load():
    if "gkway_amd" in sys.modules:
        return sys.modules["gkway_amd"]
    spec = importlib.util.spec_from_file_location(
        "gkway_amd", os.path.join(ROOT, "dac_g-kway_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gkway_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


gk = _load()
