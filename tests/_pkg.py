"""Loads the product package (its directory name has hyphens)."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "additive-fft-reed-solomon_amd")


def load():
    if "ars_amd" in sys.modules:
        return sys.modules["ars_amd"]
    spec = importlib.util.spec_from_file_location("ars_amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ars_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
