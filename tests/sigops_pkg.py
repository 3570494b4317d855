"""Imports the product package (directory ``wgpu-sigops_amd``, not a valid identifier)."""
import importlib
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
sigops = importlib.import_module("wgpu-sigops_amd")
