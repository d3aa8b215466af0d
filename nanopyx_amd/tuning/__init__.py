"""Packaged launch-shape stores (written by ``python -m nanopyx_amd.tuning.make_store`` on the target GPU)."""
import os

PACKAGED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mi355x.json")
