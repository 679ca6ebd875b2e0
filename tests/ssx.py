"""Imports of the package modules for the tests (the package directory name
contains a hyphen, so it is put on sys.path instead of being imported)."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "sparse-seam_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import sparse_seam  # noqa: E402,F401
import ssgen  # noqa: E402,F401
