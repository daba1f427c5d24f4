"""Loads the package's binding of libhm_stitch.so (hm_stitch.py, one
directory up)."""
import os
import sys

_PARENT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _PARENT not in sys.path:
    sys.path.insert(0, _PARENT)

import hm_stitch  # noqa: E402,F401
