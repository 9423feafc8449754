"""Import-name shim: `diso`, the reference package's import name (diso/__init__.py:1-9), resolving
to diso_amd, so code written against the reference runs unchanged:

    from diso import DiffMC, DiffDMC
    import diso._C   # CUMCFloat, CUMCDouble, CUDMCFloat, CUDMCDouble (src/pybind.cpp:419-447)
"""
from diso_amd import DiffDMC, DiffMC  # noqa: F401

from . import _C  # noqa: F401

__all__ = ["DiffMC", "DiffDMC"]
