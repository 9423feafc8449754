"""`diso._C`: the reference's binding module name (src/pybind.cpp:419-447), resolving to the
ctypes binding classes of diso_amd."""
from diso_amd._C import CUDMCDouble, CUDMCFloat, CUMCDouble, CUMCFloat  # noqa: F401

__all__ = ["CUMCFloat", "CUMCDouble", "CUDMCFloat", "CUDMCDouble"]
