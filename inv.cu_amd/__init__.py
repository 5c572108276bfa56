"""inv.cu_amd — MI355X-native forward+adjoint gradient path of libccy/inv.cu.

libinv_hip.so (csrc/, C ABI in include/inv_hip.h) holds every kernel; this
package holds its ctypes binding (invhip) and synthetic setups (problems).
The directory name is not an importable identifier: load it with
importlib under the module name `inv_cu_amd`.
"""
from . import invhip, problems, serving  # noqa: F401

__all__ = ["invhip", "problems", "serving"]
