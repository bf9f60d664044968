"""MI355X-native (gfx950 HIP) frequency-domain FEM hot path for EMerge.

Host code mirrors the reference's assembler / solver plug-in interfaces and
calls hand-written HIP kernels in ``libemerge_hip.so`` through a C-ABI
(``include/emerge_hip.h``).
"""
from .mesh import Mesh3D, kuhn_box
from .basis import Nedelec2

__all__ = ["Mesh3D", "kuhn_box", "Nedelec2"]
__version__ = "0.1.0"
