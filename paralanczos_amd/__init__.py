"""paralanczos_amd - MI355X-native batched principal-submatrix Lanczos eigenvalues.

Native core: ``libparalanczos.so`` (HIP kernels for gfx950 behind the C ABI of
``include/paralanczos.h`` and the drop-in ``include/ParaLanczos.h``).

* ``paralanczos_amd.solver`` - device-tensor API (torch supplies memory/streams).
* ``paralanczos_amd.legacy`` - the reference's host entry points and the Matlab
  front-end semantics on numpy arrays.
* ``paralanczos_amd.build``  - hipcc build recipe.
"""
from ._lib import LIB_PATH, ParaLanczosError, load  # noqa: F401

__all__ = ["LIB_PATH", "ParaLanczosError", "load"]
