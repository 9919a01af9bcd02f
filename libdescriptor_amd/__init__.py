"""libdescriptor_amd -- MI355X (gfx950) symmetry-function descriptors.

Drop-in for the reference's Python package ``libdescriptor``
(python_package/libdescriptor/__init__.py re-exports the pybind module;
neighbor.py adds NeighborList / assemble_forces)::

    import libdescriptor_amd as libdescriptor

Everything computes on the GPU through hand-written HIP kernels; importing
fails loudly if the native library has not been built
(``python libdescriptor_amd/build.py``).
"""
import os

try:  # load torch's HIP runtime first so that one runtime serves both
    import torch as _torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for the NumPy API
    _torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from .libdescriptor import (AvailableDescriptors, DescriptorKind, NeighList,  # noqa: F401
                                FP32_BULK, FP64, compute, compute_batch,
                                compute_single_atom, create, create_paddings,
                                get_neigh_kim, gradient, gradient_batch,
                                gradient_single_atom, jacobian,
                                num_gradient_single_atom, param_gradient)
except ImportError as e:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "libdescriptor_amd: the native module is not built; run "
        "`python libdescriptor_amd/build.py` (" + str(e) + ")") from e

from .neighbor import NeighborList, NeighborListError, assemble_forces  # noqa: F401,E402

__all__ = [
    "AvailableDescriptors", "DescriptorKind", "NeighList", "FP64", "FP32_BULK", "compute",
    "compute_batch", "compute_single_atom", "create", "create_paddings", "get_neigh_kim",
    "gradient", "gradient_batch", "gradient_single_atom", "jacobian",
    "num_gradient_single_atom", "param_gradient", "NeighborList", "NeighborListError", "assemble_forces",
]
