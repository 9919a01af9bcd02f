"""PyTorch-ROCm autograd for the symmetry-function descriptor.

``SymmetryFunctionsFn.apply(coords, plan, precision)`` returns the
(n_centers, width) descriptor tensor; its backward is the hand-written HIP
adjoint (d coords = J^T d desc), running on torch's current stream, with the
padding rows of coords included as in the reference's gradient
(python_bindings/python_bindings.cpp:235,252).  The same Function serves one
configuration and a batch of uneven configurations (one persistent launch
over all centres of a batch plan).
"""
import torch

from . import _capi


class SymmetryFunctionsFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, plan, width, precision=_capi.FP64):
        if coords.dtype != torch.float64 or not coords.is_cuda:
            raise TypeError("coords must be a float64 tensor on the GPU")
        coords = coords.contiguous()
        desc = torch.zeros((plan.n_centers, width), dtype=torch.float64, device=coords.device)
        plan.compute(coords, desc, precision)
        ctx.save_for_backward(coords)
        ctx.plan = plan
        ctx.precision = precision
        return desc

    @staticmethod
    def backward(ctx, grad_desc):
        (coords,) = ctx.saved_tensors
        g = grad_desc.to(torch.float64).contiguous()
        dcoords = torch.zeros_like(coords)
        ctx.plan.gradient(coords, g, dcoords, ctx.precision)
        return dcoords, None, None, None


class SymmetryFunctionsLayer(torch.nn.Module):
    """Descriptor layer over a fixed neighbour list held on the device.

    ``descriptor`` is a DescriptorKind from init_descriptor /
    init_symmetry_functions.  For a batch pass ``n_atoms`` and ``config_ptr``
    (host int arrays) as compute_batch takes them."""

    def __init__(self, descriptor, species, neighlist, numneigh, n_atoms=None, config_ptr=None,
                 precision=_capi.FP64, device="cuda"):
        super().__init__()
        dev = torch.device(device)
        self.descriptor = descriptor
        self.width = descriptor.width
        self.precision = precision
        self.species = torch.as_tensor(species, dtype=torch.int32).to(dev).contiguous()
        self.neighlist = torch.as_tensor(neighlist, dtype=torch.int32).to(dev).contiguous()
        self.numneigh = torch.as_tensor(numneigh, dtype=torch.int32).to(dev).contiguous()
        self.plan = _capi.Plan(descriptor, int(self.species.numel()), self.species,
                               self.neighlist, self.numneigh, n_atoms=n_atoms,
                               config_ptr=config_ptr)

    def forward(self, coords):
        return SymmetryFunctionsFn.apply(coords, self.plan, self.width, self.precision)


class _FromPositionsFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, module):
        pos = positions.detach().contiguous()
        lay = _capi.DeviceLayout(pos, module.species, module.cell, module.pbc, module.cutoff)
        plan = _capi.Plan(module.descriptor, lay.n_contrib + lay.n_padding, lay.species,
                          lay.neighlist, lay.numneigh)
        # the adjoint's records from torch's caching allocator: a plan per call
        # then allocates no large buffer of its own
        plan.set_records(torch.empty(3 * plan.num_records, dtype=torch.float64,
                                     device=positions.device))
        desc = torch.zeros((lay.n_contrib, module.width), dtype=torch.float64,
                           device=positions.device)
        plan.compute(lay.coords, desc, module.precision)
        ctx.lay, ctx.plan, ctx.precision = lay, plan, module.precision
        module.last_layout = lay
        return desc

    @staticmethod
    def backward(ctx, grad_desc):
        g = grad_desc.to(torch.float64).contiguous()
        lay = ctx.lay
        ctx.plan.adjoint_edges(lay.coords, g, ctx.precision)
        # padding rows are their atoms shifted by lattice vectors: the gather
        # folds them back as it sums each atom's records
        return ctx.plan.gather_fold(lay), None


class SymmetryFunctionsFromPositions(torch.nn.Module):
    """Descriptors of the atoms of one configuration straight from their positions
    on the device, differentiable with respect to those positions.

    Each call builds the padding rows and the KLIFF-order neighbour list on the
    GPU (``ldsf_layout_create``, the device twin of NeighborList,
    python_package/libdescriptor/neighbor.py:35-154), a plan over it and the
    forward kernel; the backward runs the adjoint kernel and the gather with
    the fold of the padding rows onto their atoms (assemble_forces,
    neighbor.py:227-263) in its epilogue.  This is the reference's NeighborList -> compute ->
    gradient -> assemble_forces pipeline as one autograd node, with nothing
    leaving the device but the layout's sizes.

    ``species`` are int codes per atom (device or host), ``cell`` a 3x3 host
    array (rows are the cell vectors), ``pbc`` three flags, ``cutoff`` the
    neighbour-list cutoff (the influence distance)."""

    def __init__(self, descriptor, species, cell, pbc, cutoff, precision=_capi.FP64,
                 device="cuda"):
        super().__init__()
        import numpy as np
        self.descriptor = descriptor
        self.width = descriptor.width
        self.precision = precision
        self.species = torch.as_tensor(species, dtype=torch.int32).to(torch.device(device))
        self.species = self.species.contiguous()
        self.cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)
        self.pbc = np.ascontiguousarray(pbc, dtype=np.int32).reshape(3)
        self.cutoff = float(cutoff)
        self.last_layout = None

    def forward(self, positions):
        if positions.dtype != torch.float64 or not positions.is_cuda:
            raise TypeError("positions must be a float64 tensor on the GPU")
        return _FromPositionsFn.apply(positions, self)
