"""ctypes binding of the device-pointer half of the C-ABI (include/ldsf.h):
plans over device-resident neighbour lists and the forward / adjoint launches
on a caller's HIP stream.  Used by the torch autograd Function, the batched
server and the benchmark.  Raises if libldsf.so is not built.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libldsf.so")

LDSF_OK = 0
LDSF_ERR_NEIGHBOR = 5
LDSF_ERR_TIMEOUT = 9
FP64 = 0
FP32_BULK = 1

_lib = None

# every function the header declares, with (restype, argtypes)
_vp, _i, _d, _l = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
_PROTOS = {
    "ldsf_last_error": (ctypes.c_char_p, []),
    "ldsf_sf_create": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp, _vp, _i]),
    "ldsf_sf_create_from_file": (_i, [_vp, ctypes.c_char_p]),
    "ldsf_sf_destroy": (None, [_vp]),
    "ldsf_sf_width": (_i, [_vp]),
    "ldsf_sf_num_params": (_i, [_vp]),
    "ldsf_param_gradient": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_plan_param_gradient": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_sf_num_species": (_i, [_vp]),
    "ldsf_sf_species_name": (ctypes.c_char_p, [_vp, _i]),
    "ldsf_sf_set_precision": (_i, [_vp, _i]),
    "ldsf_compute": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_compute_single_atom": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp, _vp]),
    "ldsf_compute_batch": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_compute_overwrite": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_compute_batch_overwrite": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_gradient": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_gradient_single_atom": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "ldsf_gradient_batch": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_jacobian": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_num_gradient_single_atom": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp, _vp, _vp]),
    "ldsf_plan_create": (_i, [_vp, _vp, _i, _vp, _i, _vp, _vp, _vp, _vp]),
    "ldsf_plan_create_batch": (_i, [_vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "ldsf_plan_destroy": (None, [_vp]),
    "ldsf_plan_num_centers": (_i, [_vp]),
    "ldsf_plan_num_edges": (_l, [_vp]),
    "ldsf_plan_max_neighbors": (_i, [_vp]),
    "ldsf_plan_compute": (_i, [_vp, _vp, _vp, _i, _vp]),
    "ldsf_plan_gradient": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "ldsf_plan_adjoint_edges": (_i, [_vp, _vp, _vp, _i, _vp]),
    "ldsf_plan_gather": (_i, [_vp, _vp, _vp]),
    "ldsf_plan_forward_gradient": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "ldsf_plan_forward_gradient_assign": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "ldsf_plan_forward_adjoint_edges_assign": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "ldsf_plan_gather_assign": (_i, [_vp, _vp, _vp]),
    "ldsf_plan_forward_adjoint_edges": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "ldsf_plan_kernel_family": (_i, [_vp]),
    "ldsf_plan_set_kernel_family": (_i, [_vp, _i]),
    "ldsf_plan_num_records": (ctypes.c_long, [_vp]),
    "ldsf_plan_set_records": (_i, [_vp, _vp, ctypes.c_long]),
    "ldsf_nbl_create_paddings": (_i, [_i, _d, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_padding_count": (_i, [_vp]),
    "ldsf_padding_copy": (_i, [_vp, _vp, _vp, _vp]),
    "ldsf_padding_free": (None, [_vp]),
    "ldsf_nbl_create": (_vp, []),
    "ldsf_nbl_free": (None, [_vp]),
    "ldsf_nbl_build": (_i, [_vp, _i, _vp, _d, _i, _vp, _vp]),
    "ldsf_nbl_get_neigh": (_i, [_vp, _i, _vp, _i, _i, _vp, _vp]),
    "ldsf_nbl_num_lists": (_i, [_vp]),
    "ldsf_nbl_list_cutoff": (_d, [_vp, _i]),
    "ldsf_nbl_num_particles": (_i, [_vp]),
    "ldsf_nbl_total": (_l, [_vp, _i, _i]),
    "ldsf_nbl_flatten": (_i, [_vp, _i, _i, _vp, _vp]),
    "ldsf_layout_create": (_i, [_vp, _i, _d, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_layout_destroy": (None, [_vp]),
    "ldsf_layout_info": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "ldsf_layout_arrays": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_layout_copy": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ldsf_layout_assemble_forces": (_i, [_vp, _vp, _vp, _vp]),
    "ldsf_assemble_forces": (_i, [_i, _i, _vp, _vp, _vp, _vp]),
    "ldsf_plan_gather_fold": (_i, [_vp, _vp, _vp, _vp]),
    "ldsf_gradient_noalloc": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i,
                                   _vp]),
    "ldsf_server_create": (_i, [_vp, _vp, _vp]),
    "ldsf_server_num_lanes": (_i, [_vp]),
    "ldsf_server_submit": (_i, [_vp, _vp, _vp]),
    "ldsf_server_wait": (_i, [_vp, ctypes.c_longlong, _d, _vp]),
    "ldsf_server_stats": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "ldsf_server_destroy": (None, [_vp]),
}


class ServerConfig(ctypes.Structure):
    """ldsf_server_config (include/ldsf.h)."""
    _fields_ = [("precision", ctypes.c_int), ("hbm_budget_bytes", ctypes.c_size_t),
                ("lane_centers", ctypes.c_int), ("lane_rows", ctypes.c_int),
                ("lane_edges", ctypes.c_int), ("max_requests", ctypes.c_int),
                ("max_lanes", ctypes.c_int), ("dispatch_min_centers", ctypes.c_int),
                ("batch_window_us", ctypes.c_int)]


class Request(ctypes.Structure):
    """ldsf_request (include/ldsf.h)."""
    _fields_ = [("n_atoms", ctypes.c_int), ("n_total", ctypes.c_int), ("n_neigh", ctypes.c_int),
                ("max_neighbors", ctypes.c_int), ("species", _vp), ("number_of_neighs", _vp),
                ("neighbor_list", _vp), ("coordinates", _vp), ("dE_dzeta", _vp), ("desc", _vp),
                ("d_coordinates", _vp)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libdescriptor_amd: libldsf.so is not built; run "
                              "`python libdescriptor_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class LdsfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ldsf error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != LDSF_OK:
        raise LdsfError(rc, lib().ldsf_last_error().decode(errors="replace"))


def _ptr(t):
    """Data pointer of a torch tensor or a numpy array (None -> NULL)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


class Plan:
    """A device plan (ldsf_plan) over device-resident int32 species,
    neighbour list and neighbour counts.  `sf` is a DescriptorKind (kept
    alive with the plan) or a raw ldsf_sf handle (the caller keeps it alive).
    Keeps references to the tensors it borrows so they outlive the plan."""

    def __init__(self, sf, n_total, species, neighlist, numneigh, centers=None,
                 n_atoms=None, config_ptr=None, stream=None):
        import numpy as np
        sf_handle = sf._handle if hasattr(sf, "_handle") else sf
        self._keep = (sf, species, neighlist, numneigh, centers)
        out = ctypes.c_void_p()
        st = _stream_ptr(stream)
        n_centers = int(numneigh.numel()) if hasattr(numneigh, "numel") else len(numneigh)
        # the C-ABI trusts the neighbour list to hold sum(numneigh) entries
        total = int(numneigh.sum()) if n_centers else 0
        n_nl = int(neighlist.numel()) if hasattr(neighlist, "numel") else len(neighlist)
        n_sp = int(species.numel()) if hasattr(species, "numel") else len(species)
        if total > n_nl or n_sp < n_total:
            raise ValueError("neighbour list shorter than sum(numneigh), or species shorter "
                             "than n_total")
        if n_atoms is not None:
            na = np.ascontiguousarray(n_atoms, dtype=np.int32)
            cp = np.ascontiguousarray(config_ptr, dtype=np.int32)
            if len(cp) != len(na) or int(na.sum()) != n_centers:
                raise ValueError("n_atoms / config_ptr do not match numneigh")
            check(lib().ldsf_plan_create_batch(ctypes.byref(out), sf_handle, len(na), na.ctypes.data,
                                               cp.ctypes.data, n_total, _ptr(species),
                                               _ptr(neighlist), _ptr(numneigh), st))
        else:
            check(lib().ldsf_plan_create(ctypes.byref(out), sf_handle, n_centers, _ptr(centers),
                                         n_total, _ptr(species), _ptr(neighlist), _ptr(numneigh),
                                         st))
        self.handle = out.value
        self.n_total = n_total
        self.device = getattr(species, "device", None)

    @property
    def n_centers(self):
        return lib().ldsf_plan_num_centers(self.handle)

    @property
    def n_edges(self):
        return lib().ldsf_plan_num_edges(self.handle)

    @property
    def max_neighbors(self):
        return lib().ldsf_plan_max_neighbors(self.handle)

    @property
    def kernel_family(self):
        return lib().ldsf_plan_kernel_family(self.handle)

    @kernel_family.setter
    def kernel_family(self, fam):
        check(lib().ldsf_plan_set_kernel_family(self.handle, int(fam)))

    def set_records(self, records):
        """Take the adjoint's records (3 * num_records float64 values) from
        `records`, a contiguous float64 tensor on the plan's device, kept alive
        with the plan."""
        _check_records(records, self.device)
        check(lib().ldsf_plan_set_records(self.handle, _ptr(records), int(records.numel())))
        self._records = records

    @property
    def num_records(self):
        return lib().ldsf_plan_num_records(self.handle)

    def compute(self, coords, desc, precision=FP64, stream=None):
        check(lib().ldsf_plan_compute(self.handle, _ptr(coords), _ptr(desc), int(precision),
                                      _stream_ptr(stream)))

    def gradient(self, coords, dE_dzeta, dcoords, precision=FP64, stream=None):
        check(lib().ldsf_plan_gradient(self.handle, _ptr(coords), _ptr(dE_dzeta), _ptr(dcoords),
                                       int(precision), _stream_ptr(stream)))

    def adjoint_edges(self, coords, dE_dzeta, precision=FP64, stream=None):
        """First half of gradient(): the adjoint kernel alone."""
        check(lib().ldsf_plan_adjoint_edges(self.handle, _ptr(coords), _ptr(dE_dzeta),
                                            int(precision), _stream_ptr(stream)))

    def forward_gradient(self, coords, dE_dzeta, desc, dcoords, precision=FP64, stream=None):
        """The reference's gradient: desc += descriptors, dcoords += J^T dE_dzeta,
        from one fused kernel and the gather."""
        check(lib().ldsf_plan_forward_gradient(self.handle, _ptr(coords), _ptr(dE_dzeta),
                                               _ptr(desc), _ptr(dcoords), int(precision),
                                               _stream_ptr(stream)))

    def forward_gradient_assign(self, coords, dE_dzeta, desc, dcoords, precision=FP64,
                                stream=None):
        """forward_gradient() with the outputs assigned: desc = descriptors and
        dcoords = J^T dE_dzeta, bitwise the += into zeroed arrays."""
        check(lib().ldsf_plan_forward_gradient_assign(self.handle, _ptr(coords), _ptr(dE_dzeta),
                                                      _ptr(desc), _ptr(dcoords), int(precision),
                                                      _stream_ptr(stream)))

    def forward_adjoint_edges(self, coords, dE_dzeta, desc, precision=FP64, stream=None,
                              assign=False):
        """First half of forward_gradient(): the fused kernel alone (assign:
        desc = descriptors instead of +=)."""
        fn = lib().ldsf_plan_forward_adjoint_edges_assign if assign else \
            lib().ldsf_plan_forward_adjoint_edges
        check(fn(self.handle, _ptr(coords), _ptr(dE_dzeta), _ptr(desc), int(precision),
                 _stream_ptr(stream)))

    def param_gradient(self, coords, dE_dzeta, dparams, drcut=None, stream=None):
        """dparams += dE/d(parameters) (initDescriptor's flat layout), drcut +=
        dE/d(cutoff matrix); FP64, device tensors."""
        check(lib().ldsf_plan_param_gradient(self.handle, _ptr(coords), _ptr(dE_dzeta),
                                             _ptr(dparams), _ptr(drcut), _stream_ptr(stream)))

    def gather(self, dcoords, stream=None, assign=False):
        """Second half of gradient(): dcoords += the per-row sums of the records
        (assign: = them, rows without records 0)."""
        fn = lib().ldsf_plan_gather_assign if assign else lib().ldsf_plan_gather
        check(fn(self.handle, _ptr(dcoords), _stream_ptr(stream)))

    def gather_fold(self, layout, out=None, stream=None):
        """The gather and assemble_forces (neighbor.py:227-263) in one pass, for a
        plan over a DeviceLayout's rows: returns (n_contrib, 3), bitwise the
        gather into zeros followed by layout.assemble_forces."""
        import torch
        if out is None:
            out = torch.empty((layout.n_contrib, 3), dtype=torch.float64, device=self.device)
        if out.dtype != torch.float64 or not out.is_contiguous() or \
                out.shape != (layout.n_contrib, 3):
            raise ValueError("out must be a contiguous float64 tensor of shape (n_contrib, 3)")
        check(lib().ldsf_plan_gather_fold(self.handle, layout.handle, _ptr(out),
                                          _stream_ptr(stream)))
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib is not None:
            _lib.ldsf_plan_destroy(h)
            self.handle = None


def _check_records(records, device):
    """The record buffer the C side writes 3 * num_records float64 values
    into: a float64, contiguous tensor on the plan's device (the C-ABI sees
    only an address and a count)."""
    import torch
    if not isinstance(records, torch.Tensor):
        raise TypeError("records must be a torch tensor")
    if records.dtype != torch.float64:
        raise ValueError(f"records must be float64, not {records.dtype}")
    if not records.is_contiguous():
        raise ValueError("records must be contiguous")
    if not records.is_cuda or (device is not None and records.device != device):
        raise ValueError(f"records must be on the plan's device ({device}), not {records.device}")


class DeviceLayout:
    """The device layout producer (ldsf_layout): padding atoms and the KLIFF-order
    neighbour list of the contributing atoms, built on the GPU from device
    coordinates (the GPU twin of NeighborList, python_package/libdescriptor/
    neighbor.py:35-154).  After construction the arrays are torch tensors on
    the same device: coords (rows, 3) float64, species / image (rows,) int32,
    numneigh (n_contrib,) int32, neighlist (n_neigh,) int32."""

    def __init__(self, coords, species, cell, pbc, cutoff, stream=None):
        import numpy as np
        import torch
        if coords.dtype != torch.float64 or species.dtype != torch.int32:
            raise TypeError("coords must be float64 and species int32 device tensors")
        coords = coords.contiguous()
        species = species.contiguous()
        n = int(species.numel())
        if coords.numel() != 3 * n:
            raise ValueError("coords must have shape (n, 3) matching species")
        cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
        pbc = np.ascontiguousarray(pbc, dtype=np.int32).reshape(3)
        st = _stream_ptr(stream)
        out = ctypes.c_void_p()
        check(lib().ldsf_layout_create(ctypes.byref(out), n, float(cutoff), cell.ctypes.data,
                                       pbc.ctypes.data, _ptr(coords), _ptr(species), st))
        self.handle = out.value
        nc, npd, nn, mx = ctypes.c_int(), ctypes.c_int(), ctypes.c_long(), ctypes.c_int()
        check(lib().ldsf_layout_info(self.handle, ctypes.byref(nc), ctypes.byref(npd),
                                     ctypes.byref(nn), ctypes.byref(mx)))
        self.n_contrib, self.n_padding = nc.value, npd.value
        self.n_neigh, self.max_neighbors = nn.value, mx.value
        rows = self.n_contrib + self.n_padding
        dev = coords.device
        self.coords = torch.empty((rows, 3), dtype=torch.float64, device=dev)
        self.species = torch.empty(rows, dtype=torch.int32, device=dev)
        self.image = torch.empty(rows, dtype=torch.int32, device=dev)
        self.numneigh = torch.empty(self.n_contrib, dtype=torch.int32, device=dev)
        self.neighlist = torch.empty(max(self.n_neigh, 1), dtype=torch.int32, device=dev)
        check(lib().ldsf_layout_copy(self.handle, _ptr(self.coords), _ptr(self.species),
                                     _ptr(self.image), _ptr(self.numneigh), _ptr(self.neighlist),
                                     st))
        self.neighlist = self.neighlist[:self.n_neigh]

    def assemble_forces(self, forces, out=None, stream=None):
        """assemble_forces (neighbor.py:227-263) on the device: (rows, 3) -> (n_contrib, 3)."""
        import torch
        forces = forces.contiguous()
        if forces.dtype != torch.float64 or forces.shape != (self.n_contrib + self.n_padding, 3):
            raise ValueError("forces must be float64 of shape (rows, 3)")
        if out is None:
            out = torch.empty((self.n_contrib, 3), dtype=torch.float64, device=forces.device)
        check(lib().ldsf_layout_assemble_forces(self.handle, _ptr(forces), _ptr(out),
                                                _stream_ptr(stream)))
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib is not None:
            _lib.ldsf_layout_destroy(h)
            self.handle = None


def assemble_forces_device(forces, n, padding_image, stream=None):
    """assemble_forces (neighbor.py:227-263) for device tensors: forces (n + n_pad, 3)
    float64, padding_image (n_pad,) int32 -> (n, 3)."""
    import torch
    forces = forces.contiguous()
    padding_image = padding_image.contiguous()
    n_pad = int(padding_image.numel())
    if forces.dtype != torch.float64 or forces.shape != (n + n_pad, 3) or \
            padding_image.dtype != torch.int32:
        raise ValueError("forces must be float64 (n + n_pad, 3), padding_image int32")
    out = torch.empty((n, 3), dtype=torch.float64, device=forces.device)
    check(lib().ldsf_assemble_forces(int(n), n_pad, _ptr(forces), _ptr(padding_image), _ptr(out),
                                     _stream_ptr(stream)))
    return out


def _stream_ptr(stream):
    """hipStream_t of a torch stream (default: torch's current stream)."""
    if stream is None:
        try:
            import torch
            if torch.cuda.is_available():
                return torch.cuda.current_stream().cuda_stream
        except ImportError:
            pass
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Server:
    """The serving loop (ldsf_server): submit() queues a request over device
    tensors and returns a ticket; wait(ticket) blocks until it has finished
    and returns its latency in seconds (submission to completion).  The
    tensors of a request are kept alive until its wait."""

    def __init__(self, sf, precision=FP32_BULK, hbm_budget_bytes=2 << 30,
                 lane_centers=65536, lane_rows=196608, lane_edges=65536 * 48, max_requests=256,
                 max_lanes=4, dispatch_min_centers=0, batch_window_us=0):
        # sf: a DescriptorKind (kept alive with the server) or a raw ldsf_sf handle
        self._sf = sf
        sf_handle = sf._handle if hasattr(sf, "_handle") else sf
        cfg = ServerConfig(int(precision), int(hbm_budget_bytes), int(lane_centers),
                           int(lane_rows), int(lane_edges), int(max_requests), int(max_lanes),
                           int(dispatch_min_centers), int(batch_window_us))
        out = ctypes.c_void_p()
        check(lib().ldsf_server_create(ctypes.byref(out), sf_handle, ctypes.byref(cfg)))
        self.handle = out.value
        self._live = {}

    @property
    def num_lanes(self):
        return lib().ldsf_server_num_lanes(self.handle)

    def submit(self, species, numneigh, neighlist, coords, dE_dzeta, desc, dcoords,
               max_neighbors):
        """All tensors on the server's device: int32 species / numneigh /
        neighlist, float64 coords (n_total, 3), dE_dzeta and desc (n, width),
        dcoords (n_total, 3); dE_dzeta and dcoords may be None.  The server
        runs on its own streams: work that produces the inputs (or touches the
        outputs) on other streams must be finished before the call."""
        n, nt, ne = int(numneigh.numel()), int(species.numel()), int(neighlist.numel())
        for t, dt in ((species, "int32"), (numneigh, "int32"), (neighlist, "int32"),
                      (coords, "float64"), (desc, "float64")):
            if str(t.dtype).split(".")[-1] != dt or not t.is_contiguous():
                raise TypeError(f"expected a contiguous {dt} tensor")
        if coords.numel() != 3 * nt or desc.shape[0] != n:
            raise ValueError("request arrays do not match")
        req = Request(n, nt, ne, int(max_neighbors), _ptr(species), _ptr(numneigh),
                      _ptr(neighlist), _ptr(coords), _ptr(dE_dzeta), _ptr(desc), _ptr(dcoords))
        ticket = ctypes.c_longlong()
        check(lib().ldsf_server_submit(self.handle, ctypes.byref(req), ctypes.byref(ticket)))
        self._live[ticket.value] = (species, numneigh, neighlist, coords, dE_dzeta, desc, dcoords)
        return ticket.value

    def wait(self, ticket, timeout=-1.0):
        lat = ctypes.c_double()
        rc = lib().ldsf_server_wait(self.handle, int(ticket), float(timeout), ctypes.byref(lat))
        if rc != LDSF_ERR_TIMEOUT:  # finished (or unknown): its buffers are free again
            self._live.pop(int(ticket), None)
        check(rc)
        return lat.value

    def stats(self):
        v = [ctypes.c_longlong() for _ in range(4)]
        check(lib().ldsf_server_stats(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(("batches", "requests", "atoms", "graphs"), (x.value for x in v)))

    def close(self):
        h = getattr(self, "handle", None)
        if h and _lib is not None:
            _lib.ldsf_server_destroy(h)
        self.handle = None

    def __del__(self):
        self.close()
