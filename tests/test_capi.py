"""The C-ABI library loads and exports every symbol include/ldsf.h declares,
and its host-only parts (parameter sets, file reader, errors) behave as the
reference's.  CPU only: no call here touches a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import libdescriptor_amd as ld
from libdescriptor_amd import _capi
from conftest import GOLDEN, ROOT, behler, golden


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ldsf.h")).read()
    return re.findall(r"LDSF_API[^;(]*?\b(ldsf_\w+)\s*\(", text)


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 35
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes prototypes cover the header
    assert set(names) == set(_capi._PROTOS)


def _create(types, sizes, params, ns=1, rc=4.5, resolved=0):
    L = _capi.lib()
    h = ctypes.c_void_p()
    names = (ctypes.c_char_p * ns)(*[f"S{k}".encode() for k in range(ns)])
    tn = (ctypes.c_char_p * len(types))(*[t.encode() for t in types])
    sz = np.asarray(sizes, np.int32)
    par = np.asarray(params, np.float64)
    cut = np.full(ns * ns, rc)
    rc_ = L.ldsf_sf_create(ctypes.byref(h), ns, names, cut.ctypes.data, len(types), tn,
                           sz.ctypes.data, par.ctypes.data, resolved)
    return rc_, h.value


def test_parameter_set_widths():
    types, sizes, params = behler()
    rc, h = _create(types, sizes, params)
    assert rc == 0 and _capi.lib().ldsf_sf_width(h) == 24
    _capi.lib().ldsf_sf_destroy(h)
    rc, h = _create(types, sizes, params, ns=2, resolved=1)
    assert rc == 0 and _capi.lib().ldsf_sf_width(h) == 64
    _capi.lib().ldsf_sf_destroy(h)
    rc, h = _create(types, sizes, params, ns=3, resolved=1)
    assert _capi.lib().ldsf_sf_width(h) == 3 * 8 + 6 * 16
    _capi.lib().ldsf_sf_destroy(h)


def test_bad_type_is_an_error():
    rc, h = _create(["g7"], [1, 1], [0.1])
    assert rc == _capi.LDSF_OK + 1 and h is None
    assert b"g7" in _capi.lib().ldsf_last_error()
    with pytest.raises(ValueError):
        ld.DescriptorKind.init_descriptor(ld.AvailableDescriptors.SymmetryFunctions, ["Si"], "cos",
                                          4.5, ["g9"], [1, 1], [0.1])
    with pytest.raises(ValueError):
        ld.DescriptorKind.init_descriptor(ld.AvailableDescriptors.SOAP, 4, 3, 4.0, ["Si"],
                                          "polynomial", 0.1)


def test_file_reader_matches_reference_behaviour():
    pr = golden("params_ref.npz")
    f2 = os.path.join(GOLDEN, "sf_two_species.params")
    ds = ld.DescriptorKind.init_descriptor(f2, ld.AvailableDescriptors.SymmetryFunctions)
    assert ds.width == int(pr["width_two_species"])
    assert ds.species == ["Al", "Si"]  # sorted, as the reference's std::set
    assert ds.param_file == f2
    ds = ld.DescriptorKind.init_descriptor(os.path.join(GOLDEN, "sf_g2g4.params"),
                                           ld.AvailableDescriptors.SymmetryFunctions)
    assert ds.width == int(pr["width_g2g4"])
    with pytest.raises(RuntimeError):
        ld.DescriptorKind.init_descriptor("/nonexistent/file", ld.AvailableDescriptors.SymmetryFunctions)


def test_truncated_file_is_an_error(tmp_path):
    p = tmp_path / "bad.params"
    p.write_text("cos\n1\nSi Si 4.5\n2\ng2 2 2\n0.1 0.0\n")
    with pytest.raises(RuntimeError):
        ld.DescriptorKind.init_descriptor(str(p), ld.AvailableDescriptors.SymmetryFunctions)


def test_kim_get_neigh_pointer():
    assert ld.get_neigh_kim() is not None


def test_python_api_surface():
    # the reference module's names (python_bindings.cpp:50-556, libdescriptor.pyi)
    for name in ["AvailableDescriptors", "DescriptorKind", "NeighList", "compute",
                 "compute_single_atom", "compute_batch", "gradient", "gradient_single_atom",
                 "gradient_batch", "jacobian", "num_gradient_single_atom", "create",
                 "create_paddings", "get_neigh_kim", "NeighborList", "assemble_forces"]:
        assert hasattr(ld, name), name
    assert int(ld.AvailableDescriptors.SymmetryFunctions) == 0
    assert int(ld.AvailableDescriptors.Bispectrum) == 1


def test_cpp_drop_in_caller_links_against_the_library():
    """tests/cpp/drop_in.cpp (built by build()) includes the drop-in headers
    and resolves every C++ API symbol from libldsf.so."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "_build", "drop_in")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    out = subprocess.run(["ldd", "-r", exe], capture_output=True, text=True)
    assert out.returncode == 0
    assert "undefined symbol" not in out.stdout + out.stderr
    assert "not found" not in out.stdout
    # clone_empty (src/SymmetryFunctions.cpp:410-427) keeps the width, species,
    # cutoffs and descriptor count (host only; its zeroed parameters are checked
    # through the clone's descriptors in the GPU drop-in test)
    r = subprocess.run([exe, "--clone-check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stderr)


def test_record_buffer_checks():
    """Plan.set_records hands the C side only an address and a count: the
    Python side refuses any buffer the adjoint could overrun or misread."""
    import torch
    ok = torch.empty(12, dtype=torch.float64)
    with pytest.raises(TypeError):
        _capi._check_records(np.zeros(12), None)
    with pytest.raises(ValueError, match="float64"):
        _capi._check_records(torch.empty(12, dtype=torch.float32), None)
    with pytest.raises(ValueError, match="contiguous"):
        _capi._check_records(torch.empty((12, 2), dtype=torch.float64)[:, 0], None)
    with pytest.raises(ValueError, match="device"):
        _capi._check_records(ok, None)  # host memory
