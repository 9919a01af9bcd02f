"""ORACLE -- test infrastructure only.

CPU restatement of libdescriptor's Behler symmetry-function path, used as the
checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
The product (libdescriptor_amd) never imports this package.

* ``oracle.orc``  -- ctypes wrapper of ``sf_oracle.c`` (FP64 forward in the
  reference loop order, hand-derived adjoint, order-8 FD check).
* ``oracle.nbl``  -- numpy restatement of the KLIFF neighbour list order,
  periodic padding and ghost-gradient fold.
* ``oracle.ref``  -- ctypes wrapper of ``_ref/libsfref.so``, the reference's own
  forward/neighbour list compiled from its sources (present only where the
  reference tree was available at build time).

Parity status of each function is recorded in its docstring.
"""
