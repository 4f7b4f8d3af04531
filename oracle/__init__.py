"""CPU oracle for the EdgeFEM frequency-domain Nedelec path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``edgefem_amd`` package and its
HIP library) imports, links or runs anything in this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it, and
only as the checker or as the timed CPU baseline.

Every function restates the reference algorithm it names (file:line in the
EdgeFEM sources).  The arithmetic library of the reference (Eigen >= 3.4) is absent
from this machine; where the reference calls into it (sparse triplet assembly, the
dense generalised eigensolver, BiCGSTAB, IncompleteLUT, SparseLU) the oracle restates
the published algorithm or uses scipy's equivalent routine, as noted per function.

Parity pins (what checks each function independently of its own formula) are listed
in the header of ``oracle/edgefem_oracle.py``.
"""
