"""MI355X-native Direct Electrifying (DEA) conductance solve.

Drop-in for the Jacobi-PCG solve of rnaimehaom/CICY_geom's Direct_Electrifying
(Direct_Electrifying.cpp:1538-1659) behind the C-ABI of include/cicy_dea.h:

* ``dea``                 -- ctypes binding of libcicy_dea.so (HIP, gfx950)
* ``direct_electrifying`` -- host mirror of Direct_Electrifying / Electrical_analysis
* ``network``             -- seeded CNT networks and the reference-rule system assembly
* ``warmstart``           -- learned warm start (message-passing network, HIP kernels)
* ``build``               -- in-tree build of the native libraries
"""
__all__ = ["dea", "direct_electrifying", "network", "build"]
