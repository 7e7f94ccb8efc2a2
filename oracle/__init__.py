"""CPU oracle for the jaqsi_amd simulation path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
directory; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Contents:

* ``sv``     -- NumPy complex128 restatement of the reference statevector
                engine (gate matrices, same-wire fusion, the reversed-plan
                adjoint and the diagonal expectation values).
* ``dm``     -- density-matrix simulation with Kraus channels.
* ``pulse``  -- pulse envelopes, the Magnus-2 propagator product and the
                process infidelity of the pulse-level optimal-control path.
* ``adam``   -- the plain Adam update used by the reference's golden
                training trajectories.
* ``cpu_baseline.cpp`` -- C++/OpenMP restatement used as the CPU timing
                baseline (built by ``oracle/Makefile``).

Every function cites the reference file:line it restates.  How each part
is pinned (golden vectors, closed forms, brute force) is stated in the
module headers; a function without such a check says "parity unpinned".
"""
