"""FP64 CPU oracle for the RVGP connection-Laplacian GP path.

TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything from this package,
and only as the checker / the timed CPU baseline.  The product package
``rvgp_amd`` never imports it and has no CPU fallback.

Contents
  oracle/oracle.c   C restatement of the index work: FPS, kNN, the symmetrised
                    graph and the Fibonacci-heap neighbourhoods (heap pinned
                    against golden vectors written by the reference's own
                    compiled ptu_dijkstra.pyx, tests/golden/heap_*.npz).
  oracle/cpu.py     numpy/scipy restatement of everything else, each function
                    citing the reference lines it follows.

Parity status of every function is listed in the module docstring of cpu.py.
"""
