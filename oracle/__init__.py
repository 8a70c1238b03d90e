"""CPU oracle for the CT splatting path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or, for the
baseline, as the thing timed on the host). The product path in ``r2splat_amd``
never calls it: it runs the HIP library or raises.

Contents:
  * ``splat``  — float64 PyTorch (CPU) restatement of the operator mathematics,
                 differentiable with autograd, plus the exact integer tile
                 binning (numpy);
  * ``tiled_cpu.c`` — C/OpenMP restatement of the tiled rasterizer and its
                 VJP used as the timed CPU baseline (built by ``build.py``).
"""
