"""CPU FP64 oracle for the warpflow-les per-step hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``warpflow`` package
under ``warpflow-les_amd/``) imports, links or runs anything in this folder.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and there only as the checker (or, for
``cpu_baseline``, as the timed CPU restatement) -- never as the thing measured
on the GPU or shipped.

Two independent transcriptions of the reference algorithm live here:

* ``np_oracle``  -- NumPy, one function per reference kernel / method, each
  citing the reference ``file:line`` it restates.  Readable, used for
  operator-level parity and known-answer tests.
* ``c/wf_oracle.c`` -- plain C (+ OpenMP), the same schedule, used for long
  trajectories and timed as the CPU baseline.  Built by ``oracle/build.py``
  into ``oracle/_build/libwf_oracle.so``; loaded through ``c_oracle``.
"""
