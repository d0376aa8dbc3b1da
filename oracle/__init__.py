"""CPU oracle for the electroweak RK4 hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker or as the timed CPU
baseline.

* ``oracle.cref``  -- ctypes binding of the literal C transcription
  (``ewpt_oracle.c``; gpu-kernels-rk4/main_rk4.jl:52-661, 6146-6441).
* ``oracle.rhs_np`` -- an independent numpy restatement in generator-matrix
  form, used to cross-check the C transcription.
* ``oracle.fields`` -- seeded smooth test states shared by both sides.
"""
