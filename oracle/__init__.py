"""CPU oracle for the IGM ice-flow hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU (numpy / torch-CPU, float32 or float64),
the reference algorithm of ``igm/processes/iceflow`` and of the transport
step that consumes its output.  Each function cites the reference file:line
it follows.  It exists to check the HIP implementation in ``igm_amd``:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
  of ``bench.py`` may import it, and only as the checker (or the timed CPU
  baseline), never as the thing measured or shipped;
* nothing in ``igm_amd`` imports it; the product path fails loudly when its
  HIP library is missing.

The reference is TensorFlow 2.15 code and cannot be imported or run here
(TensorFlow is absent).  Autograd of this restatement plays the role of the
reference's ``tf.GradientTape``.  How each module is pinned is listed in its
own docstring, with the tests that do it (``tests/test_oracle_*.py``,
``tests/test_host_*.py``); a function without an independent pin says
"parity unpinned" in its docstring.
"""
