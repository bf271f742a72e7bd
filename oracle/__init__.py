"""CPU oracle for the SumCheck prover path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import, call, link or execute anything in this directory, and only as the checker or as the
timed CPU baseline. The product (``sumcheck-triton_amd/``) never imports it and has no CPU
fallback.

* ``sumcheck_ref`` — pure-Python restatement of the reference's big-integer prover
  (``src/zkduel/reference.py:44-110``), expression rules (``src/zkduel/_expr.py:27-110``),
  field constants and the kernels' truncated CIOS word routine (``src/zkduel/fields.py:63-121``),
  the SHA3 transcript (``benchmarks/field_sweep/triton_sweep.py:474-484``), the checksum
  (``cuda/src/field_sumcheck.cu:1043-1049``) and the affine inputs
  (``cuda/src/field_sumcheck.cu:123-128``).
* ``oracle.c`` / ``liboracle.so`` — multi-threaded C twin of the same algorithm for sizes the
  Python restatement cannot reach (num_vars 16..26), loaded by ``cbind``.
"""
