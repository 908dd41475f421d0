"""FP64 CPU oracle for the StoPEPS.jl Born-sampling hot path.

TEST INFRASTRUCTURE ONLY. This package is a plain-NumPy restatement of the
reference algorithms (StoPEPS.jl v0.2.0, files cited per function). It exists
to check the HIP product path. Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or the timed CPU baseline; the product library never calls it.

Representation: a Z2-graded tensor is a dense ``numpy`` array plus one parity
vector per leg (``repr`` of TensorKit's ``Z2Space``). Dense (``ComplexSpace``)
tensors are the special case where every parity is 0. Symmetric tensors are
neutral: an entry can be nonzero only when the parities of its indices sum to
0 mod 2. Factorizations act on the two parity blocks separately, which is
what TensorKit does for ``Z2Space`` block storage.

Parity pins (what fixes each part independently of this restatement) are in
``tests/test_oracle_*.py``. Functions without such a pin say so in their
docstring with the words "parity unpinned".
"""
