"""CPU FP64 oracle: a NumPy restatement of WannierNLQG's dense-mesh k-loop.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline), never as the product path.  The product
path (``wanniernlqg_amd``) runs the HIP library and fails loudly without it.

Every function cites the reference file:line it restates.  Pinning of each
function (golden vectors, closed forms, invariants, brute force) is listed in
``oracle/kloop.py``'s module docstring.
"""
