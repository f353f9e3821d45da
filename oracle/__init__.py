"""
Oracle: a CPU FP64 restatement of the reference D4 energy path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline). The product (``tad_dftd4_amd``) never
imports it and has no CPU fallback.

Modules
-------
``dense``
    Dense torch (CPU, float64 by default) restatement that mirrors the
    reference op for op, including the dependency stages the reference takes
    from tad-mctc 0.7.0 (``cn_d4``, ``erf_count``, masks) and tad-multicharge
    0.5.0 (``get_eeq_charges``). Gradients come from torch autograd, i.e. the
    reference's own gradient mechanism.
``brute``
    Pure-Python loop versions of the pair and triple sums for molecules with
    a handful of atoms, used to pin the dense restatement.

Every function in ``dense`` is pinned by at least one check in ``tests/``
(golden vectors from the reference's examples and fixtures, closed forms,
brute force, finite differences). The exceptions are listed as
"parity unpinned" in the module docstring of ``dense``.
"""
