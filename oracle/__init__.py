"""CPU oracle for the Schmidt-form (Gamma-Lambda) MPS update path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path may import, call or
execute anything in this package: only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker
or the timed CPU baseline, never as the thing measured or shipped.

Contents
--------
``mps_oracle``  literal numpy/LAPACK restatement of tensor-net's ``MPS``
                (tensor-net/lib/mps.rs), SVD through LAPACK ``zgesvd`` with
                JOBU=JOBVT='A' exactly as ``svd_into(true, true)`` does.
``dense``       dense state-vector brute force (independent of the MPS
                algebra) used to pin the oracle: Schmidt spectra, Born
                probabilities, projections, entropies for N <= ~16.

Parity pins (what fixes each oracle function independently of itself) are
listed in the ``tests/test_oracle_pins.py`` docstrings.  Every function here
is pinned by at least one of: a closed form, a dense brute force, or an
invariant chosen so that a dropped term, wrong index or transposed operand
fails it.
"""
