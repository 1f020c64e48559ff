"""CPU oracle for the adaptive-MPS hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
directory: only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` may use it, and only as the checker / the timed CPU
baseline, never as the thing measured on the GPU.

Contents
--------
``ref_mps``      torch-CPU restatement of ATLAS-Q's ``AdaptiveMPS`` two-site
                 update, rank rule, canonical forms and batched sweep sampler
                 (op for op, so float rounding follows the reference's).
``statevector``  exact NumPy state-vector simulator for n <= 20 (independent
                 mathematical check of the MPS path).
``mlp_ref``      NumPy forward/backward/AdamW of the rank-predictor MLP.
``svd_ref``      FP64 NumPy reference for the truncated-SVD variants.

Parity pinning (what checks each function) is listed in each module header.
"""
