"""atlas_q -- MI355X-native adaptive-MPS hot path.

Drop-in for ATLAS-Q's ``AdaptiveMPS`` two-site update and batched sweep
sampler: same module and class names, same arguments and conventions, with
the arithmetic in hand-written gfx950 HIP kernels (``libatlasq.so``, C ABI
in ``include/atlasq.h``).  Nothing here runs on the CPU.
"""
__version__ = "0.1.0"


def get_adaptive_mps():
    """Export map of the reference's ``atlas_q.get_adaptive_mps()`` (__init__.py:47-66)."""
    from .adaptive_mps import AdaptiveMPS, DTypePolicy
    from .diagnostics import MPSStatistics, bond_entropy_from_S, effective_rank, spectral_gap
    from .linalg_robust import condition_number, robust_qr, robust_svd
    from .truncation import check_entropy_sanity, choose_rank_from_sigma, compute_global_error_bound
    return {
        "AdaptiveMPS": AdaptiveMPS,
        "DTypePolicy": DTypePolicy,
        "robust_svd": robust_svd,
        "robust_qr": robust_qr,
        "condition_number": condition_number,
        "choose_rank_from_sigma": choose_rank_from_sigma,
        "compute_global_error_bound": compute_global_error_bound,
        "check_entropy_sanity": check_entropy_sanity,
        "MPSStatistics": MPSStatistics,
        "bond_entropy_from_S": bond_entropy_from_S,
        "effective_rank": effective_rank,
        "spectral_gap": spectral_gap,
    }


def get_mps_pytorch():
    from .mps_pytorch import MatrixProductStatePyTorch
    return MatrixProductStatePyTorch
