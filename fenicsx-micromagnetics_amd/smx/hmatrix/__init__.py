"""Lindholm H-matrix: cluster tree, device construction, packed device matvec."""
from .tree import ClusterNode, ClusterTree, admissible, bbox_distance, build_cluster_tree  # noqa: F401
from .storage import HMatrixCPUData, LowRankBlockCPU  # noqa: F401
from .builder import HMatrixBuildConfig, build_hmatrix_device, build_or_load_hmatrix  # noqa: F401
from .backend import HMatrixGPUPackedFused  # noqa: F401
from .provider import LindholmEntryProviderGPU  # noqa: F401


def build_device_hmatrix(geo, device, epsilon=1e-6, eta=2.0, leaf_size=64, max_rank=None,
                         max_temporary_block_bytes=256 * 1024 ** 2, use_fp32=True, cache_path=None,
                         force_rebuild=False, verbose=False, near_fp32=False, compressor="fullaca",
                         near_in_z="auto", **gpu_options):
    """Build (or load) the H-matrix of a boundary and pack it on the device with
    the boundary gather of the demag path.  gpu_options: the gpu_* fields of
    HMatrixBuildConfig."""
    cfg = HMatrixBuildConfig(epsilon=epsilon, eta=eta, leaf_size=leaf_size, max_rank=max_rank,
                             max_temporary_block_bytes=max_temporary_block_bytes, compressor=compressor,
                             **gpu_options)
    data = build_or_load_hmatrix(cache_path, geo, config=cfg, device=device, force_rebuild=force_rebuild,
                                 verbose=verbose)
    gpu = HMatrixGPUPackedFused(data, use_fp32=use_fp32, device=device, near_fp32=near_fp32,
                                boundary_gather=geo.bverts, near_in_z=near_in_z)
    return data, gpu
