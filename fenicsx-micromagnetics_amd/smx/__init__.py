"""smx: MI355X-native Fredkin-Koehler demag + LLG hot path (HIP kernels behind a C ABI).

Public names follow the reference's GPU backend: ``DemagFieldLindholmGPU``,
``EffectiveField``, ``LLG_GPU`` and ``HMatrixGPUPackedFused``.
"""
from .mesh import TetMesh, kuhn_box, cube_c1, cube_c2, film_c3, disk  # noqa: F401

__all__ = ["TetMesh", "kuhn_box", "cube_c1", "cube_c2", "film_c3", "disk",
           "DemagFieldLindholmGPU", "EffectiveField", "LLG_GPU"]


def __getattr__(name):
    if name == "DemagFieldLindholmGPU":
        from .demag import DemagFieldLindholmGPU
        return DemagFieldLindholmGPU
    if name in ("EffectiveField", "LLG_GPU"):
        from . import llg
        return getattr(llg, name)
    if name == "HMatrixGPUPackedFused":
        from .hmatrix import HMatrixGPUPackedFused
        return HMatrixGPUPackedFused
    raise AttributeError(name)
