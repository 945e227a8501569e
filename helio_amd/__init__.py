"""hv: MI355X-native GPU-driven visibility (instance cull + LOD, Hi-Z pyramid,
Hi-Z occlusion, meshlet cull, compaction into DrawIndexedIndirect args).

The compute path is libhv.so (hand-written HIP kernels for gfx950 behind the C
ABI of include/hv/hv.h); this package is its Python host side.
"""
__all__ = ["layouts", "passes", "synth"]
