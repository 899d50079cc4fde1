"""Python host side of the MI355X Gaussian-splatting renderer.

`gsr.lib` binds the C ABI (include/gsr.h) of libgsr.so; `gsr.scenes` builds
synthetic scenes and cameras in the reference's layouts; `gsr.build` compiles
the HIP library and the Node addon.
"""
from . import scenes  # noqa: F401
