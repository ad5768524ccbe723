"""fopis_amd — MI355X-native electrostatic PIC step of hsimyu/fopis (TDPIC).

The compute path is libfopis_hip.so (hand-written gfx950 HIP kernels behind
the C ABI in include/fopis.h). `fopis_amd.lib` is its ctypes binding;
`fopis_amd.build` compiles it in-tree with hipcc.
"""
from . import lib  # noqa: F401

__all__ = ["lib"]
