"""The engine's Cython binding (hfsolver_amd._hfx, built from _hfx.pyx and
c_hfx.pxd in the style of hfsolver/special.pyx over hfsolver/c_hfsolver.pxd).

torch for ROCm ships its own HIP runtime under the same soname as /opt/rocm's;
it is imported first so that libhfx.so, which the extension links, binds to it
and torch device buffers can be handed to the engine. There is no fallback: a
missing extension raises ImportError naming the build step.
"""
try:
    import torch  # noqa: F401
except ImportError:
    pass

try:
    from ._hfx import Engine, HfxError, Stats  # noqa: F401
except ImportError as e:
    raise ImportError("the Cython binding hfsolver_amd/_hfx is not built (python -m hfsolver_amd.build): %s" % e)
