"""4dzb_amd — MI355X (gfx950) soft-shadow visibility path of galaxy-wolf/4DZB.

The product is the C-ABI library ``libfd4d.so`` (include/fd4d.h) built from
``csrc/``; ``fd4d`` is its ctypes binding, ``workloads`` the deterministic
input generators and ``autotune`` the layout tuner of the batched and served
forms.  Import this package with
``importlib.import_module("4dzb_amd")`` (the name starts with a digit).
"""
from . import autotune, fd4d, workloads  # noqa: F401

__all__ = ["autotune", "fd4d", "workloads"]
