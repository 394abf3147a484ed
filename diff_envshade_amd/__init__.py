"""MI355X-native (gfx950) brute-force environment-light shader.

Drop-in for zjumsj/diff_envshade's shader bindings:

    from diff_envshade_amd import _diff_envshade      # same names / argument orders
    from diff_envshade_amd import autograd as des_autograd  # torch.autograd.Function wrappers

The compute path is hand-written HIP (csrc/) behind the C ABI of
include/diff_envshade.h; build it with `python -m diff_envshade_amd.build`.
"""
from . import _diff_envshade, autograd  # noqa: F401
from ._lib import lib_path, load as load_library  # noqa: F401

__version__ = "0.1.0"
