"""pokerrust_amd: gfx950 (MI355X) CFR+ solver for the pokerrust hot path.

- drop-in: the seven C-ABI symbols of the reference's river layer
  (src/cuda_interface.rs:8-31), mirrored host-side in `pokerrust_amd.dropin`;
- solver: batched hand-level river/turn/flop subgame CFR+ (`pokerrust_amd.solver`).

All compute runs in libpokerrust_hip.so; there is no CPU fallback.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
