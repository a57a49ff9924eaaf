"""torch_amd -- MI355X-native direct-sum gravity field for Torch's gas/star bridge.

The product is libtgrav.so (hand-written HIP for gfx950 behind the C ABI in
include/tgrav.h); this package is its Python host side:

    from torch_amd import StellarGravityField, GasGravityField
    f = StellarGravityField(mode="mixed")
    f.set_stars(x, y, z, m, eps)
    ax, ay, az = f.get_gravity_at_point(eps_t, xt, yt, zt)

There is no CPU fallback: without the built library every call raises.
"""
from ._lib import TgravError, MIXED, PLAIN_FP64
from .field import G_PC_MSUN_KMS, GasGravityField, GravityField, StellarGravityField
from .pm import LeafBlocks, ParticleMesh

__all__ = ["GravityField", "StellarGravityField", "GasGravityField", "ParticleMesh", "LeafBlocks",
           "TgravError", "G_PC_MSUN_KMS", "MIXED", "PLAIN_FP64"]
