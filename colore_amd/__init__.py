"""MI355X (gfx950) implementation of CoLoRe's field -> lognormal -> sources
path.  The compute lives in libcolore_amd.so (HIP, C ABI in
include/colore_amd.h); the host side (param files, cosmology tables,
writers and the `CoLoRe` driver) in libcolore_host.so.  This package only
binds them."""
from .lib import Context, Run, load_amd, load_host  # noqa: F401

__all__ = ["Context", "Run", "load_amd", "load_host"]
