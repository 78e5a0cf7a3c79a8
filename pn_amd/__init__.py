"""pn_amd: MI355X (gfx950) implementation of the Limit-sticky9 force/energy path.

The compute lives in libpn1.so (HIP kernels + C ABI, include/pn1.h);
this package only binds it (pn_amd.lib) and builds it (pn_amd.build).
"""
from .lib import (FP64, MIXED, Context, Ensemble, Pn1Error, check_input, ic_plummer,  # noqa: F401
                  ic_uniform_sphere, legacy_force, load, read_checkpoint, read_dat,
                  set_c, set_mode, write_checkpoint, write_snapshot)
