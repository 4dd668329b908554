"""espins_amd -- MI355X-native engine for the ESpinS Metropolis sweep and
parallel-tempering replica exchange.

The compute path is ``libespins_hip.so`` (hand-written HIP for gfx950) behind the C ABI in
``include/espins_hip.h``; ``espins_amd.engine`` is its Python mirror and
``espins_amd/driver`` the Fortran driver.  Importing ``espins_amd.engine`` raises if the
library is not built: there is no CPU fallback.
"""
__version__ = "0.1.0"
