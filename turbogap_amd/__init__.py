"""turbogap_amd — MI355X-native SOAP descriptor and GAP energy/force path.

Host-side mirror of TurboGAP's soap_turbo / get_gap_soap call surface over the
HIP library libtgx (turbogap_amd/csrc, C ABI in include/tgx.h). Every compute
entry point runs on the GPU; importing works without a GPU, and a missing
libtgx.so raises at first use.
"""
from .hypers import SoapTurbo  # noqa: F401
from .lib import TgxError, load  # noqa: F401
from .context import Context  # noqa: F401

__all__ = ["SoapTurbo", "Context", "TgxError", "load"]
