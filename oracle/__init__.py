"""CPU oracle of the DPD hot path: test infrastructure only.

Imported by tests/, by the smoke check and by the CPU-baseline leg of the
benchmark; the product package (ctc_amd) never imports it.
"""
from .oracle import *  # noqa: F401,F403
