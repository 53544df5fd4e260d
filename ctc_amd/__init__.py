"""ctc_amd: MI355X-native uDeviceX DPD hot path (HIP kernels behind a C ABI).

The compute path is libudx.so (ctc_amd/csrc, declared in include/udx.h); this
package only loads it and mirrors the reference's module interfaces for
Python callers.  Importing the bindings fails loudly when the library has not
been built: there is no CPU fallback.
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libudx.so")


def build():
    """Compile libudx.so for gfx950 (hipcc cross-compiles without a GPU)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(PKG_DIR, "csrc"), "-j8"], check=True)
    return LIB_PATH
