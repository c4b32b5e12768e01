"""MI355X-native Fourier structure-factor path of DISCUS (DiffuseCode).

The compute path is libdiffuse_turbo.so (HIP kernels for gfx950 behind a C
ABI, include/diffuse_turbo.h); this package holds its Python binding
(``_lib``), the host mirror of DISCUS's ``fourier`` command set
(``fourier``), the seeded crystal generators (``crystal``) and the build
(``build``).
"""
from . import crystal  # noqa: F401

__all__ = ["crystal", "fourier", "_lib", "build"]


def __getattr__(name):
    if name in ("fourier", "_lib", "build"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
