"""pbi4_amd: the PBI4 measurement path (power-spectrum multipoles and bispectrum of particle
catalogues) on AMD Instinct MI355X.

    import pbi4_amd as pb      # drop-in for `import pbi4 as pb`
    io = pb.IO(config_file='config.json', verbose=True, scriptmode=True)
    ...

Submodules: pbi4 (IO / DENSITIES / MEASUREMENTS), powerI4 and bispectrumI4 (the f2py-named
functions), core (device-level API over the C ABI), pipeline and serving (batched, device-resident
measurement of many catalogues), catalogue (seeded synthetic catalogues), build (hipcc build).
"""
import os as _os

__version__ = "1.0.0"


def _rocfft_kernel_build():
    """rocFFT compiles its kernels at run time (hiprtc) and also takes prebuilt ones from a
    system kernel cache; built with default flags they contain packed FP32 VALU operations, which
    on this hardware have given wrong results beside matrix-core work.  Every rocFFT kernel this
    package runs is therefore built with packed FP32 off, into a kernel cache of the package's
    own, and none comes from the system cache.  rocFFT reads these settings when it first plans a
    transform in the process, so they are set here, on import, before torch or the HIP library
    is touched (libpbi4hip.so sets the same ones when it is loaded)."""
    env = _os.environ
    opt = "-Xclang -target-feature -Xclang -packed-fp32-ops"
    cur = env.get("HIPRTC_COMPILE_OPTIONS_APPEND", "")
    if "packed-fp32-ops" not in cur:
        env["HIPRTC_COMPILE_OPTIONS_APPEND"] = (cur + " " + opt).strip()
    env["ROCFFT_RTC_SYS_CACHE_PATH"] = "/nonexistent/pbi4_amd/no_system_kernel_cache.db"
    path = env.get("PBI4_ROCFFT_CACHE", "")
    if not path:
        d = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "_rocfft_cache")
        try:
            _os.makedirs(d, exist_ok=True)
            ok = _os.access(d, _os.W_OK | _os.X_OK)
        except OSError:
            ok = False
        if not ok:
            d = _os.path.join(env.get("TMPDIR") or "/tmp", f"pbi4_amd_rocfft_{_os.getuid()}")
            _os.makedirs(d, exist_ok=True)
        path = _os.path.join(d, "kernels.db")
        env["PBI4_ROCFFT_CACHE"] = path
    env["ROCFFT_RTC_CACHE_PATH"] = path


_rocfft_kernel_build()


def __getattr__(name):
    # the drop-in classes import torch and load the HIP library; defer that to first use
    if name in ("IO", "DENSITIES", "MEASUREMENTS", "NumpyArrayEncoder"):
        from . import pbi4 as _p
        return getattr(_p, name)
    raise AttributeError(name)
