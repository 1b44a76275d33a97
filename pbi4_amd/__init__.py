"""pbi4_amd: the PBI4 measurement path (power-spectrum multipoles and bispectrum of particle
catalogues) on AMD Instinct MI355X.

    import pbi4_amd as pb      # drop-in for `import pbi4 as pb`
    io = pb.IO(config_file='config.json', verbose=True, scriptmode=True)
    ...

Submodules: pbi4 (IO / DENSITIES / MEASUREMENTS), powerI4 and bispectrumI4 (the f2py-named
functions), core (device-level API over the C ABI), pipeline and serving (batched, device-resident
measurement of many catalogues), catalogue (seeded synthetic catalogues), build (hipcc build).
"""
__version__ = "1.0.0"


def __getattr__(name):
    # the drop-in classes import torch and load the HIP library; defer that to first use
    if name in ("IO", "DENSITIES", "MEASUREMENTS", "NumpyArrayEncoder"):
        from . import pbi4 as _p
        return getattr(_p, name)
    raise AttributeError(name)
