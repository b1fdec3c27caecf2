"""doki_amd — MI355X-native state-vector core behind Doki's `doki` API.

``doki_amd.doki`` is the drop-in CPython extension (same functions, capsule
names and ``doki.error`` as the reference's ``doki`` module); it calls the C ABI
of ``libdoki_hip.so`` (declared in include/doki_hip.h), whose HIP kernels run
on gfx950.  Putting this folder on ``sys.path`` also makes ``import doki``
resolve to it.

Nothing here computes on the CPU: without the built library or without a
gfx950 device, calls raise ``doki.error`` rather than falling back.
"""
import glob
import os
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdoki_hip.so")
EXT_PATH = os.path.join(HERE, "doki" + sysconfig.get_config_var("EXT_SUFFIX"))


def built():
    return os.path.exists(LIB_PATH) and os.path.exists(EXT_PATH)


def _load():
    if not built():
        raise ImportError(
            "doki_amd native parts are not built: run `python -m doki_amd.build`")
    import importlib
    return importlib.import_module(__name__ + ".doki")


def __getattr__(name):
    if name == "doki":
        return _load()
    raise AttributeError(name)


def native_files():
    """Native files this package loads (for provenance checks)."""
    return [LIB_PATH, EXT_PATH] + glob.glob(os.path.join(HERE, "*.so"))
