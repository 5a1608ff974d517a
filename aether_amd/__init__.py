"""aether_amd -- MI355X (gfx950) engine for aether's data-parallel hot path.

Host-side mirror of the reference interface over the C ABI of libaether_hip.so (include/aether_hip.h).
The torch-facing API (``Array``, ``Generator``, ``eval_cfg0``, ``reduce``, ...) lives in ``_api`` and the
submodules (``expr``, ``fft``, ``accum``, ``runtime``, ``serve``); both load on first attribute access, so
``import aether_amd.dsc`` (the reference's torch-free ``aether`` array package) never imports torch.

Layout: an ``Array`` of ``ncomp`` components and ``n`` samples is a device buffer of ``ncomp`` planes of
``capacity`` elements, capacity = n rounded up to 32 (array/Array.h:80,100-110,279-295); element (c, i) is
at c*capacity + i.
"""
import importlib as _importlib
import os as _os

_HERE = _os.path.dirname(__file__)


def _is_submodule(name: str) -> bool:
    return _os.path.isfile(_os.path.join(_HERE, name + ".py")) or _os.path.isfile(
        _os.path.join(_HERE, name, "__init__.py"))


def __getattr__(name):
    if name.startswith("__"):
        raise AttributeError(name)
    if _is_submodule(name):
        return _importlib.import_module(f"{__name__}.{name}")
    api = _importlib.import_module(f"{__name__}._api")
    try:
        value = getattr(api, name)
    except AttributeError:
        raise AttributeError(f"module {__name__!r} has no attribute {name!r}") from None
    globals()[name] = value
    return value
