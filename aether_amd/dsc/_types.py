"""Element types and memory places of aether arrays -- plain Python, no numpy or torch import.

Every element type is one interned ``DType`` object, so ``a.dtype is DType("float32")`` holds and a
dtype compares equal to any spelling of it. What a caller may pass as ``dtype=`` is the public
interface of the reference package (dsc/aether/_dtype.py): a name ("float32"), one of Python's
``float`` / ``int`` / ``bool``, or an object that names a supported type -- a numpy dtype or scalar type,
a torch dtype -- recognised by its ``name`` / ``__name__`` attribute or its printed form, without
importing the library it comes from. Type codes are DLPack's (int 0, uint 1, float 2, bfloat 4, bool 6).
"""

# (name prefix, DLPack type code, widths in bits); "bool" is added below. The order is the one the
# error message lists.
_FAMILIES = (("float", 2, (64, 32, 16)), ("bfloat", 4, (16,)), ("int", 0, (64, 32, 16, 8)),
             ("uint", 1, (64, 32, 16, 8)))
_BUILTIN = {float: "float64", int: "int64", bool: "bool"}
_ALIAS = {"bool_": "bool"}  # numpy's scalar type name


class DType:
    """An element type: ``name``, DLPack ``code``, ``bits`` and ``itemsize``. One object per type."""

    __slots__ = ("name", "code", "bits")

    def __new__(cls, name):
        known = _BY_NAME.get(name)
        if known is None:
            raise _unsupported(name)
        return known

    @property
    def itemsize(self):
        return self.bits // 8

    def __eq__(self, other):
        if isinstance(other, DType):
            return other is self
        try:
            return resolve(other) is self
        except Exception:  # not something that names a type
            return False

    def __hash__(self):
        return hash(self.name)

    def __reduce__(self):
        return (DType, (self.name,))

    def __str__(self):
        return self.name

    def __repr__(self):
        return f"aether.DType({self.name!r})"


def _intern(name, code, bits):
    t = object.__new__(DType)
    object.__setattr__(t, "name", name)
    object.__setattr__(t, "code", code)
    object.__setattr__(t, "bits", bits)
    return t


_BY_NAME = {}
for _prefix, _code, _widths in _FAMILIES:
    for _w in _widths:
        _BY_NAME[f"{_prefix}{_w}"] = _intern(f"{_prefix}{_w}", _code, _w)
_BY_NAME["bool"] = _intern("bool", 6, 8)
NAMES = tuple(_BY_NAME)


def _unsupported(spec):
    from ._core import Error

    return Error(f"dtype {spec!r} is not supported ({' '.join(NAMES)})")


def _candidate_names(spec):
    """The names a dtype spelling may stand for, most specific first."""
    if isinstance(spec, str):
        return [spec]
    names = []
    try:
        builtin = _BUILTIN.get(spec)
    except TypeError:  # unhashable
        builtin = None
    if builtin:
        return [builtin]
    for attr in ("name", "__name__"):
        v = getattr(spec, attr, None)
        if isinstance(v, str):
            names.append(v)
    names.append(str(spec).rpartition(".")[2])  # "torch.float32" -> "float32"
    return names


def resolve(spec):
    """The DType a spelling names; ``aether.Error`` listing the supported names otherwise."""
    if isinstance(spec, DType):
        return spec
    for name in _candidate_names(spec):
        t = _BY_NAME.get(_ALIAS.get(name, name))
        if t is not None:
            return t
    raise _unsupported(spec)


class Device:
    """Where an array's memory lives: ``kind`` is "cpu", "pinned" (page-locked host memory) or "cuda" (a
    HIP device, named as torch names it), ``index`` the device ordinal."""

    __slots__ = ("kind", "index")
    # the seam's memory kinds, by position: CPU 0, DEVICE 1, PINNED 2 (dsc/src/seam.h:17)
    _SEAM_KINDS = ("cpu", "cuda", "pinned")

    def __init__(self, kind, index=0):
        if kind not in self._SEAM_KINDS:
            from ._core import Error

            raise Error(f"a device kind is 'cpu', 'pinned' or 'cuda', not {kind!r}")
        self.kind = kind
        self.index = int(index)

    @property
    def seam_kind(self):
        return self._SEAM_KINDS.index(self.kind)

    def __eq__(self, other):
        return isinstance(other, Device) and other.kind == self.kind and other.index == self.index

    def __hash__(self):
        return hash((self.kind, self.index))

    def __repr__(self):
        return f"Device({self.kind!r}, {self.index})"

    def __str__(self):
        return self.kind if self.kind != "cuda" else f"cuda:{self.index}"
