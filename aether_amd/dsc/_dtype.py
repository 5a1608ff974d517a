"""Element types by name and ``Device``, in plain Python (dsc/aether/_dtype.py:1-78 is the interface).

A dtype is named by a string ("float32"), a Python type (float, int, bool) or any object whose ``__name__``
or ``str()`` is such a name (numpy and torch dtypes included, without importing either). The codes are
DLPack's type codes (int 0, uint 1, float 2, bfloat 4, bool 6)."""

_TABLE = {
    "float64": (2, 64), "float32": (2, 32), "float16": (2, 16), "bfloat16": (4, 16),
    "int64": (0, 64), "int32": (0, 32), "int16": (0, 16), "int8": (0, 8),
    "uint64": (1, 64), "uint32": (1, 32), "uint16": (1, 16), "uint8": (1, 8), "bool": (6, 8),
}
_PYTHON = {float: "float64", int: "int64", bool: "bool"}


class DType:
    """An element type: ``name``, ``code`` (DLPack), ``bits``, ``itemsize``. Compares equal to its name."""
    __slots__ = ("name", "code", "bits")

    def __init__(self, name):
        self.name = name
        self.code, self.bits = _TABLE[name]

    @property
    def itemsize(self):
        return self.bits // 8

    def __eq__(self, other):
        if isinstance(other, DType):
            return other.name == self.name
        try:
            return parse(other).name == self.name
        except Exception:  # not a dtype spelling at all
            return False

    def __hash__(self):
        return hash(self.name)

    def __str__(self):
        return self.name

    def __repr__(self):
        return f"aether.DType({self.name!r})"


def _spellings(dtype):
    if isinstance(dtype, str):
        yield dtype
        return
    try:
        if dtype in _PYTHON:
            yield _PYTHON[dtype]
            return
    except TypeError:  # unhashable
        pass
    name = getattr(dtype, "__name__", None)
    if isinstance(name, str):
        yield name
    text = str(dtype)
    yield text
    if text.startswith("torch."):
        yield text[len("torch."):]


def parse(dtype):
    """The DType of a spelling; ``aether.Error`` naming the supported set otherwise."""
    if isinstance(dtype, DType):
        return dtype
    for s in _spellings(dtype):
        s = "bool" if s == "bool_" else s
        if s in _TABLE:
            return DType(s)
    from ._core import Error
    raise Error(f"dtype {dtype!r} is not supported ({' '.join(_TABLE)})")


class Device:
    """Where an array lives: ``Device(kind, index)``, kind "cpu", "pinned" or "cuda" (a HIP device)."""
    __slots__ = ("kind", "index")

    def __init__(self, kind, index=0):
        self.kind = kind
        self.index = int(index)

    def __eq__(self, other):
        return isinstance(other, Device) and (other.kind, other.index) == (self.kind, self.index)

    def __hash__(self):
        return hash((self.kind, self.index))

    def __repr__(self):
        return f"Device({self.kind!r}, {self.index})"

    def __str__(self):
        return f"cuda:{self.index}" if self.kind == "cuda" else self.kind
