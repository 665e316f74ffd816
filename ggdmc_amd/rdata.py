"""Reader for R's `.rda` save files (gzip + "RDX3" + XDR serialization v3).

Only what the reference's likelihood fixtures use is supported: S4 objects,
generic vectors, atomic vectors (logical/integer/double/character), pairlists
with attributes and tags, symbol/reference tables and the ALTREP compact
sequences and wrappers that R >= 3.5 writes. Environments, closures and
bytecode are skipped structurally (the fixtures hold none on the path we read).

Format: R Internals manual, "Serialization Formats"; the top-level object of a
`save()` file is a tagged pairlist of (variable name, value).
"""
import gzip
import struct

_NILVALUE = 254
_REFSXP = 255
_GLOBALENV, _UNBOUNDVALUE, _MISSINGARG, _BASENAMESPACE, _EMPTYENV, _BASEENV = (
    253, 252, 251, 247 + 0, 242, 241)
_SPECIAL_NO_PAYLOAD = {253, 252, 251, 250, 242, 241}
_PERSIST, _PACKAGE, _NAMESPACE = 247, 248, 249
_ALTREP = 238
_ATTRLANG, _ATTRLIST = 240, 239


class RValue:
    """A decoded R object: `kind` is one of 'real', 'int', 'lgl', 'str',
    'list', 'S4', 'pairlist', 'lang', 'sym', 'other'; `value` holds the payload
    (a python list for vectors); `attrs` maps attribute names to RValues."""

    __slots__ = ("kind", "value", "attrs")

    def __init__(self, kind, value=None, attrs=None):
        self.kind = kind
        self.value = value
        self.attrs = attrs or {}

    def attr(self, name):
        return self.attrs.get(name)

    def names(self):
        n = self.attrs.get("names")
        return list(n.value) if n is not None else None

    def dim(self):
        d = self.attrs.get("dim")
        return list(d.value) if d is not None else None

    def slot(self, name):
        """S4 slots are stored as attributes."""
        return self.attrs[name]

    def __getitem__(self, key):
        if isinstance(key, str):
            return self.value[self.names().index(key)]
        return self.value[key]

    def __len__(self):
        return len(self.value)


class _Stream:
    def __init__(self, buf):
        self.buf = buf
        self.pos = 0
        self.refs = []

    def int32(self):
        (v,) = struct.unpack_from(">i", self.buf, self.pos)
        self.pos += 4
        return v

    def length(self):
        n = self.int32()
        if n == -1:
            hi = self.int32()
            lo = self.int32()
            n = (hi << 32) + lo
        return n

    def raw(self, n):
        b = self.buf[self.pos:self.pos + n]
        self.pos += n
        return b

    def read_header(self):
        magic = self.raw(5)
        if magic not in (b"RDX3\n", b"RDX2\n"):
            raise ValueError("not an R XDR save file")
        if self.raw(2) != b"X\n":
            raise ValueError("only XDR (big-endian) serialization is supported")
        version = self.int32()
        self.int32()  # writer R version
        self.int32()  # minimal reader version
        if version == 3:
            n = self.int32()
            self.raw(n)  # native encoding name

    def _pairlist_attrs(self, pl):
        out = {}
        while isinstance(pl, RValue) and pl.kind == "pairlist":
            tag, car, cdr = pl.value
            out[tag] = car
            pl = cdr
        return out

    def item(self):
        flags = self.int32()
        sxptype = flags & 0xFF
        has_attr = bool(flags & (1 << 9))
        has_tag = bool(flags & (1 << 10))

        if sxptype == _NILVALUE:
            return None
        if sxptype in _SPECIAL_NO_PAYLOAD:
            return RValue("other", sxptype)
        if sxptype == _REFSXP:
            idx = flags >> 8
            if idx == 0:
                idx = self.int32()
            return self.refs[idx - 1]
        if sxptype in (_PERSIST, _PACKAGE, _NAMESPACE):
            self.int32()
            n = self.int32()
            obj = RValue("other", [self.item() for _ in range(n)])
            self.refs.append(obj)
            return obj
        if sxptype == 1:  # SYMSXP: a CHARSXP follows
            name = self.item()
            sym = name if isinstance(name, str) else ""
            self.refs.append(sym)
            return sym
        if sxptype in (2, 3, 5, 6, 17, _ATTRLIST, _ATTRLANG):
            attrs = None
            if has_attr:
                attrs = self._pairlist_attrs(self.item())
            tag = self.item() if has_tag else None
            car = self.item()
            cdr = self.item()
            kind = "lang" if sxptype in (6, _ATTRLANG) else "pairlist"
            return RValue(kind, (tag, car, cdr), attrs)
        if sxptype == 4:  # ENVSXP
            env = RValue("other", None)
            self.refs.append(env)
            self.int32()  # locked
            self.item()  # enclosure
            self.item()  # frame
            self.item()  # hash table
            self.item()  # attributes
            return env
        if sxptype == 9:  # CHARSXP
            n = self.int32()
            if n == -1:
                return None  # NA_character_
            return self.raw(n).decode("utf-8", "replace")
        if sxptype in (10, 13):
            n = self.length()
            vals = list(struct.unpack_from(">%di" % n, self.buf, self.pos))
            self.pos += 4 * n
            obj = RValue("lgl" if sxptype == 10 else "int", vals)
        elif sxptype == 14:
            n = self.length()
            vals = list(struct.unpack_from(">%dd" % n, self.buf, self.pos))
            self.pos += 8 * n
            obj = RValue("real", vals)
        elif sxptype == 16:
            n = self.length()
            obj = RValue("str", [self.item() for _ in range(n)])
        elif sxptype in (19, 20):
            n = self.length()
            obj = RValue("list", [self.item() for _ in range(n)])
        elif sxptype == 24:
            n = self.length()
            obj = RValue("other", self.raw(n))
        elif sxptype == 25:
            obj = RValue("S4", None)
        elif sxptype in (7, 8):  # SPECIALSXP / BUILTINSXP
            n = self.int32()
            return RValue("other", self.raw(n).decode())
        elif sxptype == 22:  # EXTPTRSXP
            obj = RValue("other", None)
            self.refs.append(obj)
            self.item()
            self.item()
            if has_attr:
                obj.attrs = self._pairlist_attrs(self.item())
            return obj
        elif sxptype == 23:  # WEAKREFSXP
            obj = RValue("other", None)
            self.refs.append(obj)
            if has_attr:
                obj.attrs = self._pairlist_attrs(self.item())
            return obj
        elif sxptype == _ALTREP:
            return self._altrep()
        else:
            raise NotImplementedError("SEXP type %d at byte %d" % (sxptype, self.pos))
        if has_attr:
            obj.attrs = self._pairlist_attrs(self.item())
        return obj

    def _altrep(self):
        info = self.item()
        state = self.item()
        attr = self.item()
        cls = info.value[1]  # (tag, car=class symbol, cdr)
        if cls == "compact_intseq":
            n, start, inc = state.value
            obj = RValue("int", [int(start + k * inc) for k in range(int(n))])
        elif cls == "compact_realseq":
            n, start, inc = state.value
            obj = RValue("real", [start + k * inc for k in range(int(n))])
        elif cls.startswith("wrap_"):
            inner = state.value[1] if state.kind == "pairlist" else state
            if state.kind == "pairlist":
                inner = state.value[1]
            obj = RValue(inner.kind, inner.value, dict(inner.attrs))
        elif cls == "deferred_string":
            arg = state.value[1] if state.kind == "pairlist" else state
            obj = RValue("str", [str(x) for x in arg.value])
        else:
            raise NotImplementedError("ALTREP class " + cls)
        if attr is not None:
            obj.attrs.update(self._pairlist_attrs(attr))
        return obj


def load_rda(path):
    """Return {variable name: RValue} for an `.rda` file written by save()."""
    with gzip.open(path, "rb") as f:
        s = _Stream(f.read())
    s.read_header()
    top = s.item()
    out = {}
    while isinstance(top, RValue) and top.kind == "pairlist":
        tag, car, cdr = top.value
        out[tag] = car
        top = cdr
    return out
