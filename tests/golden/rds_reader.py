"""Minimal reader for R's XDR serialization format (``saveRDS`` files).

Reads gzip/xz/bzip2-compressed or plain RDS streams into plain Python values:
numeric vectors become lists of floats (NA -> None), integer/logical vectors
lists of ints (NA -> None), character vectors lists of str, generic vectors
``RList`` objects carrying their R attributes (``names``, ``dim``, ``class``,
``row.names`` ...). Only the subset of the format that data fixtures use is
supported (no closures or bytecode); ALTREP compact sequences and wrappers are
expanded. Nothing in a file is executed.
"""
import bz2
import gzip
import lzma
import struct

NA_INT = -2147483648


class RObj:
    """An R value with its attributes."""

    def __init__(self, rtype, value, attrs=None):
        self.rtype = rtype
        self.value = value
        self.attrs = attrs or {}

    def names(self):
        return self.attrs.get("names")

    def __getitem__(self, key):
        if isinstance(key, str):
            names = self.names() or []
            for i, nm in enumerate(names):
                if nm == key:
                    return self.value[i]
            raise KeyError(key)
        return self.value[key]

    def get(self, key, default=None):
        try:
            return self[key]
        except KeyError:
            return default

    def __repr__(self):
        return "RObj(%s, len=%s, attrs=%s)" % (
            self.rtype, len(self.value) if hasattr(self.value, "__len__") else "-",
            list(self.attrs))


def _decompress(raw):
    if raw[:2] == b"\x1f\x8b":
        return gzip.decompress(raw)
    if raw[:6] == b"\xfd7zXZ\x00":
        return lzma.decompress(raw)
    if raw[:3] == b"BZh":
        return bz2.decompress(raw)
    return raw


class _Reader:
    def __init__(self, data):
        self.d = data
        self.pos = 0
        self.refs = []

    def int(self):
        v = struct.unpack_from(">i", self.d, self.pos)[0]
        self.pos += 4
        return v

    def length(self):
        n = self.int()
        if n == -1:
            hi = self.int()
            lo = self.int()
            n = (hi << 32) + lo
        return n

    def doubles(self, n):
        v = struct.unpack_from(">%dd" % n, self.d, self.pos)
        self.pos += 8 * n
        return list(v)

    def ints(self, n):
        v = struct.unpack_from(">%di" % n, self.d, self.pos)
        self.pos += 4 * n
        return [None if x == NA_INT else x for x in v]

    def bytes_(self, n):
        v = self.d[self.pos:self.pos + n]
        self.pos += n
        return v

    def attrs_to_dict(self, pl):
        out = {}
        while pl is not None:
            tag, car, cdr = pl
            out[tag] = car
            pl = cdr
        return out

    def item(self):
        flags = self.int()
        t = flags & 0xFF
        has_attr = bool(flags & (1 << 9))
        has_tag = bool(flags & (1 << 10))
        if t == 254:  # NILVALUE
            return None
        if t in (253, 242, 241, 251, 252, 250):
            return None
        if t == 255:  # REFSXP
            idx = flags >> 8
            if idx == 0:
                idx = self.int()
            return self.refs[idx - 1]
        if t == 1:  # SYMSXP
            name = self.item()
            self.refs.append(name)
            return name
        if t in (2, 6, 239, 240, 17, 5, 3):  # pairlist-like
            attrs = self.item() if has_attr else None
            tag = self.item() if has_tag else None
            car = self.item()
            cdr = self.item()
            return (tag, car, cdr) if t in (2, 239) else ("<lang>", car, cdr)
        if t == 4:  # ENVSXP
            self.int()
            env = {}
            self.refs.append(env)
            self.item(); self.item(); self.item(); self.item()
            return env
        if t == 9:  # CHARSXP
            n = self.int()
            if n == -1:
                return None
            return self.bytes_(n).decode("utf-8", "replace")
        if t in (10, 13):
            n = self.length()
            v = self.ints(n)
            if t == 10:
                v = [None if x is None else bool(x) for x in v]
            return self._wrap(t, v, has_attr)
        if t == 14:
            n = self.length()
            return self._wrap(t, self.doubles(n), has_attr)
        if t == 15:
            n = self.length()
            d = self.doubles(2 * n)
            return self._wrap(t, [complex(d[2 * i], d[2 * i + 1]) for i in range(n)], has_attr)
        if t == 16:
            n = self.length()
            return self._wrap(t, [self.item() for _ in range(n)], has_attr)
        if t in (19, 20):
            n = self.length()
            return self._wrap(t, [self.item() for _ in range(n)], has_attr)
        if t == 24:
            n = self.length()
            return self._wrap(t, self.bytes_(n), has_attr)
        if t == 25:  # S4SXP
            return self._wrap(t, None, has_attr)
        if t == 238:  # ALTREP
            info = self.item()
            state = self.item()
            attr = self.item()
            cls = info[1] if isinstance(info, tuple) else None
            val = self._altrep(cls, state)
            obj = val if isinstance(val, RObj) else RObj(0, val)
            if attr is not None:
                obj.attrs.update(self.attrs_to_dict(attr))
            return obj
        raise ValueError("unsupported SEXP type %d at offset %d" % (t, self.pos))

    def _wrap(self, t, v, has_attr):
        attrs = self.attrs_to_dict(self.item()) if has_attr else {}
        for k, a in list(attrs.items()):
            if isinstance(a, RObj) and a.rtype in (16, 13, 14, 10, 0):
                attrs[k] = a.value
        return RObj(t, v, attrs)

    def _altrep(self, cls, state):
        if cls == "compact_intseq":
            n, start, inc = state.value
            return RObj(13, [int(start + i * inc) for i in range(int(n))])
        if cls == "compact_realseq":
            n, start, inc = state.value
            return RObj(14, [start + i * inc for i in range(int(n))])
        if cls and cls.startswith("wrap_"):
            return state.value[0]
        if cls == "deferred_string":
            arg = state[1] if isinstance(state, tuple) else state
            vals = arg.value if isinstance(arg, RObj) else arg
            return RObj(16, [None if v is None else _r_num_str(v) for v in vals])
        raise ValueError("unsupported ALTREP class %r" % cls)


def _r_num_str(v):
    if isinstance(v, float) and v.is_integer():
        return str(int(v))
    return repr(v)


def read_rds(path):
    with open(path, "rb") as fh:
        data = _decompress(fh.read())
    if data[:2] != b"X\n":
        raise ValueError("only XDR-format RDS files are supported")
    r = _Reader(data)
    r.pos = 2
    version = r.int()
    r.int(); r.int()
    if version == 3:
        n = r.int()
        r.bytes_(n)
    return r.item()
