"""
Value ranges of generated C statements.

The fused kernel runs ``depth`` steps per launch, and only the colour left
by the last of them is ever read.  When the colour a step writes does not
depend on the colour before it, the colour of the earlier steps can be
left out.  For a plain :class:`~xentica_amd.core.color_effects.ColorEffect`
that always holds.  For ``MovingAverage`` (the stored colour moves towards
the new one by at most ``FADE_IN`` up and ``FADE_OUT`` down,
``xentica/core/color_effects.py:106-130``) it holds when every value a
channel can take lies in ``[lo, hi]``, ``FADE_IN`` and ``FADE_OUT`` are at
least ``hi - lo``, and the colour stored before the launch lies in
``[hi - FADE_IN, lo + FADE_OUT]``.  Then the clamp returns the new value at
every step, so the colour after the launch is the last step's new value.

:func:`channel_bounds` finds ``[lo, hi]`` by interval arithmetic over the
generated statements, with C's integer types, promotions and conversions
(LP64, two's complement, as clang compiles for AMDGPU and for the host).
A value it cannot follow (an unknown function, a reinterpreting cast, a
possible overflow) spans its whole type; a statement of a shape it does
not know makes the analysis give up.
"""
import math
import re

__all__ = ['Value', 'CType', 'Analysis', 'channel_bounds', 'type_named',
           'Unknown', 'GiveUp', 'CTYPES']


class Unknown(Exception):
    """An expression whose value the analysis cannot follow."""


class GiveUp(Exception):
    """A statement of unknown shape: no result at all."""


class CType:
    """An arithmetic C type: width, signedness, floating point."""

    def __init__(self, name, bits, signed, floating=False):
        self.name = name
        self.bits = bits
        self.signed = signed
        self.floating = floating
        if floating:
            self.lo, self.hi = -math.inf, math.inf
        elif signed:
            self.lo, self.hi = -(1 << (bits - 1)), (1 << (bits - 1)) - 1
        else:
            self.lo, self.hi = 0, (1 << bits) - 1

    def __repr__(self):
        return self.name


_BOOL = CType("bool", 1, False)
_INT = CType("int", 32, True)
_UINT = CType("unsigned int", 32, False)
_LONG = CType("long", 64, True)
_ULONG = CType("unsigned long", 64, False)
_FLOAT = CType("float", 32, True, True)
_DOUBLE = CType("double", 64, True, True)

#: Type names (keyword sets, ``int`` implied where C allows) of the types
#: the generated code declares.
CTYPES = {
    ("bool", ): _BOOL,
    ("char", ): CType("char", 8, True),
    ("signed", "char"): CType("signed char", 8, True),
    ("unsigned", "char"): CType("unsigned char", 8, False),
    ("short", ): CType("short", 16, True),
    ("unsigned", "short"): CType("unsigned short", 16, False),
    ("int", ): _INT,
    ("signed", ): _INT,
    ("unsigned", ): _UINT,
    ("long", ): _LONG,
    ("long", "long"): _LONG,
    ("unsigned", "long"): _ULONG,
    ("unsigned", "long", "long"): _ULONG,
    ("__int128", ): CType("__int128", 128, True),
    ("unsigned", "__int128"): CType("unsigned __int128", 128, False),
    ("float", ): _FLOAT,
    ("double", ): _DOUBLE,
}
_TYPE_WORDS = {"bool", "char", "signed", "unsigned", "short", "int", "long",
               "float", "double", "__int128"}

# float results are widened outward by this relative amount (a float32
# rounding is 2^-24 relative) and by the smallest normal float
_FLOAT_SLACK = 2.0 ** -20
_FLOAT_TINY = 2.0 ** -126
_FLOAT_MAX = 3.4e38


class Value:
    """A closed interval ``[lo, hi]`` of values of C type ``ctype``."""

    __slots__ = ('ctype', 'lo', 'hi', 'integral')

    def __init__(self, ctype, lo, hi, integral=False):
        self.ctype = ctype
        self.lo = lo
        self.hi = hi
        #: a float that holds an integer of [lo, hi] exactly
        self.integral = integral

    @classmethod
    def full(cls, ctype):
        """Every value of the type."""
        return cls(ctype, ctype.lo, ctype.hi)

    def __repr__(self):
        return "Value(%s, %r, %r)" % (self.ctype, self.lo, self.hi)


class _Pointer:
    """A pointer ``depth`` levels above values of type ``ctype``."""

    def __init__(self, ctype, depth):
        self.ctype = ctype
        self.depth = depth


def _name_type(words):
    words = tuple(w for w in words if w != "const")
    if words and words[-1] == "int" and len(words) > 1:
        words = words[:-1]
    return CTYPES.get(words)


def type_named(text):
    """The :class:`CType` of a C type name such as ``"unsigned char"``."""
    return _name_type(text.split())


def _fit(ctype, lo, hi):
    """``[lo, hi]`` as a value of ``ctype``: the whole type if outside."""
    if ctype.floating:
        if (math.isnan(lo) or math.isnan(hi) or lo < -_FLOAT_MAX
                or hi > _FLOAT_MAX):
            return Value.full(ctype)
        return Value(ctype, lo, hi)
    if lo < ctype.lo or hi > ctype.hi:
        return Value.full(ctype)
    return Value(ctype, lo, hi)


def _widen(ctype, lo, hi):
    """A float result rounded outward."""
    if math.isinf(lo) or math.isinf(hi) or math.isnan(lo) or math.isnan(hi):
        return Value.full(ctype)
    return _fit(ctype, lo - abs(lo) * _FLOAT_SLACK - _FLOAT_TINY,
                hi + abs(hi) * _FLOAT_SLACK + _FLOAT_TINY)


def convert(val, ctype):
    """C conversion of ``val`` to ``ctype`` (assignment, cast)."""
    src = val.ctype
    if ctype is _BOOL:
        if val.lo == val.hi == 0:
            return Value(_BOOL, 0, 0)
        if val.lo > 0 or val.hi < 0:
            return Value(_BOOL, 1, 1)
        return Value(_BOOL, 0, 1)
    if ctype.floating:
        if src.floating:
            if src.bits > ctype.bits:
                return _widen(ctype, val.lo, val.hi)
            return _fit(ctype, val.lo, val.hi)
        lo, hi = float(val.lo), float(val.hi)
        if max(abs(val.lo), abs(val.hi)) < (1 << 24):
            return Value(ctype, lo, hi, integral=True)
        return _widen(ctype, lo, hi)
    if src.floating:
        if math.isinf(val.lo) or math.isinf(val.hi):
            return Value.full(ctype)
        # truncation toward zero is monotone; out of range is undefined
        return _fit(ctype, int(val.lo), int(val.hi))
    return _fit(ctype, val.lo, val.hi)


def _promote(ctype):
    if ctype.floating or ctype.bits >= 32:
        return ctype
    return _INT


def _common(a, b):
    """The usual arithmetic conversions of types ``a`` and ``b``."""
    if a.floating or b.floating:
        return _DOUBLE if _DOUBLE in (a, b) else _FLOAT
    a, b = _promote(a), _promote(b)
    if a is b:
        return a
    if a.signed == b.signed:
        return a if a.bits >= b.bits else b
    uns, sig = (a, b) if not a.signed else (b, a)
    if uns.bits >= sig.bits:
        return uns
    if sig.bits > uns.bits:
        return sig
    return _ULONG if sig.bits == 64 else _UINT


def _corners(func, a, b):
    vals = [func(x, y) for x in (a.lo, a.hi) for y in (b.lo, b.hi)]
    return min(vals), max(vals)


def _tdiv(x, y):
    q = abs(x) // abs(y)
    return q if (x >= 0) == (y >= 0) else -q


def _bitlen_mask(n):
    return (1 << max(0, int(n)).bit_length()) - 1


def binary(op, a, b):
    """``a op b`` with C's conversions."""
    if op in ("<", "<=", ">", ">=", "==", "!=", "&&", "||"):
        return Value(_INT, 0, 1)
    if op in ("<<", ">>"):
        if a.ctype.floating or b.ctype.floating:
            raise Unknown("shift of a float")
        ctype = _promote(a.ctype)
        a = convert(a, ctype)
        if b.lo < 0 or b.hi >= ctype.bits:
            return Value.full(ctype)
        if op == ">>":
            return _fit(ctype, *_corners(lambda x, y: x >> y, a, b))
        if a.lo < 0:
            return Value.full(ctype)
        return _fit(ctype, a.lo << b.lo, a.hi << b.hi)
    ctype = _common(a.ctype, b.ctype)
    a, b = convert(a, ctype), convert(b, ctype)
    if ctype.floating:
        if op == "+":
            return _widen(ctype, a.lo + b.lo, a.hi + b.hi)
        if op == "-":
            return _widen(ctype, a.lo - b.hi, a.hi - b.lo)
        if op == "*":
            if any(math.isinf(v) for v in (a.lo, a.hi, b.lo, b.hi)):
                return Value.full(ctype)
            return _widen(ctype, *_corners(lambda x, y: x * y, a, b))
        if op == "/":
            if b.lo <= 0 <= b.hi or any(
                    math.isinf(v) for v in (a.lo, a.hi, b.lo, b.hi)):
                return Value.full(ctype)
            return _widen(ctype, *_corners(lambda x, y: x / y, a, b))
        raise Unknown("float %s" % op)
    if a.lo == a.hi and b.lo == b.hi and op in ("%", "&", "|", "^"):
        # two known values: the exact result
        x, y = a.lo, b.lo
        if op == "%":
            if y == 0:
                return Value.full(ctype)
            return _fit(ctype, x - _tdiv(x, y) * y, x - _tdiv(x, y) * y)
        exact = {"&": x & y, "|": x | y, "^": x ^ y}[op]
        return _fit(ctype, exact, exact)
    if op == "+":
        return _fit(ctype, a.lo + b.lo, a.hi + b.hi)
    if op == "-":
        return _fit(ctype, a.lo - b.hi, a.hi - b.lo)
    if op == "*":
        return _fit(ctype, *_corners(lambda x, y: x * y, a, b))
    if op == "/":
        if b.lo <= 0 <= b.hi:
            return Value.full(ctype)
        return _fit(ctype, *_corners(_tdiv, a, b))
    if op == "%":
        if b.lo <= 0 <= b.hi:
            return Value.full(ctype)
        m = max(abs(b.lo), abs(b.hi)) - 1
        if a.lo >= 0:
            if a.hi < min(abs(b.lo), abs(b.hi)):
                return a
            return _fit(ctype, 0, min(a.hi, m))
        if a.hi <= 0:
            return _fit(ctype, max(a.lo, -m), 0)
        return _fit(ctype, max(a.lo, -m), min(a.hi, m))
    if op == "&":
        if a.lo >= 0 and b.lo >= 0:
            return Value(ctype, 0, min(a.hi, b.hi))
        if a.lo >= 0:
            return Value(ctype, 0, a.hi)
        if b.lo >= 0:
            return Value(ctype, 0, b.hi)
        return Value.full(ctype)
    if op in ("|", "^"):
        if a.lo >= 0 and b.lo >= 0:
            lo = max(a.lo, b.lo) if op == "|" else 0
            return _fit(ctype, lo, _bitlen_mask(max(a.hi, b.hi)))
        return Value.full(ctype)
    raise Unknown(op)


def unary(op, a):
    """``op a`` with C's promotions."""
    if op == "!":
        return Value(_INT, 0, 1)
    if op == "+":
        return convert(a, _promote(a.ctype))
    ctype = _promote(a.ctype)
    a = convert(a, ctype)
    if op == "-":
        if ctype.floating:
            return Value(ctype, -a.hi, -a.lo)
        return _fit(ctype, -a.hi, -a.lo)
    if op == "~":
        if ctype.floating:
            raise Unknown("~ of a float")
        if ctype.signed:
            return Value(ctype, -a.hi - 1, -a.lo - 1)
        return Value(ctype, ctype.hi - a.hi, ctype.hi - a.lo)
    raise Unknown(op)


def _minmax(func, a, b):
    ctype = _common(a.ctype, b.ctype)
    a, b = convert(a, ctype), convert(b, ctype)
    return Value(ctype, func(a.lo, b.lo), func(a.hi, b.hi))


def _abs(a):
    ctype = _promote(a.ctype)
    a = convert(a, ctype)
    if a.lo >= 0:
        return a
    if a.hi <= 0:
        return _fit(ctype, -a.hi, -a.lo)
    return _fit(ctype, 0, max(-a.lo, a.hi))


_FUNCTIONS = {
    "min": (2, lambda a, b: _minmax(min, a, b)),
    "max": (2, lambda a, b: _minmax(max, a, b)),
    "abs": (1, _abs),
}

_TOKEN = re.compile(r"""\s*(?:
    (?P<float>(?:\d+\.\d*|\.\d+)(?:[eE][-+]?\d+)?[fF]?|\d+[eE][-+]?\d+[fF]?)
  | (?P<int>0[xX][0-9a-fA-F]+[uUlL]*|\d+[uUlL]*)
  | (?P<name>[A-Za-z_]\w*)
  | (?P<op><<=|>>=|<<|>>|<=|>=|==|!=|&&|\|\||[-+*/%&|^]=
         |[-+*/%&|^~!<>?:()\[\],.=;{}])
)""", re.VERBOSE)

_PRECEDENCE = {
    "||": 1, "&&": 2, "|": 3, "^": 4, "&": 5, "==": 6, "!=": 6,
    "<": 7, "<=": 7, ">": 7, ">=": 7, "<<": 8, ">>": 8,
    "+": 9, "-": 9, "*": 10, "/": 10, "%": 10,
}


def tokenize(text):
    """C tokens of ``text``; raises :class:`GiveUp` on anything else."""
    tokens = []
    pos = 0
    text = text.rstrip()
    while pos < len(text):
        match = _TOKEN.match(text, pos)
        if match is None or match.end() == pos:
            raise GiveUp("token at %r" % text[pos:pos + 20])
        kind = match.lastgroup
        tokens.append((kind, match.group(kind)))
        pos = match.end()
    return tokens


def _int_literal(text):
    """An integer literal with the type C gives it."""
    body = text.rstrip("uUlL")
    suffix = text[len(body):].lower()
    decimal = not body.startswith("0") or body == "0"
    if body[:2].lower() == "0x":
        value = int(body, 16)
    elif not decimal:
        value = int(body, 8)
    else:
        value = int(body)
    unsigned = "u" in suffix
    if unsigned:
        candidates = (_UINT, _ULONG) if "l" not in suffix else (_ULONG, )
    elif decimal:
        candidates = (_INT, _LONG) if "l" not in suffix else (_LONG, )
    else:
        candidates = ((_INT, _UINT, _LONG, _ULONG) if "l" not in suffix
                      else (_LONG, _ULONG))
    for ctype in candidates:
        if value <= ctype.hi:
            return Value(ctype, value, value)
    return Value(_ULONG, value, value)


class _Parser:
    """
    Evaluates one expression's tokens over an environment of ranges, and
    writes the expression out again: fully parenthesised, and with
    ``selects`` a product of integers one of which is 0 or 1 as a select
    (``a * b`` with ``a`` in [0, 1] is ``a ? (T) b : (T) 0``, ``T`` the
    product's type), which the GPU does in one instruction where a 32-bit
    multiply takes four.
    """

    def __init__(self, tokens, lookup, type_of, selects=False):
        self.tokens = tokens
        self.pos = 0
        self.lookup = lookup
        self.type_of = type_of
        self.selects = selects

    def peek(self, offset=0):
        idx = self.pos + offset
        return self.tokens[idx] if idx < len(self.tokens) else (None, None)

    def take(self, text=None):
        tok = self.peek()
        if text is not None and tok[1] != text:
            raise Unknown("expected %r, got %r" % (text, tok[1]))
        self.pos += 1
        return tok

    def expression(self):
        cond, ctext = self.binary(1)
        if self.peek()[1] != "?":
            return cond, ctext
        self.take("?")
        first, ftext = self.expression()
        self.take(":")
        second, stext = self.expression()
        if _is_pointer(cond) or _is_pointer(first) or _is_pointer(second):
            raise Unknown("pointer in ?:")
        ctype = _common(first.ctype, second.ctype)
        first, second = convert(first, ctype), convert(second, ctype)
        return (Value(ctype, min(first.lo, second.lo),
                      max(first.hi, second.hi)),
                "(%s ? %s : %s)" % (ctext, ftext, stext))

    def binary(self, level):
        left, ltext = self.unary()
        while True:
            op = self.peek()[1]
            prec = _PRECEDENCE.get(op) if self.peek()[0] == "op" else None
            if prec is None or prec < level:
                return left, ltext
            self.take()
            right, rtext = self.binary(prec + 1)
            text = "(%s %s %s)" % (ltext, op, rtext)
            if _is_pointer(left):
                if op in ("+", "-") and not _is_pointer(right):
                    ltext = text   # pointer arithmetic keeps the pointer
                    continue
                raise Unknown("pointer operand")
            if _is_pointer(right):
                raise Unknown("pointer operand")
            value = binary(op, left, right)
            if op == "*" and self.selects:
                text = _select_product(left, ltext, right, rtext,
                                       value.ctype) or text
            elif op == "/" and self.selects:
                text = _constant_division(left, ltext, right,
                                          value.ctype) or text
            left, ltext = value, text

    def cast_type(self):
        """The type (and its words) of a ``(type)`` cast at the cursor."""
        idx = 1
        words = []
        while True:
            kind, text = self.peek(idx)
            if kind == "name" and (text in _TYPE_WORDS or text == "const"
                                   or self.type_of(text) is not None
                                   and not words):
                words.append(text)
                idx += 1
                continue
            break
        if not words or self.peek(idx)[1] != ")":
            return None, None
        if len(words) == 1 and self.type_of(words[0]) is not None:
            ctype = self.type_of(words[0])
        else:
            ctype = _name_type(words)
        if ctype is None:
            return None, None
        self.pos += idx + 1
        return ctype, " ".join(words)

    def unary(self):
        kind, text = self.peek()
        if kind == "op" and text in ("-", "+", "!", "~"):
            self.take()
            operand, otext = self.unary()
            if _is_pointer(operand):
                raise Unknown("pointer operand")
            return unary(text, operand), "(%s%s)" % (text, otext)
        if kind == "op" and text in ("*", "&"):
            raise Unknown("dereference or address")
        if kind == "op" and text == "(":
            ctype, words = self.cast_type()
            if ctype is not None:
                operand, otext = self.unary()
                if _is_pointer(operand):
                    raise Unknown("pointer cast")
                return convert(operand, ctype), "((%s) %s)" % (words, otext)
        return self.postfix()

    def postfix(self):
        value, text = self.primary()
        while True:
            nxt = self.peek()[1]
            if nxt == "[":
                text += self.skip_brackets("[", "]")
                if not _is_pointer(value):
                    raise Unknown("subscript of a value")
                if value.depth == 1:
                    value = Value.full(value.ctype)
                else:
                    value = _Pointer(value.ctype, value.depth - 1)
            elif nxt in (".", "("):
                raise Unknown("member or call")
            else:
                return value, text

    def skip_brackets(self, opening, closing):
        """The tokens up to the matching ``closing``, as text."""
        depth = 0
        texts = []
        while True:
            kind, text = self.take()
            if kind is None:
                raise Unknown("unbalanced")
            texts.append(text)
            if text == opening:
                depth += 1
            elif text == closing:
                depth -= 1
                if depth == 0:
                    return " ".join(texts)

    def primary(self):
        kind, text = self.take()
        if kind == "int":
            return _int_literal(text), text
        if kind == "float":
            ctype = _FLOAT if text[-1] in "fF" else _DOUBLE
            value = float(text.rstrip("fF"))
            if ctype is _FLOAT and _is_float32(value):
                return Value(ctype, value, value), text
            return _widen(ctype, value, value), text
        if kind == "op" and text == "(":
            value, inner = self.expression()
            self.take(")")
            return value, "(%s)" % inner
        if kind == "name":
            if self.peek()[1] == "(" and text in _FUNCTIONS:
                count, func = _FUNCTIONS[text]
                self.take("(")
                args = [self.expression()]
                while self.peek()[1] == ",":
                    self.take(",")
                    args.append(self.expression())
                self.take(")")
                if len(args) != count or any(
                        _is_pointer(a) for a, _ in args):
                    raise Unknown(text)
                return (func(*(a for a, _ in args)),
                        "%s(%s)" % (text, ", ".join(t for _, t in args)))
            return self.lookup(text), text
        raise Unknown("token %r" % text)


def _select_product(left, ltext, right, rtext, ctype):
    """``left * right`` as a select when one integer factor is 0 or 1."""
    if ctype.floating or left.ctype.floating or right.ctype.floating:
        return None
    for cond, ctext, other in ((left, ltext, rtext), (right, rtext, ltext)):
        if cond.lo >= 0 and cond.hi <= 1:
            return "(%s ? (%s) %s : (%s) 0)" % (ctext, ctype.name, other,
                                               ctype.name)
    return None


#: a float divided by a constant: a product with the rounded reciprocal and
#: one fused correction, x * r + fma(x - x * r * c, r), for float kernels
FDIV_HELPER = (
    "__device__ __forceinline__ float _xen_fdiv(float x, float c, float r) {\n"
    "    const float q = x * r;\n"
    "    return __builtin_fmaf(__builtin_fmaf(-q, c, x), r, q);\n"
    "}\n")

_EXACT_DIVISIONS = {}


def division_by_constant_exact(c, lo, hi, recip=None):
    """
    Whether ``_xen_fdiv(x, c, RN(1/c))`` is the correctly rounded float
    quotient ``x / c`` for every integer ``x`` in ``[lo, hi]`` (float32,
    round to nearest even), checked on every one of them.

    The quotient is the hardware's IEEE division (NumPy).  The product
    ``q = RN(x * r)`` is NumPy's float32 product; the residual
    ``x - q * c`` is exact in float64 (``q * c`` has 48 significant bits
    and lies within a factor 2 of ``x``) and the first fused operation
    rounds it once; ``e * r`` is exact in float64 and the second fused
    operation's sum is carried as a float64 pair (TwoSum), which rounds
    to float32 as its leading part does unless that part is a float32
    midpoint, where the trailing part's sign decides.
    """
    key = (float(c), int(lo), int(hi))
    if recip is None and key in _EXACT_DIVISIONS:
        return _EXACT_DIVISIONS[key]
    import numpy as np
    c32 = np.float32(c)
    r32 = np.float32(np.float64(1.0) / np.float64(c32)) if recip is None \
        else np.float32(recip)
    ok = bool(np.isfinite(r32)) and float(c32) != 0.0
    if ok and hi >= lo:
        x = np.arange(lo, hi + 1, dtype=np.float64).astype(np.float32)
        want = x / c32
        q = x * r32
        e = (x.astype(np.float64)
             - q.astype(np.float64) * np.float64(c32)).astype(np.float32)
        a = q.astype(np.float64)
        b = e.astype(np.float64) * np.float64(r32)
        s = a + b
        bb = s - a
        err = (a - (s - bb)) + (b - bb)
        got = s.astype(np.float32)
        # s exactly halfway between two floats: the trailing part decides
        up = np.nextafter(got, np.float32(np.inf))
        down = np.nextafter(got, np.float32(-np.inf))
        half_up = (up.astype(np.float64) + got.astype(np.float64)) / 2
        half_down = (down.astype(np.float64) + got.astype(np.float64)) / 2
        got = np.where((s == half_up) & (err > 0), up, got)
        got = np.where((s == half_down) & (err < 0), down, got)
        ok = bool(np.array_equal(got, want))
    result = (ok, float(r32)) if ok else (False, None)
    if recip is None:
        _EXACT_DIVISIONS[key] = result
    return result


def _constant_division(left, ltext, right, ctype):
    """``left / right`` through ``_xen_fdiv`` where that is exact: a float
    holding an integer of a range of at most 2^20 values, divided by a
    float constant."""
    exact_int = (not left.ctype.floating
                 and max(abs(left.lo), abs(left.hi)) < (1 << 24))
    if (ctype.name != "float" or not (left.integral or exact_int)
            or right.lo != right.hi or left.hi - left.lo > (1 << 20)):
        return None
    ok, recip = division_by_constant_exact(right.lo, int(left.lo),
                                           int(left.hi))
    if not ok:
        return None
    # the divisor the check used: the constant rounded to float32, written
    # with enough digits to parse back to that same float
    import numpy as np
    divisor = float(np.float32(right.lo))
    return "_xen_fdiv(%s, %.9ef, %.9ef)" % (ltext, divisor, recip)


def _is_float32(value):
    import struct
    return struct.unpack("f", struct.pack("f", value))[0] == value


def _is_pointer(value):
    return isinstance(value, _Pointer)


class Analysis:
    """
    Ranges of the variables of straight-line generated statements.

    :param names: ``{identifier: Value or None}`` known before the code:
        parameters, constants, ``i``; None means "declared, value unknown"
    :param pointers: ``{identifier: (CType, depth)}`` of the arrays the
        code reads (their elements span their whole type)
    :param types: ``{type name: CType}`` of typedefs (``_word``)
    :param stores: names of further arrays the code stores into (``col``)
    """

    def __init__(self, names=None, pointers=None, types=None,
                 stores=("col", )):
        self.env = dict(names or {})
        self.pointers = dict(pointers or {})
        self.types = dict(types or {})
        self.arrays = set(self.pointers) | set(stores)
        self.declared = {}

    def type_of(self, name):
        """A typedef's type, or None."""
        return self.types.get(name)

    def lookup(self, name):
        """The range of identifier ``name``."""
        if name in self.pointers:
            ctype, depth = self.pointers[name]
            return _Pointer(ctype, depth)
        value = self.env.get(name)
        if value is None:
            raise Unknown(name)
        return value

    def evaluate(self, text, selects=False):
        """The range of a C expression (raises :class:`Unknown`); with
        ``selects``, ``(range, text written out again)``."""
        parser = _Parser(tokenize(text), self.lookup, self.type_of, selects)
        value, out = parser.expression()
        if parser.pos != len(parser.tokens) or _is_pointer(value):
            raise Unknown("trailing tokens")
        return (value, out) if selects else value

    def _eval_or_full(self, text, ctype):
        try:
            value, out = self.evaluate(text, selects=True)
            return convert(value, ctype), out
        except Unknown:
            return Value.full(ctype), text.strip()

    def run(self, code):
        """Follow every statement of ``code`` (raises on unknown shapes)."""
        for raw in code.split(";"):
            stmt = raw.strip()
            if stmt:
                self.statement(stmt)

    def rewrite(self, code):
        """
        ``(text, complete)``: ``code`` with every product of integers one
        of which the ranges show to be 0 or 1 written as a select, and
        whether every statement was followed (the statements from the
        first of unknown shape on are kept as they are).  Follows the code
        as :meth:`run` does.
        """
        out = []
        pieces = code.split(";")
        for idx, raw in enumerate(pieces):
            stmt = raw.strip()
            if not stmt:
                out.append(raw)
                continue
            try:
                out.append("\n" + self.statement(stmt))
            except GiveUp:
                out.append(";".join(pieces[idx:]))
                return ";".join(out), False
        return ";".join(out), True

    def statement(self, stmt):
        """Follow one statement (without its ``;``); returns it written
        out again with products of a 0/1 factor as selects."""
        tokens = tokenize(stmt)
        if any(text in ("{", "}") or text in ("if", "for", "while", "do",
                                                "return", "goto", "switch")
               for _, text in tokens):
            raise GiveUp("control flow: %r" % stmt)
        # declaration: type words, a name, optionally "= expr"
        idx = 0
        words = []
        while idx < len(tokens) and tokens[idx][0] == "name" and (
                tokens[idx][1] in _TYPE_WORDS or tokens[idx][1] == "const"
                or (not words and tokens[idx][1] in self.types)):
            words.append(tokens[idx][1])
            idx += 1
        if words and idx < len(tokens) and tokens[idx][0] == "name":
            name = tokens[idx][1]
            if len(words) == 1 and words[0] in self.types:
                ctype = self.types[words[0]]
            else:
                ctype = _name_type(words)
            if ctype is None:
                raise GiveUp("type %r" % (words, ))
            rest = tokens[idx + 1:]
            self.declared[name] = ctype
            if not rest:
                self.env[name] = Value.full(ctype)
                return stmt
            if rest[0][1] != "=":
                raise GiveUp("declaration %r" % stmt)
            cut = stmt.index("=")
            self.env[name], out = self._eval_or_full(stmt[cut + 1:], ctype)
            return "%s= %s" % (stmt[:cut], out)
        if tokens and tokens[0][0] == "name" and len(tokens) > 1 and (
                tokens[1][1] == "=" or tokens[1][1].endswith("=")
                and tokens[1][1] not in ("==", "!=", "<=", ">=")):
            name, op = tokens[0][1], tokens[1][1]
            if name not in self.declared:
                raise GiveUp("assignment to undeclared %r" % name)
            ctype = self.declared[name]
            cut = stmt.index(op)
            rhs = stmt[cut + len(op):]
            if op == "=":
                self.env[name], out = self._eval_or_full(rhs, ctype)
                return "%s= %s" % (stmt[:cut], out)
            try:
                value, out = self.evaluate(rhs, selects=True)
                self.env[name] = convert(
                    binary(op[:-1], self.env[name], value), ctype)
            except Unknown:
                self.env[name] = Value.full(ctype)
                out = rhs.strip()
            return "%s%s %s" % (stmt[:cut], op, out)
        # a store into one of the arrays (never a local: its target starts
        # with an array's name and takes no address) changes no local
        texts = [text for _, text in tokens]
        if "=" in texts:
            target = tokens[:texts.index("=")]
            names = [text for kind, text in target if kind == "name"]
            if (names and names[0] in self.arrays
                    and "&" not in texts[:len(target)]
                    and target[-1][1] == "]"):
                return stmt
        raise GiveUp("statement %r" % stmt)


def channel_bounds(code, names, pointers, types, channels):
    """
    Ranges of the C variables ``channels`` after the straight-line
    ``code``: a list of ``(lo, hi)`` integer pairs, or None when the code
    has a statement the analysis does not know or a channel is not a
    bounded integer.
    """
    analysis = Analysis(names, pointers, types)
    try:
        analysis.run(code)
    except GiveUp:
        return None
    out = []
    for name in channels:
        value = analysis.env.get(name)
        if value is None or value.ctype.floating:
            return None
        out.append((value.lo, value.hi))
    return out
