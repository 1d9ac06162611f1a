"""
Expression objects of the DSL.

:class:`DeferredExpression` turns Python operators into parenthesised C
text, with the operator set and formatting of
``xentica/core/expressions.py:9-120``:

* binary ``a op b`` gives ``"(a op b)"`` and the reflected form swaps the
  operands (``expressions.py:66-81``);
* unary ``op a`` gives ``"(op(a))"`` (``expressions.py:106-116``);
* ``a op= b`` on a variable or property emits the statement
  ``"name op= b;"`` and returns the variable; on a bare expression it is an
  error (``expressions.py:83-97``).

:class:`PatternExpression` is the lazy arithmetic used by seeds
(``expressions.py:123-204``): operators record a tree that is evaluated
with NumPy when a pattern asks for its value through ``__get__``.
"""
import operator

from xentica_amd.core.exceptions import XenticaException

__all__ = ['DeferredExpression', 'PatternExpression', ]

#: (dunder stem, C operator) pairs, in the reference's order.
_BINARY_OPS = (
    ('add', '+'), ('sub', '-'), ('mul', '*'), ('truediv', '/'),
    ('mod', '%'), ('rshift', '>>'), ('lshift', '<<'), ('and', '&'),
    ('xor', '^'), ('or', '|'), ('lt', '<'), ('le', '<='), ('eq', '=='),
    ('ne', '!='), ('gt', '>'), ('ge', '>='),
)
_UNARY_OPS = (
    ('neg', '-'), ('pos', '+'), ('abs', 'abs'), ('invert', '~'),
    ('int', '(int)'), ('float', '(float)'), ('round', 'round'),
)


def _binary(c_op):
    def method(self, other):
        return DeferredExpression("(%s %s %s)" % (self, c_op, other))
    return method


def _reflected(c_op):
    def method(self, other):
        return DeferredExpression("(%s %s %s)" % (other, c_op, self))
    return method


def _augmented(c_op):
    def method(self, other):
        if type(self) is DeferredExpression:
            raise XenticaException("Can't assign to DeferredExpression")
        self.declare_once()
        self.bsca.append_code("%s %s= %s;\n" % (self.var_name, c_op, other))
        return self
    return method


def _unary(c_op):
    def method(self):
        return DeferredExpression("(%s(%s))" % (c_op, self))
    return method


class DeferredExpression:
    """
    A piece of C code usable in mixed Python expressions.

    Allowed binary ops
        ``+``, ``-``, ``*``, ``/``, ``%``, ``>>``, ``<<``, ``&``,
        ``^``, ``|``, ``<``, ``<=``, ``>``, ``>=``, ``==``, ``!=``

    Allowed unary ops
        ``+``, ``-``, ``~``, ``abs``, ``int``, ``float``, ``round``

    Allowed augmented assigns
        ``+=``, ``-=``, ``*=``, ``/=``, ``%=``, ``<<=``, ``>>=``,
        ``&=``, ``^=``, ``|=``
    """

    def __init__(self, code=''):
        """Store the C code of the expression."""
        self.code = code

    def __str__(self):
        """Return the accumulated C code."""
        return self.code

    # ``__eq__`` is overloaded below, which would otherwise drop hashing.
    __hash__ = object.__hash__


for _name, _c_op in _BINARY_OPS:
    setattr(DeferredExpression, "__%s__" % _name, _binary(_c_op))
    setattr(DeferredExpression, "__r%s__" % _name, _reflected(_c_op))
    setattr(DeferredExpression, "__i%s__" % _name, _augmented(_c_op))
for _name, _c_op in _UNARY_OPS:
    setattr(DeferredExpression, "__%s__" % _name, _unary(_c_op))
del _name, _c_op


_PATTERN_BINARY = (
    'add', 'sub', 'mul', 'truediv', 'mod',
    'rshift', 'lshift', 'and', 'xor', 'or',
)
_PATTERN_UNARY = ('neg', 'pos', 'invert')


class PatternExpression:
    """
    Lazy arithmetic over seed values.

    Allowed binary ops
        ``+``, ``-``, ``*``, ``/``, ``%``, ``>>``, ``<<``, ``&``,
        ``^``, ``|``

    Allowed unary ops
        ``-``, ``~``
    """

    def __init__(self, arg1=None, arg2=None, cur_oper=None):
        """Record one node of the operator tree."""
        self.arg1 = arg1
        self.arg2 = arg2
        self.cur_oper = cur_oper

    def __get__(self, instance, owner):
        """Evaluate the tree for ``instance`` (a pattern holding a stream)."""
        def value(arg):
            if isinstance(arg, PatternExpression):
                return arg.__get__(instance, owner)
            return arg
        oper = getattr(operator, "__%s__" % self.cur_oper)
        if self.arg2 is None:
            return oper(value(self.arg1))
        return oper(value(self.arg1), value(self.arg2))


def _pattern_binary(name):
    def method(self, other):
        return PatternExpression(self, other, name)
    return method


def _pattern_reflected(name):
    def method(self, other):
        return PatternExpression(other, self, name)
    return method


def _pattern_unary(name):
    def method(self):
        return PatternExpression(self, None, name)
    return method


for _name in _PATTERN_BINARY:
    setattr(PatternExpression, "__%s__" % _name, _pattern_binary(_name))
    setattr(PatternExpression, "__r%s__" % _name, _pattern_reflected(_name))
for _name in _PATTERN_UNARY:
    setattr(PatternExpression, "__%s__" % _name, _pattern_unary(_name))
del _name
