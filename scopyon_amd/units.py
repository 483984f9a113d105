"""Physical units for the configuration surface.

The reference converts ``{value, units}`` config leaves with pint
(src/scopyon/config.py:101-147, src/scopyon/constants.py:1-8).  pint is used
when it is importable; otherwise this module supplies a small registry with the
same behaviour for the units scopyon configurations use: ``to_base_units()``
to SI magnitudes, dimensionality checks raising ``DimensionalityError`` and the
``'{:~}'`` short format.
"""
import math
import numbers
import re

import numpy

try:  # pragma: no cover - exercised only where pint is installed
    import pint as _pint
except ImportError:  # pint is optional
    _pint = None

# base dimensions: length, mass, time, current, temperature, substance
_DIMS = ('m', 'kg', 's', 'A', 'K', 'mol')

# unit -> (factor to SI, {dim: power})
_UNITS = {}


def _def(names, factor, **dims):
    for n in names:
        _UNITS[n] = (factor, dims)


_def(('m', 'meter', 'metre', 'meters'), 1.0, m=1)
_def(('g', 'gram', 'grams'), 1e-3, kg=1)
_def(('s', 'sec', 'second', 'seconds'), 1.0, s=1)
_def(('min', 'minute', 'minutes'), 60.0, s=1)
_def(('h', 'hour', 'hours'), 3600.0, s=1)
_def(('A', 'ampere'), 1.0, A=1)
_def(('K', 'kelvin'), 1.0, K=1)
_def(('mol', 'mole'), 1.0, mol=1)
_def(('rad', 'radian', 'radians'), 1.0)
_def(('deg', 'degree', 'degrees'), math.pi / 180.0)
_def(('dimensionless', ''), 1.0)
_def(('percent', '%'), 0.01)
_def(('Hz', 'hertz'), 1.0, s=-1)
_def(('N', 'newton'), 1.0, kg=1, m=1, s=-2)
_def(('J', 'joule'), 1.0, kg=1, m=2, s=-2)
_def(('W', 'watt'), 1.0, kg=1, m=2, s=-3)
_def(('M', 'molar'), 1e3, mol=1, m=-3)
_def(('L', 'l', 'liter', 'litre'), 1e-3, m=3)

_PREFIXES = {'Y': 1e24, 'Z': 1e21, 'E': 1e18, 'P': 1e15, 'T': 1e12, 'G': 1e9, 'M': 1e6,
             'k': 1e3, 'h': 1e2, 'da': 1e1, 'd': 1e-1, 'c': 1e-2, 'm': 1e-3, 'u': 1e-6,
             'µ': 1e-6, 'μ': 1e-6, 'n': 1e-9, 'p': 1e-12, 'f': 1e-15, 'a': 1e-18}
_LONG_PREFIXES = {'yotta': 1e24, 'zetta': 1e21, 'exa': 1e18, 'peta': 1e15, 'tera': 1e12,
                  'giga': 1e9, 'mega': 1e6, 'kilo': 1e3, 'hecto': 1e2, 'deca': 1e1,
                  'deci': 1e-1, 'centi': 1e-2, 'milli': 1e-3, 'micro': 1e-6, 'nano': 1e-9,
                  'pico': 1e-12, 'femto': 1e-15, 'atto': 1e-18}
_BASE_NAMES = {'m': 'meter', 'kg': 'kilogram', 's': 'second', 'A': 'ampere', 'K': 'kelvin',
               'mol': 'mole'}


class DimensionalityError(ValueError):
    """Raised when a quantity is assigned to a leaf of another dimension."""

    def __init__(self, units1, units2, dim1='', dim2=''):
        super().__init__("Cannot convert from '{}' ({}) to '{}' ({})".format(
            units1, dim1, units2, dim2))
        self.units1, self.units2, self.dim1, self.dim2 = units1, units2, dim1, dim2


def _lookup(name):
    if name in _UNITS:
        return _UNITS[name]
    for pre, f in _LONG_PREFIXES.items():
        if name.startswith(pre) and name[len(pre):] in _UNITS:
            base, dims = _UNITS[name[len(pre):]]
            return f * base, dims
    for pre in sorted(_PREFIXES, key=len, reverse=True):
        if name.startswith(pre) and name[len(pre):] in _UNITS and name[len(pre):]:
            base, dims = _UNITS[name[len(pre):]]
            return _PREFIXES[pre] * base, dims
    raise ValueError("unknown unit '{}'".format(name))


_TOKEN = re.compile(r'\s*(\*\*|\^|\*|/|\(|\)|[-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?|[A-Za-zµμ%_]+)')


def _tokens(text):
    pos, out = 0, []
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ValueError("cannot parse units '{}'".format(text))
        out.append(m.group(1))
        pos = m.end()
    return out


def parse_units(text):
    """'kg / s ** 3' -> (factor to SI, {dim: power})."""
    if text is None:
        return 1.0, {}
    toks = _tokens(str(text))
    pos = [0]

    def peek():
        return toks[pos[0]] if pos[0] < len(toks) else None

    def take():
        t = peek()
        pos[0] += 1
        return t

    def atom():
        t = take()
        if t == '(':
            f, d = expr()
            if take() != ')':
                raise ValueError("unbalanced parentheses in '{}'".format(text))
        elif t is not None and re.match(r'^[-+]?\d', t):
            f, d = float(t), {}
        elif t is None:
            raise ValueError("unexpected end of units '{}'".format(text))
        else:
            f, d = _lookup(t)
            d = dict(d)
        if peek() in ('**', '^'):
            take()
            e = float(take())
            f = f ** e
            d = {k: v * e for k, v in d.items()}
        return f, d

    def expr():
        f, d = atom()
        while peek() in ('*', '/'):
            op = take()
            f2, d2 = atom()
            sign = 1 if op == '*' else -1
            f = f * f2 if sign > 0 else f / f2
            for k, v in d2.items():
                d[k] = d.get(k, 0) + sign * v
        return f, {k: v for k, v in d.items() if v != 0}

    result = expr()
    if pos[0] != len(toks):
        raise ValueError("cannot parse units '{}'".format(text))
    return result


def _base_string(dims):
    if not dims:
        return 'dimensionless'
    num = [(_BASE_NAMES[k], v) for k, v in dims.items() if v > 0]
    den = [(_BASE_NAMES[k], -v) for k, v in dims.items() if v < 0]

    def fmt(name, p):
        p = int(p) if float(p).is_integer() else p
        return name if p == 1 else '{} ** {}'.format(name, p)

    s = ' * '.join(fmt(n, p) for n, p in num) or '1'
    if den:
        s += ' / ' + ' / '.join(fmt(n, p) for n, p in den)
    return s


class Quantity:
    """A magnitude with units (the subset of pint.Quantity the config uses)."""

    def __init__(self, value, units='dimensionless'):
        if isinstance(value, (list, tuple)):
            value = numpy.asarray(value, dtype=float)
        self.magnitude = value
        self.units = units if units is not None else 'dimensionless'
        self._factor, self._dims = parse_units(self.units)

    @property
    def dimensionality(self):
        return dict(self._dims)

    def check(self, other):
        other_dims = other._dims if isinstance(other, Quantity) else parse_units(other)[1]
        return self._dims == other_dims

    def to_base_units(self):
        mag = self.magnitude
        if self._factor != 1.0:
            mag = numpy.asarray(mag) * self._factor if not isinstance(mag, numbers.Number) \
                else mag * self._factor
        return Quantity(mag, _base_string(self._dims))

    def to(self, units):
        f, d = parse_units(units)
        if d != self._dims:
            raise DimensionalityError(self.units, units, self._dims, d)
        mag = self.magnitude
        scale = self._factor / f
        return Quantity(mag * scale if isinstance(mag, numbers.Number) else numpy.asarray(mag) * scale,
                        units)

    def __format__(self, spec):
        return '{} {}'.format(self.magnitude, self.units)

    def __repr__(self):
        return '<Quantity({}, {!r})>'.format(self.magnitude, self.units)


if _pint is not None:  # pragma: no cover
    ureg = _pint.UnitRegistry()
    Quantity = ureg.Quantity  # noqa: F811
    from pint.errors import DimensionalityError  # noqa: F811,E402

    def is_quantity(value):
        return isinstance(value, _pint.Quantity)

    def short_units(q):
        return '{:~}'.format(q.units)

    def base_magnitude(value, units):
        given = Quantity(value, units)
        v = given.to_base_units()
        return v.magnitude, given.units != v.units, '{:~}'.format(given.units), '{:~}'.format(v.units)
else:
    ureg = None

    def is_quantity(value):
        return isinstance(value, Quantity)

    def short_units(q):
        return q.units

    def base_magnitude(value, units):
        given = Quantity(value, units)
        v = given.to_base_units()
        changed = given._factor != 1.0
        mag = v.magnitude
        if isinstance(mag, numpy.ndarray) and mag.ndim == 0:
            mag = float(mag)
        return mag, changed, given.units, v.units

Q_ = Quantity
