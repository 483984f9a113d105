"""YAML configuration with unit-aware leaves.

Same surface as the reference's ``Configuration``/``DefaultConfiguration``
(src/scopyon/config.py:19-165): attribute access descends into sub-trees,
``{value, units}`` leaves read back as SI magnitudes (with a warning when a
conversion happened), assignment checks the dimensionality of quantities, and
``update`` deep-merges a dict, another Configuration or a YAML string.
"""
import collections.abc
import os
import pathlib
import warnings

import yaml

from . import units as _units

__all__ = ["Configuration", "DefaultConfiguration"]

_DEFAULT_YAML = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'scopyon.yaml')


def _loader():
    try:
        return yaml.CSafeLoader
    except AttributeError:  # pragma: no cover
        return yaml.SafeLoader


def _load_yaml(text):
    return yaml.load(text, Loader=_loader())


def dict_merge(dct, merge_dct):
    """Recursive in-place merge (config.py:19-24)."""
    for k, v in merge_dct.items():
        if isinstance(dct.get(k), dict) and isinstance(v, collections.abc.Mapping):
            dict_merge(dct[k], v)
        else:
            dct[k] = v


class Configuration(collections.abc.Mapping):
    """A view of one YAML (sub-)tree."""

    def __init__(self, filename=None, yaml=None):
        if filename is not None:
            assert yaml is None
            with open(filename) as f:
                self.__yaml = _load_yaml(f.read())
        elif yaml is not None:
            self.__yaml = yaml
        else:
            self.__yaml = None

    def __repr__(self):
        return yaml.safe_dump(self.__yaml, default_flow_style=False)

    def save(self, file):
        """Write the tree as YAML (config.py:56-65)."""
        assert isinstance(file, (str, pathlib.PurePath))
        with open(str(file), 'w') as f:
            f.write(repr(self))

    def update(self, conf):
        """Deep-merge a Configuration, a dict or a YAML string (config.py:67-78)."""
        if isinstance(conf, Configuration):
            dict_merge(self.__yaml, conf.yaml)
        elif isinstance(conf, dict):
            dict_merge(self.__yaml, conf)
        else:
            dict_merge(self.__yaml, _load_yaml(conf))

    @property
    def yaml(self):
        return self.__yaml

    def get(self, key, defaultobj=None):
        return self.__yaml.get(key, defaultobj)

    def __getitem__(self, key):
        return getattr(self, key)

    def __len__(self):
        return len(self.__yaml)

    def __iter__(self):
        return (key for key, value in self.__yaml.items()
                if not isinstance(value, dict) or 'value' in value)

    def __getattr__(self, key):
        if key.startswith('_Configuration__'):
            raise AttributeError(key)
        if key not in self.__yaml:
            raise KeyError("'{}'".format(key))
        value = self.__yaml[key]
        if isinstance(value, dict):
            if 'value' not in value:
                return Configuration(yaml=value)
            if value['value'] is not None and 'units' in value:
                mag, changed, src, dst = _units.base_magnitude(value['value'], value['units'])
                if changed:
                    warnings.warn("Unit conversion in '{}' from [{}] to [{}]".format(key, src, dst))
                return mag
            return value['value']
        return value

    def __setattr__(self, key, value):
        if key.startswith('_'):
            object.__setattr__(self, key, value)
            return
        if key not in self.__yaml:
            raise KeyError("'{}'".format(key))
        if isinstance(value, dict):
            raise TypeError("The given value for '{}' has wrong type: {}".format(key, value))
        original = self.__yaml[key]
        if isinstance(original, dict) and 'value' not in original:
            raise ValueError("Cannot update '{}'.".format(key))
        if _units.is_quantity(value):
            if isinstance(original, dict) and 'units' in original:
                other = _units.Q_(original['value'] if original['value'] is not None else 0.0,
                                  original['units'])
                if not value.check(other):
                    raise _units.DimensionalityError(
                        value.units, other.units, value.dimensionality, other.dimensionality)
            mag = value.magnitude
            if hasattr(mag, 'tolist'):
                mag = mag.tolist()
            self.__yaml[key] = dict(value=mag, units=_units.short_units(value))
        else:
            self.__yaml[key] = value


class DefaultConfiguration(Configuration):
    """The package's default configuration (scopyon.yaml)."""

    def __init__(self):
        with open(_DEFAULT_YAML) as f:
            data = _load_yaml(f.read())
        Configuration.__init__(self, yaml=data)


def to_si_dict(conf):
    """Resolve every leaf of a Configuration (sub-)tree to plain SI values."""
    out = {}
    for key, value in conf.yaml.items():
        if isinstance(value, dict) and 'value' not in value:
            out[key] = to_si_dict(getattr(conf, key))
        else:
            with warnings.catch_warnings():
                warnings.simplefilter('ignore')
                v = getattr(conf, key)
            if hasattr(v, 'tolist') and not isinstance(v, (int, float)):
                v = v.tolist()
            out[key] = v
    return out
