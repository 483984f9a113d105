"""Spectral catalog reader (src/scopyon/io.py:29-66).

Catalogs are looked up first in ``$SCOPYON_CATALOG_DIR`` (a scopyon
``catalog/`` tree) and then in this package's ``catalog/`` folder, which ships
the TRITC fluorophore spectrum that the default configuration needs.
"""
import csv
import io as _io
import os

_HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'catalog')


def _isfloat(val):
    try:
        float(val)
        return True
    except ValueError:
        return False


def _find(relpath):
    for root in (os.environ.get('SCOPYON_CATALOG_DIR'), _HERE):
        if root and os.path.exists(os.path.join(root, relpath)):
            return os.path.join(root, relpath)
    raise FileNotFoundError(
        "catalog file '{}' not found (set SCOPYON_CATALOG_DIR to a scopyon catalog)".format(relpath))


def read_catalog(data_path):
    """5 header lines, then key rows followed by data rows (io.py:29-52)."""
    with open(_find(data_path)) as f:
        stream = _io.StringIO(f.read())
    for _ in range(5):
        stream.readline()
    reader = csv.reader(stream)
    row = next(reader)
    if len(row) != 1 or _isfloat(row[0]):
        raise RuntimeError('A catalog in invalid format was given [{}]'.format(data_path))
    key, table = row[0], {}
    for row in reader:
        if len(row) == 1 and not _isfloat(row[0]):
            key = row[0]
        elif len(row) != 0:
            table.setdefault(key, []).append(row)
    return table


def read_fluorophore_catalog(fluorophore_type):
    data = read_catalog('fluorophore/{}.csv'.format(fluorophore_type))
    return data.get('Excitation', []), data.get('Emission', [])


def read_dichroic_catalog(dm):
    return read_catalog('dichroic/{}.csv'.format(dm))['wavedataset']


def read_excitation_catalog(excitation):
    return read_catalog('excitation/{}.csv'.format(excitation))['wavedataset']


def read_emission_catalog(emission):
    return read_catalog('emission/{}.csv'.format(emission))['wavedataset']
