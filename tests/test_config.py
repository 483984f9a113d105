"""The configuration surface, as pinned by the reference's test/test_config.py."""
import numpy
import pytest

import scopyon_amd
from scopyon_amd.config import DefaultConfiguration, Configuration
from scopyon_amd.constants import Q_
from scopyon_amd import units


def test_default_loads():
    config = DefaultConfiguration()
    assert config.default.detector.exposure_time == pytest.approx(0.1)
    config.default.detector.exposure_time = 0.033
    assert config.default.detector.exposure_time == pytest.approx(0.033)


def test_update_with_units():
    config = DefaultConfiguration()
    config.update("""
    default:
        detector:
            exposure_time:
                value: 0.033
                units: s
    """)
    assert config.default.detector.exposure_time == pytest.approx(0.033)
    config.update("""
    default:
        detector:
            exposure_time:
                value: 100
                units: ms
    """)
    with pytest.warns(UserWarning):
        assert config.default.detector.exposure_time == pytest.approx(0.1)
    config.default.detector.exposure_time = 0.033
    assert config.default.detector.exposure_time == pytest.approx(0.033)


def test_update_plain_value():
    config = DefaultConfiguration()
    config.update("""
    default:
        detector:
            exposure_time: 0.033
    """)
    assert config.default.detector.exposure_time == pytest.approx(0.033)


def test_quantity_assignment():
    config = DefaultConfiguration()
    config.default.detector.exposure_time = Q_(33, 'ms')
    with pytest.warns(UserWarning):
        assert config.default.detector.exposure_time == pytest.approx(0.033)


def test_dimensionality_error():
    config = DefaultConfiguration()
    with pytest.raises(units.DimensionalityError):
        config.default.detector.exposure_time = Q_(33, 'm')


def test_missing_key_and_subtree():
    config = DefaultConfiguration()
    with pytest.raises(KeyError):
        config.default.no_such_key
    assert isinstance(config.default, Configuration)
    with pytest.raises(ValueError):
        config.default.detector = 3
    with pytest.raises(TypeError):
        config.default.detector.QE = {'a': 1}


def test_si_values_and_lists():
    config = DefaultConfiguration()
    assert config.default.light_source.flux_density == 400000.0
    assert config.default.light_source.angle == 1.2566370614359172
    assert numpy.array_equal(config.default.detector.focal_point, [0.0, 0.0, 0.0])
    assert config.default.detector.image_size == [512, 512]
    assert config.default.fluorophore.radial_width is None


def test_units_parser():
    f, d = units.parse_units('kg / s ** 3')
    assert f == 1.0 and d == {'kg': 1, 's': -3}
    assert units.parse_units('um**2/s')[0] == pytest.approx(1e-12)
    assert units.parse_units('degree')[0] == pytest.approx(numpy.pi / 180)
    with pytest.raises(ValueError):
        units.parse_units('furlong')


def test_save_roundtrip(tmp_path):
    config = DefaultConfiguration()
    config.default.detector.exposure_time = 0.05
    path = tmp_path / 'c.yaml'
    config.save(path)
    again = Configuration(str(path))
    assert again.default.detector.exposure_time == 0.05


def test_public_names():
    for name in ['EnvironSettings', 'EPIFMSimulator', 'form_image', 'generate_images',
                 'create_simulator', 'Configuration', 'DefaultConfiguration', 'Image', 'Video',
                 'sample_inputs', 'sample', 'constants']:
        assert hasattr(scopyon_amd, name)


def test_create_simulator_errors():
    config = DefaultConfiguration()
    config.update({'default': {'type': 'confocal'}})
    with pytest.raises(ValueError):
        scopyon_amd.create_simulator(config)
    with pytest.raises(TypeError):
        scopyon_amd.EPIFMSimulator(config=3.0, rng=numpy.random.RandomState(0))
