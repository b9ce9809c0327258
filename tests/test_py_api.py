"""The Python front against the reference module's own tests
(crates/py-propag/tests/test_basic.py): the same constructors, names
(`PyGeoReference`, `PySettings`, `PyTimeFeature` and the short names),
checks and error messages, so code written against the reference's
`py_propag` runs unchanged.  CPU only."""

from unittest.mock import patch

import numpy as np
import pytest

import py_propag

POINT_WKB = bytes([1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 240, 63, 0, 0, 0, 0, 0, 0, 0, 64])  # POINT(1 2)
TRANSFORM = np.array([0.0, 1.0, 0.0, 0.0, 0.0, -1.0], dtype=np.float64)


def test_reference_names_are_exported():
    for name in ("Settings", "PySettings", "GeoReference", "PyGeoReference", "TimeFeature", "PyTimeFeature",
                 "PropagResults", "PyPropagResults", "propagate", "load_terrain_data", "check_cuda_available",
                 "PropagCudaError", "PropagGdalError", "PropagValidationError", "PropagMPIError"):
        assert name in py_propag.__all__ and hasattr(py_propag, name), name
    assert py_propag.PySettings is py_propag.Settings
    assert py_propag.PyGeoReference is py_propag.GeoReference
    assert py_propag.__version__ == "0.1.0"
    # exception hierarchy as the reference declares it (lib.rs create_exception!)
    assert issubclass(py_propag.PropagValidationError, ValueError)
    for e in (py_propag.PropagCudaError, py_propag.PropagGdalError, py_propag.PropagMPIError):
        assert issubclass(e, RuntimeError)


class TestGpuAvailability:
    def test_reports_a_bool(self):
        assert isinstance(py_propag.check_cuda_available(), bool)

    def test_no_device_node_means_no_gpu(self):
        with patch("subprocess.run", side_effect=FileNotFoundError), patch("os.path.exists", return_value=False):
            assert py_propag.check_cuda_available() is False


class TestGeoReference:
    def test_creation(self):
        g = py_propag.PyGeoReference(100, 200, "EPSG:4326", TRANSFORM)
        assert (g.width, g.height) == (100, 200)
        assert g.is_north_up

    def test_invalid_dimensions(self):
        with pytest.raises(py_propag.PropagValidationError, match="Grid dimensions must be positive"):
            py_propag.PyGeoReference(0, 100, "EPSG:4326", TRANSFORM)
        with pytest.raises(py_propag.PropagValidationError, match="too large"):
            py_propag.PyGeoReference(32769, 100, "EPSG:4326", TRANSFORM)

    def test_invalid_transform(self):
        with pytest.raises(py_propag.PropagValidationError, match="exactly 6 elements"):
            py_propag.PyGeoReference(100, 200, "EPSG:4326", np.array([0.0, 1.0, 0.0]))
        with pytest.raises(py_propag.PropagValidationError, match="pixel size cannot be zero"):
            py_propag.PyGeoReference(100, 200, "EPSG:4326", [0.0, 0.0, 0.0, 0.0, 0.0, -1.0])
        with pytest.raises(py_propag.PropagValidationError, match="pixel size cannot be zero"):
            py_propag.PyGeoReference(100, 200, "EPSG:4326", [0.0, 1.0, 0.0, 0.0, 0.0, 0.0])
        with pytest.raises(py_propag.PropagValidationError, match="non-zero and finite"):
            py_propag.PyGeoReference(100, 200, "EPSG:4326", [0.0, float("inf"), 0.0, 0.0, 0.0, -1.0])

    def test_projection_edge_cases(self):
        with pytest.raises(py_propag.PropagValidationError, match="Projection string cannot be empty"):
            py_propag.PyGeoReference(100, 200, "", TRANSFORM)
        at_limit = "EPSG:4326+" + "a" * (1023 - 10)
        assert py_propag.PyGeoReference(100, 200, at_limit, TRANSFORM).width == 100
        with pytest.raises(py_propag.PropagValidationError, match=r"Projection string too large \(max 1023 characters\)"):
            py_propag.PyGeoReference(100, 200, at_limit + "a", TRANSFORM)
        # the limit is on UTF-8 bytes, not characters
        assert py_propag.PyGeoReference(100, 200, "EPSG:4326 with émojis 🌍 and spëcial chars: αβγ", TRANSFORM).height == 200
        two_byte = "EPSG:4326+" + "é" * ((1023 - 10) // 2)
        two_byte += "a" * (1023 - len(two_byte.encode()))
        assert len(two_byte.encode()) == 1023
        py_propag.PyGeoReference(100, 200, two_byte, TRANSFORM)
        three_byte = "EPSG:4326+" + "€" * 338
        assert len(three_byte.encode()) >= 1024 and len(three_byte) < 1024
        with pytest.raises(py_propag.PropagValidationError, match="Projection string too large"):
            py_propag.PyGeoReference(100, 200, three_byte, TRANSFORM)


class TestSettings:
    def test_creation(self):
        g = py_propag.PyGeoReference(100, 200, "EPSG:4326", TRANSFORM)
        s = py_propag.PySettings(g, 3600.0)
        assert s.max_time == 3600.0
        s.max_time = 7200.0
        assert s.max_time == 7200.0

    @pytest.mark.parametrize("bad, msg", [(-1.0, "must be positive"), (0.0, "must be positive"),
                                          (float("inf"), "finite number"), (float("nan"), "finite number"),
                                          (10.1 * 365.0 * 24.0 * 3600.0, "too large")])
    def test_invalid_max_time(self, bad, msg):
        g = py_propag.PyGeoReference(100, 200, "EPSG:4326", TRANSFORM)
        with pytest.raises(py_propag.PropagValidationError, match=msg):
            py_propag.PySettings(g, bad)
        s = py_propag.PySettings(g, 60.0)
        with pytest.raises(py_propag.PropagValidationError, match=msg):
            s.max_time = bad
        assert s.max_time == 60.0


class TestTimeFeature:
    def test_creation(self):
        f = py_propag.PyTimeFeature(300.0, POINT_WKB)
        assert f.time == 300.0 and f.geom_wkb == POINT_WKB
        assert py_propag.TimeFeature(0.0, POINT_WKB).time == 0.0

    def test_invalid_time(self):
        with pytest.raises(py_propag.PropagValidationError, match="non-negative"):
            py_propag.PyTimeFeature(-1.0, POINT_WKB)
        with pytest.raises(py_propag.PropagValidationError, match="finite number"):
            py_propag.PyTimeFeature(float("inf"), POINT_WKB)

    def test_invalid_geometry(self):
        with pytest.raises(py_propag.PropagValidationError, match="cannot be empty"):
            py_propag.PyTimeFeature(300.0, b"")
        with pytest.raises(py_propag.PropagValidationError, match="too short"):
            py_propag.PyTimeFeature(300.0, b"short")
        with pytest.raises(py_propag.PropagValidationError, match="too short"):
            py_propag.PyTimeFeature(300.0, b"x" * 8)

    def test_cells_instead_of_wkb(self):
        f = py_propag.TimeFeature(5.0, cells=[(1, 2)])
        assert f.geom_wkb == b"" and f.cells == [(1, 2)]


class TestLoadTerrainData:
    def test_basic(self):
        r = np.random.default_rng(0)
        w, h = 10, 10
        g = py_propag.load_terrain_data(w, h, "EPSG:4326", TRANSFORM,
                                        r.uniform(0, 1000, (h, w)).astype(np.float32),
                                        r.uniform(0, 45, (h, w)).astype(np.float32),
                                        r.uniform(0, 360, (h, w)).astype(np.float32),
                                        r.integers(1, 14, (h, w), dtype=np.uint8))
        assert (g.width, g.height) == (w, h)

    def test_invalid(self):
        with pytest.raises(py_propag.PropagValidationError, match="must be positive"):
            py_propag.load_terrain_data(0, 10, "EPSG:4326", TRANSFORM, None, None, None, None)
        with pytest.raises(py_propag.PropagValidationError, match="exceeds maximum"):
            py_propag.load_terrain_data(10, 10, "EPSG:4326", TRANSFORM, np.full((10, 10), 15000.0, np.float32))
        with pytest.raises(py_propag.PropagValidationError, match="below minimum"):
            py_propag.load_terrain_data(10, 10, "EPSG:4326", TRANSFORM, None, np.full((10, 10), -10.0, np.float32))
        with pytest.raises(py_propag.PropagValidationError, match="shape mismatch"):
            py_propag.load_terrain_data(100, 200, "EPSG:4326", TRANSFORM, np.zeros((210, 100), np.float32))
        with pytest.raises(py_propag.PropagValidationError, match="Invalid fuel model ID"):
            py_propag.load_terrain_data(10, 10, "EPSG:4326", TRANSFORM, None, None, None,
                                        np.full((10, 10), 20, np.uint8))
        with pytest.raises(py_propag.PropagValidationError, match="too large"):
            py_propag.load_terrain_data(50000, 50000, "EPSG:4326", TRANSFORM)
