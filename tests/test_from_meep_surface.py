"""The user entry without MEEP: lifting an mp.Simulation refuses by naming MEEP, the result
object carries the reference's read-out names, and a stop condition is exactly one of two."""

import importlib.util

import pytest

from meep_gpu import from_meep


@pytest.mark.skipif(importlib.util.find_spec("meep") is not None, reason="MEEP is installed")
def test_lifting_without_meep_names_meep():
    with pytest.raises(ImportError, match="MEEP"):
        from_meep.lift_simulation(object())
    with pytest.raises(ImportError, match="MEEP"):
        from_meep.gpu_compatibility(object())
    with pytest.raises(ImportError, match="MEEP"):
        from_meep.run_on_gpu(object(), until=1.0)


def test_result_surface_names():
    for name in ("get_flux_spectrum", "get_flux_data", "get_array", "get_field", "close",
                 "monitors"):
        assert hasattr(from_meep.GpuRunResult, name)
    assert {"driver", "steps", "meep_time", "wall_time_s"} <= set(
        from_meep.GpuRunResult.__dataclass_fields__)


def test_stop_condition_is_exactly_one():
    with pytest.raises(ValueError):
        from_meep.run_driver(object(), until=None, steps=None)
    with pytest.raises(ValueError):
        from_meep.run_driver(object(), until=1.0, steps=3)


def test_compatibility_verdict_reads_as_bool():
    ok = from_meep.GpuCompatibility(True, ())
    bad = from_meep.GpuCompatibility(False, ("symmetry",))
    assert ok and not bad
    assert "symmetry" in str(bad)
    err = from_meep.MeepSimulationNotLiftable(["a", "b"])
    assert isinstance(err, ValueError) and err.reasons == ("a", "b")
