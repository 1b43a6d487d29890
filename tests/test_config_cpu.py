"""RunConfig validation before any device work (CPU only): the reference's
range checks on the dynamics coefficients (gpuwm/config.py:1220-1262,
tests/test_dynamics_key_validation.py) for the fields the step reads."""
import pytest

from arwen_amd.config import RunConfig, validate_for_step

BASE = RunConfig(nx=8, ny=8, nz=10, dx=1000.0, dy=1000.0, ztop=10000.0,
                 dt=6.0)


def test_defaults_pass():
    validate_for_step(BASE)
    validate_for_step(BASE.with_(diff_6th_opt=2, diff_6th_factor=0.0))


@pytest.mark.parametrize("kw,match", [
    ({"diff_6th_factor": -0.12, "diff_6th_opt": 2}, "diff_6th_factor"),
    ({"dampcoef": -0.2, "damp_opt": 3}, "dampcoef"),
    ({"epssm": 0.0}, "epssm"),
    ({"epssm": 1.5}, "epssm"),
    ({"smdiv": -0.1}, "smdiv"),
    ({"damp_opt": 2}, "damp_opt"),
    ({"time_step_sound": 5}, "time_step_sound"),
    ({"khdif": -1.0}, "khdif"),
    ({"diff_6th_opt": 3}, "diff_6th_opt"),
])
def test_out_of_range_values_raise(kw, match):
    with pytest.raises(ValueError, match=match):
        validate_for_step(BASE.with_(**kw))
