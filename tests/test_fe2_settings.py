"""RVEExecutionSettings validation (src/fem/rve/data.py:116-138) on the CPU."""

import pytest

from concrete_impact_ai_amd.fe2 import RVEExecutionSettings


def test_execution_settings_validation():
    RVEExecutionSettings()
    RVEExecutionSettings("streams", 4)
    RVEExecutionSettings("batched", 1)
    with pytest.raises(ValueError, match="workers must be positive"):
        RVEExecutionSettings("streams", 0)
    with pytest.raises(ValueError, match="threads_per_worker"):
        RVEExecutionSettings("serial", 1, 0)
    with pytest.raises(ValueError, match="exactly one worker"):
        RVEExecutionSettings("serial", 2)
    with pytest.raises(ValueError, match="half of available CPUs"):
        RVEExecutionSettings("process_pool", 10_000)
    with pytest.raises(ValueError, match="backend"):
        RVEExecutionSettings("threads", 1)
