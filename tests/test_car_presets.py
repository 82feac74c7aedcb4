"""The stock car bodies (rg_car_preset_config) against the reference build's OctaneConfig() ... MercConfig(): the
records must be the same bytes (the reference CPU build is the checker; it is skipped where it is absent)."""
import numpy as np
import pytest

from conftest import oracle_or_skip


def test_presets_match_reference():
    R = oracle_or_skip()
    import rocketgpu_amd as rga

    ref = R.RefSim(1, 6)
    ref.set_env_car_count(np.array([6], np.int32))
    ref.set_car_presets(np.arange(6, dtype=np.uint8))
    want = ref.download_car_configs(6)
    ref.close()
    pad = np.ones(120, bool)
    pad[81] = False  # the padding byte between three_wheels and body
    for i in range(6):
        got = rga.car_preset_config(i)
        assert np.array_equal(got[pad], want[i][pad]), rga.CAR_PRESETS[i]


def test_preset_names_and_errors():
    import rocketgpu_amd as rga

    assert np.array_equal(rga.car_preset_config("Octane"), rga.car_preset_config(0))
    three = rga.car_preset_config("psyclops").view(np.uint8)
    assert three.size == 120
    with pytest.raises(rga.RocketGpuError):
        rga.car_preset_config(8)
