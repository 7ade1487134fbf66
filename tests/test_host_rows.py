"""Host-side logic of the rows either side of the step (no GPU): the
reference's config constants the classes carry, and the loud failure of
every device class when no HIP device is visible (there is no CPU
fallback)."""
from __future__ import annotations

import math

import pytest
import torch

from pourover_amd import forcing as FC
from pourover_amd import multiphase as MP
from pourover_amd import thermal as TH


def test_reference_constants():
    # config/physics.py: flow 4 ml/s through a 5 mm nozzle, +5 %, lattice units, capped at 0.05
    base = 4e-6 / (math.pi * 0.0025 ** 2) * 1.05 * (0.14 / 224 / 0.01 * 1.2) / (0.14 / 224)
    assert base > 0.05 and FC.inlet_velocity() == 0.05
    assert abs(MP.SURFACE_TENSION_LU - 2.4 ** 2 * 0.000625) < 1e-15
    assert abs(MP.RHO_AIR - 1.204 / 965.3) < 1e-18
    assert FC.GRID_SIZE_CM == pytest.approx(0.0625)
    assert 0.5 / FC.GRID_SIZE_CM == pytest.approx(8.0)        # pour diameter in cells
    cfg = TH.CouplingConfig()
    assert (cfg.coupling_frequency, cfg.thermal_subcycles, cfg.enable_diagnostics) == (1, 1, True)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the behaviour without a device")
def test_device_classes_fail_loudly_without_a_device():
    with pytest.raises(RuntimeError, match="HIP device"):
        TH.ThermalLBM(grid=(8, 8, 8))
    with pytest.raises(RuntimeError, match="HIP device"):
        MP.MultiphaseFlow3D(object())
