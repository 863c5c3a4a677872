"""Host-side checks of the benchmark workloads (CPU only)."""
import math

import numpy as np

from spade_amd import studies


def test_ensemble_distribution():
    specs = studies.ensemble_specs(300, 500, seed=2024)
    assert len(specs) == 300
    assert all(c in (32, 48, 64, 96) for s in specs for c in s.cells)
    assert all(0.1 <= s.mach <= 0.5 and 100.0 <= s.reynolds <= 5000.0 for s in specs)
    kinds = [s.kind for s in specs]
    assert 100 < kinds.count("tgv") < 200
    # uneven block counts per case, about 52.7 blocks on average ((2+3+4+6)/4)^3
    blocks = [s.blocks for s in specs]
    assert len(set(blocks)) > 10
    assert 40 < np.mean(blocks) < 65
    # seeded: the same list every time
    again = studies.ensemble_specs(300, 500, seed=2024)
    assert [(s.kind, s.cells, s.mach) for s in specs] == [(s.kind, s.cells, s.mach) for s in again]


def test_case_constants_match_definitions():
    from spade_amd import CaseSpec
    k = studies.case_constants(CaseSpec("tgv", (32,) * 3, mach=0.1, reynolds=1600.0))
    assert math.isclose(k["R"] * 1.4, 100.0)          # c0^2 = gamma R T0 = (1/M)^2
    assert math.isclose(k["mu"], 1 / 1600.0)
    v = studies.case_constants(CaseSpec("vortex", (32,) * 3, mach=0.3, reynolds=500.0))
    assert math.isclose(v["R"], 1 / 1.4) and math.isclose(v["mu"], 0.3 / 500.0)


def test_channel_initial_state():
    c = studies.channel()
    h = c["host"]
    assert h.shape == (576, 20, 20, 20, 5)
    inner = h[:, 2:-2, 2:-2, 2:-2]
    assert np.all(inner[..., 1] == 1.0)
    assert np.allclose(inner[..., 0], c["R"])
    # Poiseuille profile plus perturbations of at most 5% of the bulk velocity
    assert float(np.max(inner[..., 2])) <= 1.5 + 0.05
    assert float(np.max(np.abs(inner[..., 3:5]))) <= 0.05
    # Mach 0.2 at the wall temperature: c = 5
    assert math.isclose(math.sqrt(c["gamma"] * c["R"] * c["T_wall"]), 5.0)
    assert 1e-4 < c["dt"] < 1e-3
