"""Decaying isotropic turbulence with entropy viscosity (16^3 hex8, p = 6,
Ma_t = 0.3, Re_lambda = 100, c_E = 1, c_max = 0.5), 500 steps with FP32
contractions: the run stays finite, the kinetic energy decays, the entropy
viscosity is active, and the state stays close to the FP64 run."""
import numpy as np
import pytest

from dsem_amd import iso
from dsem_amd.mixed_study import field_errors

pytestmark = pytest.mark.gpu


def test_isotropic_turbulence_ev_mixed_500_steps():
    print("\n[case] iso 16^3 p=6 500 steps fp32 vs fp64", flush=True)
    r32 = iso.run(n=16, p=6, steps=500, precision="fp32")
    r64 = iso.run(n=16, p=6, steps=500, precision="fp64")
    for r in (r32, r64):
        h = r["history"]
        assert np.all(np.isfinite(r["JQ"]))
        assert h[-1, 2] < 0.95 * h[0, 2]                 # energy decays
        assert np.all(np.diff(h[:, 2]) < 0)
        assert r["mu_h_max"] > 0                          # EV active
    e = field_errors(r32["JQ"], r64["JQ"])
    print({k: v["rel_l2"] for k, v in e.items()}, "ms/step", r32["wall_s"] * 2, r64["wall_s"] * 2, flush=True)
    assert all(v["rel_l2"] < 1e-5 for v in e.values())
