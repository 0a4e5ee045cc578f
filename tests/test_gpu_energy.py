"""Energy meter (lap_amd.energy): around a second of device work the GPU's
energy accumulator advances, and the mean power it implies and the sampled
socket power (the reference's time x mean-power method) are both physical
for an MI355X (between 50 W and the 1.4 kW board limit).  The accumulator
advances in firmware steps; over two seconds of work the meter must see at
least two of them and take its power between them.  The two sources are
reported side by side, not required to agree (the sampled socket power is the
firmware's own running average and lags changes of load).
Skipped where the AMD SMI library cannot read the counters."""
import time

import pytest
import torch

from lap_amd.energy import EnergyMeter

pytestmark = pytest.mark.gpu


def test_energy_meter_reads_the_accumulator():
    x = torch.randn(1 << 26, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    with EnergyMeter() as em:
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < 2.0:
            for _ in range(20):
                torch.mul(x, 1.0000001, out=y)
            torch.cuda.synchronize()
    if not em.available:
        pytest.skip(em.error)
    r = em.report()
    print(f"[diag] energy meter: {r}", flush=True)
    assert r["energy_j_per_solve"] > 0
    assert r["counter_method"] == "between_steps", r
    assert 50.0 <= r["mean_power_w_counter"] <= 1400.0
    if "mean_power_w_sampled" in r and r["power_samples"] >= 10:
        assert 50.0 <= r["mean_power_w_sampled"] <= 1400.0
