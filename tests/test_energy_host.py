"""Energy meter host logic (lap_amd.energy) against a simulated AMD SMI whose
energy accumulator advances in steps, as the GPU firmware's does: a constant
power P with steps every `step` seconds must read as P from the steps, while
the plain two-read difference over a short region may be off by a step."""
import time

import pytest

from lap_amd import energy as E

RES_UJ = 15.3


class StepSmi:
    def __init__(self, watts, step, phase=0.0):
        self.watts, self.step, self.t0 = watts, step, time.perf_counter() - phase

    def amdsmi_get_energy_count(self, h):
        k = int((time.perf_counter() - self.t0) / self.step)
        joules = k * self.step * self.watts
        return {"energy_accumulator": int(round(joules / (RES_UJ * 1e-6))),
                "counter_resolution": RES_UJ}

    def amdsmi_get_power_info(self, h):
        return {"current_socket_power": self.watts}


@pytest.mark.parametrize("step,phase", [(0.1, 0.0), (0.1, 0.07), (0.25, 0.2)])
def test_meter_reads_power_between_accumulator_steps(monkeypatch, step, phase):
    smi = StepSmi(800.0, step, phase)
    monkeypatch.setattr(E, "_load_smi", lambda: (smi, object(), "0000:00:00.0", None))
    with E.EnergyMeter(period=0.005) as em:
        time.sleep(1.2)
    r = em.report(solves=2)
    assert r["counter_method"] == "between_steps"
    assert r["counter_steps"] >= 3
    assert abs(r["mean_power_w_counter"] - 800.0) <= 0.05 * 800.0, r
    assert abs(r["energy_j_per_solve"] - 800.0 * r["seconds"] / 2) <= 0.05 * 400.0 * r["seconds"]
    assert r["sources_agree"] and r["power_field"] == "current_socket_power"
    assert abs(r["counter_step_interval_s"] - step) <= 0.3 * step


def test_meter_falls_back_to_two_reads_without_steps(monkeypatch):
    smi = StepSmi(500.0, 10.0)  # no step inside a 0.2 s region
    monkeypatch.setattr(E, "_load_smi", lambda: (smi, object(), "0000:00:00.0", None))
    with E.EnergyMeter(period=0.005) as em:
        time.sleep(0.2)
    r = em.report()
    assert r["counter_method"] == "two_reads"
    assert r["energy_j_per_solve"] == r["energy_j_per_solve_two_reads"]


def test_meter_unavailable_reports_error(monkeypatch):
    monkeypatch.setattr(E, "_load_smi", lambda: (None, None, None, "no GPU handle"))
    with E.EnergyMeter() as em:
        pass
    assert not em.available
    assert "no GPU handle" in em.report()["error"]
