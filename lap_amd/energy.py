"""Energy to solution on the GPU (the reference's power-sampling methodology,
script_hip/continous.sh:5 and script/analyzeEnergy_cuda.py:183-201).

    with EnergyMeter() as em:
        ... timed GPU work ...
    em.joules, em.mean_watts, em.seconds

The reference samples `rocm-smi --csv -P` once a second while the solver
runs and reports energy = solve time x mean power.  Here the GPU's own
energy accumulator is read before and after the region through the AMD SMI
library (amdsmi_get_energy_count: microjoule counter x resolution), and the
socket power is sampled every `period` seconds on a thread beside it, so
both the accumulator's energy and the reference's time x mean-power figure
are reported.  When the AMD SMI library cannot be loaded or does not expose
the counters, the meter reports itself unavailable (no fallback figure).

The accumulator does not advance smoothly: the firmware adds to it in steps
(tens to hundreds of milliseconds apart), so a difference of two reads taken
at arbitrary instants over a window of a second or two can be off by a whole
step, in either direction.  The sampling thread therefore also reads the
accumulator, and the meter takes the mean power between the first and the
last step it saw inside the region (energy between two update instants over
the time between them); that power times the region's length is the energy.
The plain two-read difference is kept beside it.
"""
from __future__ import annotations

import os
import sys
import threading
import time

_SMI_DIR = "/opt/rocm/share/amd_smi"


def _our_bus():
    """PCI bus number of the GPU torch uses as device 0, or None."""
    try:
        import torch
        return int(torch.cuda.get_device_properties(0).pci_bus_id)
    except Exception:
        return None


def _load_smi():
    """(amdsmi module, handle of our GPU, bdf string, error)."""
    try:
        if os.path.isdir(_SMI_DIR) and _SMI_DIR not in sys.path:
            sys.path.append(_SMI_DIR)
        import amdsmi  # noqa: PLC0415
        amdsmi.amdsmi_init()
        handles = amdsmi.amdsmi_get_processor_handles()
        if not handles:
            return None, None, None, "no GPU handle"
        bus = _our_bus()
        for h in handles:  # the handle whose PCI bus is torch's device 0
            bdf = amdsmi.amdsmi_get_gpu_device_bdf(h)
            if len(handles) == 1 or (bus is not None and int(bdf.split(":")[1], 16) == bus):
                return amdsmi, h, bdf, None
        return None, None, None, f"device 0 (bus {bus}) not among {len(handles)} GPUs"
    except Exception as exc:  # library absent, no permission, no device
        return None, None, None, f"{type(exc).__name__}: {exc}"


class EnergyMeter:
    def __init__(self, period: float = 0.02):
        self.period = period
        self.smi, self.h, self.bdf, self.error = _load_smi()
        self.samples: list[tuple[float, float]] = []
        self.energy_samples: list[tuple[float, float]] = []
        self.power_key = None
        self.edge_watts = None
        self.updates = 0
        self.joules = None
        self.seconds = None
        self.mean_watts = None
        self._stop = threading.Event()
        self._thread = None

    @property
    def available(self) -> bool:
        return self.smi is not None

    def _energy_j(self):
        c = self.smi.amdsmi_get_energy_count(self.h)
        return c["energy_accumulator"] * c["counter_resolution"] * 1e-6

    def _power_w(self):
        p = self.smi.amdsmi_get_power_info(self.h)
        for key in ("current_socket_power", "socket_power", "average_socket_power"):
            v = p.get(key)
            if isinstance(v, (int, float)) and v not in (0, 0xFFFF):
                self.power_key = key
                return float(v)
        return None

    def _sample(self):
        while not self._stop.is_set():
            try:
                e = self._energy_j()
                self.energy_samples.append((time.perf_counter(), e))
                w = self._power_w()
                if w is not None:
                    self.samples.append((time.perf_counter(), w))
            except Exception:
                pass
            self._stop.wait(self.period)

    def _edge_power(self):
        """Mean power between the first and the last accumulator step seen
        in the region, with the number of steps; (None, steps) when fewer
        than two steps span at least a quarter of the region."""
        es = [(self._t0, self._e0)] + self.energy_samples + [(self._t1, self._e1)]
        steps = [(t, e) for (tp, ep), (t, e) in zip(es, es[1:]) if e != ep]
        if len(steps) < 2 or steps[-1][0] - steps[0][0] < 0.25 * self.seconds:
            return None, len(steps)
        (ta, ea), (tb, eb) = steps[0], steps[-1]
        return (eb - ea) / (tb - ta), len(steps)

    def __enter__(self):
        if not self.available:
            return self
        self._e0 = self._energy_j()
        self._t0 = time.perf_counter()
        self._thread = threading.Thread(target=self._sample, daemon=True)
        self._thread.start()
        return self

    def __exit__(self, *exc):
        if not self.available:
            return False
        self._t1 = time.perf_counter()
        self._e1 = self._energy_j()
        self._stop.set()
        self._thread.join()
        self.seconds = self._t1 - self._t0
        self.edge_watts, self.updates = self._edge_power()
        self.joules = (self.edge_watts * self.seconds if self.edge_watts is not None
                       else self._e1 - self._e0)
        if self.samples:
            self.mean_watts = sum(w for _, w in self.samples) / len(self.samples)
        return False

    def report(self, solves: int = 1) -> dict:
        if not self.available:
            return {"error": f"AMD SMI energy counters unavailable ({self.error})"}
        out = {
            "energy_j_per_solve": self.joules / solves,
            "mean_power_w_counter": self.joules / self.seconds,
            "counter_method": "between_steps" if self.edge_watts is not None else "two_reads",
            "counter_steps": self.updates,
            "energy_j_per_solve_two_reads": (self._e1 - self._e0) / solves,
            "seconds": self.seconds, "solves": solves,
            "power_samples": len(self.samples), "gpu_bdf": self.bdf,
        }
        if self.updates >= 2:
            es = [t for (tp, ep), (t, e) in zip(self.energy_samples, self.energy_samples[1:])
                  if e != ep]
            if len(es) >= 2:
                out["counter_step_interval_s"] = (es[-1] - es[0]) / (len(es) - 1)
        if self.mean_watts is not None:
            out["mean_power_w_sampled"] = self.mean_watts
            out["power_field"] = self.power_key
            out["energy_j_per_solve_time_x_power"] = self.mean_watts * self.seconds / solves
            # the two sources should tell the same story; a counter that implies
            # a mean power far from the sampled one (or above the board's limit)
            # is flagged rather than trusted
            counter_w = out["mean_power_w_counter"]
            out["sources_agree"] = bool(abs(counter_w - self.mean_watts)
                                        <= 0.25 * max(self.mean_watts, 1.0))
        return out
