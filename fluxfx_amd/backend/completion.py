"""Step-completion fence (backend/completion.py:5-20): waits for the stream
and checks one texel of every state field for finiteness."""
from math import isfinite


class StepCompletion:
    def __init__(self, device):
        self.device = device

    def wait(self, solver):
        fields = [solver.density, solver.temperature, *solver._velocity, solver.projector.after]
        if solver.combustion:
            fields += [solver.combustion.fuel, solver.combustion.flame]
        torch = self.device.torch
        sample = torch.stack([f.tensor[0] for f in fields]).cpu()
        if not all(isfinite(v) for v in sample.tolist()):
            raise RuntimeError("Nonfinite completion sample; reset simulation")

    def close(self):
        self.device = None
