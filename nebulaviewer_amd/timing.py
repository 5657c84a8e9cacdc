"""Per-kernel timing with HIP events recorded on the stream the kernels are launched on.

Sim launches every kernel on torch's current stream of its device, so torch.cuda.Event pairs recorded on that
stream bracket exactly the launches between them.
"""
import torch


class KernelTimer:
    def __init__(self, device="cuda:0"):
        self.device = torch.device(device)
        self.pairs = []

    def start(self, name):
        s = torch.cuda.Event(enable_timing=True)
        s.record(torch.cuda.current_stream(self.device))
        return name, s

    def stop(self, token):
        name, s = token
        e = torch.cuda.Event(enable_timing=True)
        e.record(torch.cuda.current_stream(self.device))
        self.pairs.append((name, s, e))

    def summary(self):
        """{name: dict(count, total_ms, avg_ms)}; synchronises the device."""
        torch.cuda.synchronize(self.device)
        out = {}
        for name, s, e in self.pairs:
            d = out.setdefault(name, dict(count=0, total_ms=0.0))
            d["count"] += 1
            d["total_ms"] += s.elapsed_time(e)
        for d in out.values():
            d["avg_ms"] = d["total_ms"] / max(d["count"], 1)
        return out
