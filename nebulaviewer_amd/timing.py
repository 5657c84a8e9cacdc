"""Per-kernel timing with HIP events recorded on the stream the kernels are launched on.

Sim launches every kernel on torch's current stream of its device, so torch.cuda.Event pairs recorded on that
stream bracket exactly the launches between them.
"""
import torch


class KernelTimer:
    """reserve: events created up front, so that a timed region records events without creating any (creating
    events between launches has stretched some measured steps by a third)"""

    def __init__(self, device="cuda:0", reserve=0):
        self.device = torch.device(device)
        self.pairs = []
        self._pool = [torch.cuda.Event(enable_timing=True) for _ in range(reserve)]

    def _event(self):
        return self._pool.pop() if self._pool else torch.cuda.Event(enable_timing=True)

    def start(self, name):
        s = self._event()
        s.record(torch.cuda.current_stream(self.device))
        return name, s

    def stop(self, token):
        name, s = token
        e = self._event()
        e.record(torch.cuda.current_stream(self.device))
        self.pairs.append((name, s, e))

    def summary(self):
        """{name: dict(count, total_ms, avg_ms, min_ms, max_ms)}; synchronises the device."""
        torch.cuda.synchronize(self.device)
        out = {}
        for name, s, e in self.pairs:
            d = out.setdefault(name, dict(count=0, total_ms=0.0, min_ms=float("inf"), max_ms=0.0))
            ms = s.elapsed_time(e)
            d["count"] += 1
            d["total_ms"] += ms
            d["min_ms"] = min(d["min_ms"], ms)
            d["max_ms"] = max(d["max_ms"], ms)
        for d in out.values():
            d["avg_ms"] = d["total_ms"] / max(d["count"], 1)
        return out
