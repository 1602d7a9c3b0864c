"""An uneven batch of independent 2.5-D lines on one GPU.

A set of survey lines of different sizes (e.g. 16 lines of 3k-40k cells)
cannot share one batched plan: their patterns differ.  `LineBatch` keeps one
`TorchDC2D` (one plan) per line and gets throughput from concurrency and
launch economy:

* streams: every line's eval+grad is issued on one of `streams`
  default-priority streams, so the small lines' level launches overlap and
  fill the chip together instead of running one line after another;
* graphs: `capture()` would record each line's whole eval+grad as a HIP
  graph through `torch.cuda.CUDAGraph` and replay it; it is switched off and
  raises: replays of FP64 graphs matched the eager results bit for bit, but
  a replay of mixed-precision graphs ended the process with a host
  segmentation fault inside hipGraphLaunch whose cause is not found;
* autotuning: `autotune_leaf` times the nested-dissection leaf size per size
  class (cells rounded to a power of two) on one line of that class and
  `LineBatch(leaf="auto")` builds every line with its class's choice.
  Tuning is not part of a timed step.

The nearest reference analogue is cuDSS's batched sequence (dctorch/
solver.py:181-185), which needs one pattern; there is no uneven-batch code in
the reference.
"""
import time

import numpy as np
import torch

from .simulation2d import TorchDC2D


def size_class(n_cells):
    """Power-of-two bucket of a line's cell count."""
    return int(np.round(np.log2(max(n_cells, 1))))


def _eval_grad(eng, x, dobs, w):
    x.grad = None
    phi = eng.misfit(x, dobs, w)
    phi.backward()
    return phi, x.grad


def autotune_leaf(mesh, survey, m, dobs, w, candidates=(32, 48, 64), reps=3, nky=12,
                  precision="mixed", refine=0, device="cuda"):
    """Leaf size with the fastest eval+grad on this line (median of `reps`
    after one warm-up).  Returns (best leaf, {leaf: ms})."""
    dev = torch.device(device)
    times = {}
    for leaf in candidates:
        eng = TorchDC2D(mesh, survey, nky=nky, device=dev, precision=precision,
                        refine=refine, leaf=leaf)
        x = torch.tensor(m, dtype=torch.float64, device=dev, requires_grad=True)
        d_t = torch.as_tensor(dobs, dtype=torch.float64, device=dev)
        w_t = torch.as_tensor(w, dtype=torch.float64, device=dev)
        _eval_grad(eng, x, d_t, w_t)
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize(dev)
            t0 = time.perf_counter()
            _eval_grad(eng, x, d_t, w_t)
            torch.cuda.synchronize(dev)
            ts.append(1e3 * (time.perf_counter() - t0))
        times[leaf] = sorted(ts)[len(ts) // 2]
        del eng
    return min(times, key=times.get), times


class LineBatch:
    """eval+grad of many independent lines per call.

    lines: list of (mesh, survey); leaf: None (the solver default), an int, a
    {size_class: leaf} dict, or "auto" together with `tune_data` =
    [(m, dobs, w) per line] to run `autotune_leaf` on one line per class."""

    def __init__(self, lines, device="cuda", nky=12, precision="mixed", refine=0, streams=4,
                 leaf=None, tune_data=None):
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.tuning = {}
        if leaf == "auto":
            if tune_data is None:
                raise ValueError("leaf='auto' needs tune_data")
            leaf = {}
            for i, (mesh, survey) in enumerate(lines):
                c = size_class(mesh.nC)
                if c not in leaf:
                    m, d, w = tune_data[i]
                    leaf[c], self.tuning[c] = autotune_leaf(mesh, survey, m, d, w, nky=nky,
                                                            precision=precision, refine=refine,
                                                            device=self.device)
        self.leaf = leaf
        self.engines = []
        for mesh, survey in lines:
            lf = leaf.get(size_class(mesh.nC)) if isinstance(leaf, dict) else leaf
            self.engines.append(TorchDC2D(mesh, survey, nky=nky, device=self.device,
                                          precision=precision, refine=refine, leaf=lf))
        self.streams = [torch.cuda.Stream(self.device) for _ in range(max(1, streams))]
        self._data = None
        self._graphs = None

    def set_data(self, dobs, w):
        """Observed data and weights per line (kept on the device)."""
        self._data = [(torch.as_tensor(d, dtype=torch.float64, device=self.device),
                       torch.as_tensor(x, dtype=torch.float64, device=self.device))
                      for d, x in zip(dobs, w)]
        self._graphs = None

    def _stream(self, i):
        return self.streams[i % len(self.streams)]

    def capture(self, ms, warmup=2):
        """Switched off: raises NotImplementedError.  A replay of the
        mixed-precision graphs that `_capture_graphs` records ended the process
        with a host segmentation fault inside hipGraphLaunch (FP64 replays
        matched the eager results bit for bit); the cause is not found."""
        raise NotImplementedError(
            "graph capture of the eval+grad is switched off: a replay of mixed-precision "
            "graphs ended in a host segmentation fault inside hipGraphLaunch (cause not found)")

    def _capture_graphs(self, ms, warmup=2):
        """Record each line's eval+grad at models `ms` as a graph (after
        `warmup` eager runs on its stream, which also sizes every buffer the
        library keeps, so nothing is allocated during capture).  Not called:
        see `capture`."""
        if self._data is None:
            raise RuntimeError("set_data first")
        graphs = []
        for i, eng in enumerate(self.engines):
            s = self._stream(i)
            s.wait_stream(torch.cuda.current_stream(self.device))
            x = torch.tensor(np.asarray(ms[i], float), dtype=torch.float64, device=self.device,
                             requires_grad=True)
            dobs, w = self._data[i]
            with torch.cuda.stream(s):
                for _ in range(warmup):
                    _eval_grad(eng, x, dobs, w)
            s.synchronize()
            g = torch.cuda.CUDAGraph()
            x.grad = None
            with torch.cuda.graph(g, stream=s):
                phi = eng.misfit(x, dobs, w)
                phi.backward()
            graphs.append((g, x, phi, x.grad))
        torch.cuda.synchronize(self.device)
        self._graphs = graphs

    def eval_grad(self, ms):
        """[(phi, grad)] for the models `ms` (one per line), every line on its
        stream; graph replays when captured.  The results are ordered after
        the caller's stream: phi and grad of a captured line are its graph's
        output buffers, overwritten by the next call."""
        if self._data is None:
            raise RuntimeError("set_data first")
        cur = torch.cuda.current_stream(self.device)
        for s in self.streams:
            s.wait_stream(cur)
        out = []
        for i, eng in enumerate(self.engines):
            s = self._stream(i)
            with torch.cuda.stream(s):
                if self._graphs is not None:
                    g, x, phi, grad = self._graphs[i]
                    with torch.no_grad():
                        x.copy_(torch.as_tensor(ms[i], dtype=torch.float64, device=self.device))
                    g.replay()
                    out.append((phi, grad))
                else:
                    x = torch.as_tensor(ms[i], dtype=torch.float64, device=self.device)
                    x = x.detach().clone().requires_grad_(True)
                    dobs, w = self._data[i]
                    phi, grad = _eval_grad(self.engines[i], x, dobs, w)
                    phi, grad = phi.detach(), grad
                    phi.record_stream(cur)
                    grad.record_stream(cur)
                    out.append((phi, grad))
        for s in self.streams:
            cur.wait_stream(s)
        return out
