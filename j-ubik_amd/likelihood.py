"""Poisson likelihood on the GPU: the ``nifty.re.Poissonian(data)`` that
J-UBIK composes with its response (chandra_likelihood.py:37,
erosita_likelihood.py:89).  Energy sum(lam - d log lam)
(demos/x-ray-imaging.py:312-313), reduced in FP64; gradient 1 - d/lam;
metric diag(1/lam) and its left square root diag(lam^-1/2), composed with the
linearisation of an amended forward.
"""
import torch

from . import _device as dev
from . import _lib
from .data import Vector


def _leaves(v):
    if isinstance(v, Vector):
        return [v.tree[k] for k in v.tree]
    if isinstance(v, dict):
        return list(v.values())
    return [v]


class Poissonian:
    """Poisson negative log-likelihood of integer counts ``data``."""

    def __init__(self, data):
        self.data = data
        self._forward = None

    def amend(self, forward):
        """Compose with a response: ``lh.amend(R)(x) = lh(R(x))``."""
        new = Poissonian(self.data)
        inner = self._forward
        new._forward = forward if inner is None else (lambda x: forward(inner(x)))
        return new

    def _lam(self, x):
        return self._forward(x) if self._forward is not None else x

    def energy_and_gradient(self, lam):
        """(NLL, dNLL/dlam) for lam shaped like the data (Vector or array)."""
        dev.require_device()
        leaves_l = _leaves(lam)
        leaves_d = _leaves(self.data)
        if len(leaves_l) != len(leaves_d):
            raise ValueError("lambda and data have different structure")
        total = 0.0
        grads = []
        for l, d in zip(leaves_l, leaves_d):
            dtype = dev.float_dtype(l)
            lt = dev.to_device(l, dtype).reshape(-1)
            dt = dev.to_device(d, dtype).reshape(-1)
            if lt.numel() != dt.numel():
                raise ValueError("lambda and data sizes differ")
            g = torch.empty_like(lt)
            nll = torch.empty(1, dtype=torch.float64, device="cuda")
            work = torch.empty(_lib.POISSON_WORK_BYTES // 8, dtype=torch.float64, device="cuda")
            _lib.call("jubik_poisson_nll_grad", dev.code(dtype), dev.ptr(lt), dev.ptr(dt),
                      lt.numel(), dev.ptr(nll), dev.ptr(g), dev.ptr(work), dev.stream())
            total += float(nll.item())
            grads.append(dev.like_input(g.reshape(tuple(l.shape)), l))
        if isinstance(lam, Vector):
            return total, Vector(dict(zip(lam.tree.keys(), grads)))
        if isinstance(lam, dict):
            return total, dict(zip(lam.keys(), grads))
        return total, grads[0]

    def __call__(self, x):
        return self.energy_and_gradient(self._lam(x))[0]

    # ---------------------------------------------------------------- metric
    def _data_metric(self, lam, t, sqrt_form):
        dev.require_device()
        leaves_l, leaves_t = _leaves(lam), _leaves(t)
        if len(leaves_l) != len(leaves_t):
            raise ValueError("lambda and tangent have different structure")
        out = []
        for l, v in zip(leaves_l, leaves_t):
            dtype = dev.float_dtype(l, v)
            lt = dev.to_device(l, dtype).reshape(-1)
            vt = dev.to_device(v, dtype).reshape(-1)
            if lt.numel() != vt.numel():
                raise ValueError("lambda and tangent sizes differ")
            o = torch.empty_like(vt)
            _lib.call("jubik_poisson_metric", dev.code(dtype), dev.ptr(lt), dev.ptr(vt),
                      lt.numel(), int(sqrt_form), dev.ptr(o), dev.stream())
            out.append(dev.like_input(o.reshape(tuple(v.shape)), v))
        if isinstance(t, Vector):
            return Vector(dict(zip(t.tree.keys(), out)))
        if isinstance(t, dict):
            return dict(zip(t.keys(), out))
        return out[0]

    def _linearised(self):
        f = self._forward
        if f is None:
            return None
        if not (hasattr(f, "jvp") and hasattr(f, "vjp")):
            raise NotImplementedError(
                "the metric of an amended likelihood needs a forward with jvp(x, t) and "
                "vjp(x, ct) (LinpatchOperator has them; XrayModel.metric is the fused product)")
        return f

    def metric(self, primals, tangents):
        """Fisher metric applied to ``tangents`` at ``primals``: diag(1/lam)
        in data space, J^T diag(1/lam) J through an amended forward, as
        nifty.re's Likelihood.metric."""
        f = self._linearised()
        if f is None:
            return self._data_metric(primals, tangents, False)
        lam = f(primals)
        out = f.vjp(primals, self._data_metric(lam, f.jvp(primals, tangents), False))
        return dev.like_input(out, tangents)

    def left_sqrt_metric(self, primals, tangents):
        """Left square root of the metric applied to data-space ``tangents``:
        diag(lam^-1/2) t, or J^T diag(lam^-1/2) t through an amended forward
        (nifty.re's Likelihood.left_sqrt_metric, which the sample draws of
        optimize_kl call)."""
        f = self._linearised()
        if f is None:
            return self._data_metric(primals, tangents, True)
        out = f.vjp(primals, self._data_metric(f(primals), tangents, True))
        return dev.like_input(out, tangents)
