"""Poisson negative log-likelihood (TEST INFRASTRUCTURE ONLY).

The reference uses nifty.re's ``Poissonian(data).amend(R)``
(jubik/instruments/chandra/chandra_likelihood.py:37,
jubik/instruments/erosita/erosita_likelihood.py:89); the formula it
evaluates is documented at demos/x-ray-imaging.py:312-313:
    NLL(lam) = sum(lam - d * log(lam)).
The gradient is 1 - d / lam and the Fisher metric is diag(1 / lam).
"""
import numpy as np


def poisson_nll(lam, d):
    lam = np.asarray(lam, dtype=np.float64)
    d = np.asarray(d, dtype=np.float64)
    return float(np.sum(lam) - np.sum(d * np.log(lam)))


def poisson_grad(lam, d):
    return 1.0 - np.asarray(d, dtype=np.float64) / np.asarray(lam, np.float64)


def poisson_metric(lam, v):
    return np.asarray(v, dtype=np.float64) / np.asarray(lam, np.float64)
