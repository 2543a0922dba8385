"""CPU float64 restatement of J-UBIK's forward model and gradient.

TEST INFRASTRUCTURE ONLY.  This package is the checker: it is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, and by nothing else.  The product (``j-ubik_amd/``) never
imports it and has no CPU fallback.

Modules
-------
convolve   restates jubik/convolve.py (weights, slice_patches,
           linpatch_convolve, convolve, integrate) plus the hand-written
           adjoint of linpatch_convolve.
response   restates jubik/response.py (exposure, readout) and the readout
           transpose.
likelihood the Poisson negative log-likelihood used through
           ``jft.Poissonian`` (formula documented at
           demos/x-ray-imaging.py:312-313), its gradient and Fisher metric.
sky        the correlated-field GP sky with a fixed amplitude field
           (jubik/sky_model/multifrequency/spectral_product_mf_sky.py:171-298,
           jubik/sky_models.py:306-341) and its adjoint.
model      composition lambda = Readout(Exposure(PSF(a * s))) as in
           jubik/instruments/chandra/chandra_response.py:136-142 and
           jubik/instruments/erosita/erosita_response.py:349-350, the NLL,
           its gradient with respect to the latent excitations, and the MGVI
           metric product.
brute      pure-Python loop versions (no FFT) for tiny inputs.

Parity status (how each function is pinned is recorded in the tests):
every function here is pinned by at least one of the reference test
fixtures, a closed form, a brute-force loop version or finite differences.
The exact additive constant of nifty.re's Poissonian and nifty.re's
amplitude normalisation are outside this restatement (the amplitude field
is an input), so no parity claim is made for them.
"""
