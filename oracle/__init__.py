"""CPU oracle for the scopyon hot path (float64 numpy/scipy restatement).

TEST INFRASTRUCTURE ONLY.  Nothing in this package is part of the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or, for the bench, as the
timed CPU baseline).  The product path in ``scopyon_amd`` never imports it and
fails loudly when its HIP library is missing.

Each function restates one reference function, in the same arithmetic order,
so that run on the same numpy/scipy it reproduces the reference's float64
results.  Citations are ``file:line`` into the scopyon source tree
(``src/scopyon/...``).

Modules
  ``sampling``  -- ``sampling2.py`` (``sample``) and ``sampling.py``
                   (``sample_inputs``).
  ``epifm``     -- ``_epifm.py``: PSF tables, ``overlay_signal_``, emitter
                   weighting, EMCCD/CCD/CMOS detectors, ADC, ``output_frame``
                   and ``generate_frames``; ``base.py`` input formatting.
  ``params``    -- ``EPIFMConfigs`` scalars from a plain config mapping.

How each function is pinned (checks live in ``tests/test_oracle_*.py``):
  * PSF radial profile: closed form at r=0, z=0 (``0.255025 alpha^2/pi``);
    the normalisation invariant ``norm * alpha^2/pi * (1 nm)^2 == 1`` for the
    TRITC emission line; z -> -z symmetry; Gaussian closed form;
    an independent mpmath evaluation of the same Riemann sum.
  * ``radial_to_cartesian``: exact grid values on the axes, corner plateau.
  * ``overlay_signal_``: brute-force per-cell loop on tiny frames; photon
    conservation when the footprint lies inside the frame; whole-pixel
    translation equivariance.
  * EMCCD pmf: the compound Poisson-Gamma closed form evaluated by an
    independent series; atom mass ``exp(-E)``; mean ``g E`` and variance
    ``2 g^2 E``; numpy ``RandomState.choice`` restated as
    cumsum/searchsorted.
  * Full ``form_image`` path (PSF, splat, emitter weights, EMCCD draw order,
    ADC, 8-bit scaling): regenerates the reference's own example output
    ``examples/tirf_000.png`` (``examples/tirf.py``) pixel for pixel,
    committed as ``tests/golden/tirf_000_u8.npz``.
  * Sampling: golden vectors written by the reference's own ``sampling.py``
    and ``sampling2.py`` (``tests/golden/make_sampling_golden.py``), plus
    MSD / KS statistics.
"""
