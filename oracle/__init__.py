"""CPU oracle for the 2DTM search hot path -- TEST INFRASTRUCTURE ONLY.

This package is a plain torch (CPU) restatement of the Leopard-EM reference path
(``src/leopard_em/backend/*`` and the third-party numerics it calls). It exists to
check the HIP implementation in ``leopard-em_amd``; it is never the thing measured
or shipped.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it.

Every function takes a ``dtype`` (float32 for parity with the reference's FP32
path, float64 for error bars) and cites the reference file:line it follows.

Modules
-------
geometry   rotations, Fourier central-slice extraction, search grids
filters    CTF, B-envelope, whitening, bandpass, image / volume pre-processing
search     projection normalisation, cross-correlation, statistics, match/refine loops
peaks      z-score cutoff and greedy peak picking
synthetic  seeded Gaussian-blob volumes and micrographs (analytic projections)

Parity pins (see tests/test_oracle_*.py):
* rotation: closed form (orthonormal, det 1, theta=0 reduces to Rz(phi+psi)).
* slice extraction: identity rotation reproduces the centre plane exactly; 90-degree
  rotations permute planes; Hermitian fold; brute-force float64 trilinear;
  Fourier-slice theorem against analytic Gaussian projections.
* correlation chain: brute-force direct circular correlation in float64.
* normalisation: closed form including the double relative-size scaling.
* statistics: brute-force numpy max/argmax/sum incl. the tie rule.
* grids: reference test counts (24 orientations at 90/90 deg; defocus vectors).
* CTF: closed-form values (k=0, zero rings) -- the exact upstream ``torch_ctf``
  sign/phase conventions are parity unpinned (package not importable here).
* whitening: white noise -> flat filter; upstream binning rule parity unpinned.
"""
