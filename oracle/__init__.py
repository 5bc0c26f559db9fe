"""FP64 CPU oracle for the micromagnetic time-stepping path.

TEST INFRASTRUCTURE ONLY.  This package is a plain numpy restatement of the
reference's algorithms (dirkpetersen/mumax-plus) in float64.  It exists to
check the HIP product path; it is never imported by the product.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker.

Modules and the reference files each one restates:

* ``newell``  -- Newell demag tensor (src/physics/newell.cu:7-139) and the
  kernel grid / PBC image sums (src/physics/strayfieldkernel.cu:25-99).
* ``demag``   -- zero-padded FFT convolution (src/physics/strayfieldfft.cu:
  51-100,191-246) and the direct-sum executor (src/physics/strayfieldbrute.cu).
* ``fields``  -- exchange (src/physics/exchange.cu:20-104), DMI
  (src/physics/dmi.cu:87-178, dmi.hpp:53-69), anisotropy
  (src/physics/anisotropy.cu:25-140), Zeeman (src/physics/zeeman.cu:53-91),
  LLG and damping torque (src/physics/torque.cu:21-77), energy densities
  (src/physics/energy.cu:18-50), effective field (effectivefield.cpp:15-28).
* ``rk``      -- Butcher tableaus (src/core/butchertableau.cpp:89-224), the
  Runge-Kutta stepper (src/core/rungekutta.cpp:25-151) and the time solver
  (src/core/timesolver.cpp:37-111, mumaxplus/timesolver.py:61-152).
* ``afm``     -- antiferromagnet sublattice terms: sister-aware exchange
  (src/physics/exchange.cu:106-202), AFM DMI (dmi.cu:112-240), homogeneous and
  inhomogeneous AFM exchange (afmexchange.cu:28-149).
* ``elastic`` -- strain, stress, elastic force, magnetoelastic field and force,
  elastic damping (straintensor.cu, stresstensor.cu, elasticforce.cu,
  magnetoelasticfield.cu, magnetoelasticforce.cu, elasticdamping.cu).
* ``model``   -- a minimal world / ferromagnet / antiferromagnet container
  that assembles the above the way src/physics/mumaxworld.cpp:159-205 and
  elastodynamics.cu:23-96 do.
* ``relax``   -- total energies, the relaxer (src/physics/relaxer.cpp:87-174)
  and the energy minimiser (src/physics/minimizer.cu:8-160).
* ``stray``   -- stray fields between grids by the direct sum
  (src/physics/strayfieldbrute.cu:15-88).

Pinning status of every function is recorded in each module's docstring.
"""

from . import newell, demag, fields, rk, afm, elastic, model, relax, stray  # noqa: F401
