"""CPU oracle for the EMerge frequency-domain FEM hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import, call, link or run anything in
this package, and only as the checker (or the timed CPU baseline) -- never as
the product path.  The product (`emerge_amd`) never imports it.

Every function restates a reference function from its source text (file:line
cited in each docstring; paths are relative to the reference repository,
``E/`` = ``src/emerge/_emerge/``, ``ASM/`` = ``E/physics/microwave/assembly/``).
The reference itself cannot be imported here (numba, gmsh, loguru and emsutil
are absent), so the restatement is pinned independently:

* topology / DOF maps: brute-force set/dict restatement on tiny meshes
  (`oracle.topology`) and the closed-form Kuhn counts;
* element matrices: brute-force quadrature of the basis as defined by the
  reference's interpolation routine (`oracle.quadrature`), symmetry, the
  curl-kernel invariant (gradients of P2 scalars lie in the curl null space),
  and the h / h^3 scale laws;
* COO -> CSR: scipy 1.15.3 itself (the version the reference's lock file pins);
* Robin / port surface terms: 2-D quadrature of the tangential basis;
* direct solve: scipy SuperLU with the reference's options;
* S-parameters and port excitation (tests/test_oracle_sparam.py):
  arg S21 = -beta L of a TE10 section, the first-order ABC reflection
  (beta - k0)/(beta + k0) exp(-2j beta L), the 1/sqrt(P_q) power law, the
  transmission-line S-matrix of a parallel-plate TEM line between two
  lumped ports (and a shorted one), and the discrete energy balance of an
  internal lumped-port sheet.
* far field (`oracle.farfield`): the closed-form far field of a Hertzian
  dipole from its exact near fields on a closed surface, the extinction
  theorem for a plane wave, and the divergence theorem for the weights.

Pinning status per function is listed in each module header.
"""
