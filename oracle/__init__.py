"""FP64 CPU oracle for the mhs half-space matrix assembler -- TEST INFRASTRUCTURE.

This package is a plain-numpy restatement of the reference algorithm
(brendanjmeade/mhs, ``src/mhs/``), written for this repository and used only
as the checker: by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``.  The product path (``mhs_amd``) never
imports it; the product runs the HIP kernels or raises.

Every function cites the reference file:line it restates.  The oracle is
pinned three ways (see ``tests/test_oracle_*.py``):

* against golden vectors written by importing the reference in the build
  container (``tests/golden/``), at 1e-12 relmax;
* against mathematics the reference's own gates use -- Gauss quadrature of the
  pointwise Cortez-Kelvin formulas, the free-surface traction, the Love disk,
  finite differences of the Boussinesq kernels;
* against identities (total - eigenstress == elastic, P1/P2 uniform == P0,
  interaction == contracted stress).

Module map (reference layer in brackets):

    geometry   frame, shape functions, Gauss rules   [fullspace/frame, shape, moments]
    edges      edge antiderivatives, solid angle     [fullspace/primitives]
    moments    closed-form / far-field moment tables [fullspace/moments]
    fullspace  lift and nodal U/H/E                  [fullspace/kernels]
    image      Mindlin image R- and Q-families       [kernels/image]
    fold       folded image of surface sources       [kernels/fold]
    assemble   per-source assembly, all entry points [kernels/assemble, matrices]
"""
