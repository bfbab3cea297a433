"""CPU oracle for the ERI + J/K Fock-build path.

TEST INFRASTRUCTURE ONLY.  This package restates, on the CPU, the reference
Gaussium algorithms for two-electron repulsion integrals (Obara-Saika), the
Boys function, basis normalisation and expansion, one-electron integrals, the
restricted J/K Fock build and the RHF/DIIS self-consistent field.  Each
function cites the reference file:line it follows (paths relative to
``src/gaussium/`` of the reference).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product package ``gaussium_amd`` never imports it.

Pinning (what checks each part, besides restating the formula):
  boys        -- 15 reference golden values, F_m(0)=1/(2m+1), the erf closed
                 form of F_0, mpmath incomplete-gamma values (tests/test_oracle_*).
  basis       -- 6 reference primitive-normalisation values, <phi|phi>=1 of
                 contracted functions, reference component order.
  os_eri      -- 6 HeH+ reference ERIs, the mpmath closed form of (ss|ss),
                 derivative identities for p and d (mpmath differentiation of
                 the closed form), 8-fold symmetry, Schwarz inequality, the
                 independent C restatement (oracle/c/os_eri.c) on every class.
  one_electron-- 11 reference S/T/V values, S_ii = 1, virial-free checks.
  jk / scf    -- reference RHF energies (HeH+/STO-3G, HeH+/6-311+G**,
                 C2H4/3-21G, mock HeH+ -4.227529), reference total-energy
                 fixture, density fixture, E_nuc fixtures.
  uhf         -- the reference's UHF energy of O2/STO-3G (-147.634028141).
  guhf        -- the reference's GUHF energy of O2/STO-3G (-147.634028141);
                 the blocked Fock sums against spatial J/K block by block
                 (off-diagonal spin blocks included); the blocked
                 diagonalisation against two spatial ones.
  mp2         -- the reference's MP2 energy of CO/STO-3G (-111.354512528);
                 the MO transformation against an einsum contraction.
"""
