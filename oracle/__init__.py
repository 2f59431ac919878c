"""CPU oracle for the MicMec micromechanical force field and integrators.

TEST INFRASTRUCTURE ONLY. This package restates the reference algorithm on the CPU in numpy
(and C, under ``oracle/c``) so that the HIP implementation in ``micmec_amd`` can be checked
against it. Only the test suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU baseline. The product
path never imports, links or calls anything here.

Modules
-------
constants   molmod 1.4.x constants used by the reference (boltzmann, femtosecond, ...).
topology    brute-force loop restatement of ``Grid``/``build_system``/``get_mic``
            (micmec/utils.py:55-263, micmec/pes/mmff.py:258-286), for tiny grids.
nanocell    per-cell energy and gradient of both nanocell models with the reference's
            derivative tables (micmec/pes/nanocell.py:40-132,
            micmec/pes/nanocell_original.py:38-132, micmec/pes/nanocell_utils.py:32-110).
mmff        ``deformation`` + gpos/vtens/epot assembly (micmec/pes/mmff.py:288-403), as a
            literal per-cell loop and as a chunked vectorised form of the same arithmetic.
cforce      ctypes front end of ``oracle/c/mm_oracle.c``: the same force-field arithmetic in C
            with OpenMP over cells (built into ``oracle/libmm_oracle.so``), for grids where the
            numpy forms are too slow (32^3 and up); pinned against the numpy oracle.
integrators velocity Verlet with NHC and MTK/TBC hooks (micmec/sampling/verlet.py:71-190,
            nvt.py:361-532, npt.py:52-177, 513-757) and the init helpers of
            micmec/sampling/utils.py.

Pinning (how the oracle itself is checked, see tests/test_oracle_*.py): every function is
checked against the reference's CHK fixtures where they hold data for it, against closed forms
(zero energy at rest, affine strain energies, uniaxial cubic stretch, equal-state LSE),
against central finite differences of its own energy, and against invariants (translation,
rotation, virial symmetry, virial = dE/d(deformation)). Parity of the CG optimiser with
molmod's minimiser is unpinned (molmod is absent); only algorithm-independent outcomes are.
"""
