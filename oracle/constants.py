"""molmod 1.4.x constants used by the reference (test infrastructure only).

angstrom and pascal are pinned by data/2x2x2_test_micmec.chk; boltzmann, femtosecond and amu
follow molmod 1.4.x and are not pinned by any reference file.
"""

bohr_m = 0.5291772083e-10
angstrom = 1.0e-10 / bohr_m
hartree_j = 4.35974381e-18
joule = 1.0 / hartree_j
meter = 1.0 / bohr_m
pascal = joule / meter**3
bar = 1.0e5 * pascal
second = 1.0 / 2.418884326500e-17
femtosecond = 1.0e-15 * second
kilogram = 1.0 / 9.10938188e-31
amu = 1.66053886e-27 * kilogram
kelvin = 1.0
boltzmann = 3.1668154051341965e-06
