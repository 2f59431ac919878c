"""Atomic-unit conversion factors and physical constants.

MicMec works in atomic units throughout and takes its constants from molmod 1.4.x
(micmec/utils.py:25, micmec/pes/mmff.py:33, micmec/sampling/verlet.py:26,
micmec/sampling/npt.py:24), which uses the CODATA-1998 Bohr radius and Hartree energy.

Two factors are fixed by numbers stored in the reference's own fixtures:

* ``angstrom`` reproduces the lattice spacing 1.889726133921252e+01 (10 Angstrom) in
  data/2x2x2_test_micmec.chk to the last printed digit;
* ``pascal`` reproduces the C44 = 10 GPa entry 3.398931582757843e-04 of the elasticity
  tensor in the same file.

``boltzmann``, ``femtosecond`` and ``amu`` appear in no reference file; their values follow
molmod 1.4.x (boltzmann = 1.3806505e-23 J/K in Hartree, second = 1/2.4188843265e-17).
"""

__all__ = [
    "au", "meter", "angstrom", "nanometer", "kilogram", "gram", "amu", "joule",
    "kjmol", "kcalmol", "electronvolt", "newton", "second", "femtosecond",
    "picosecond", "pascal", "bar", "atm", "kelvin", "boltzmann", "avogadro",
    "gigapascal", "megapascal",
]

au = 1.0

# length
meter = 1.0 / 0.5291772083e-10
angstrom = 1.0e-10 * meter
nanometer = 1.0e-9 * meter

# mass
kilogram = 1.0 / 9.10938188e-31
gram = 1.0e-3 * kilogram
avogadro = 6.0221415e23
amu = 1.66053886e-27 * kilogram

# energy
joule = 1.0 / 4.35974381e-18
kjmol = 1.0e3 * joule / avogadro
kcalmol = 4.184e3 * joule / avogadro
electronvolt = 1.602176462e-19 * joule

# force
newton = joule / meter

# time
second = 1.0 / 2.418884326500e-17
femtosecond = 1.0e-15 * second
picosecond = 1.0e-12 * second

# pressure
pascal = newton / meter**2
bar = 1.0e5 * pascal
atm = 1.01325 * bar
megapascal = 1.0e6 * pascal
gigapascal = 1.0e9 * pascal

# temperature
kelvin = 1.0

# constants
boltzmann = 3.1668154051341965e-06
