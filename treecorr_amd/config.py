"""Angle units and small parsing helpers (treecorr/config.py:88-105 parse_unit).

TreeCorr takes unit values from LSSTDESC.Coord's AngleUnit (not installed
here); the values below are that package's published definitions: radians = 1,
degrees = pi/180, hours = pi/12, arcmin = pi/10800, arcsec = pi/648000.
"""
import math

ANGLE_UNITS = {
    'rad': 1.0,
    'deg': math.pi / 180.,
    'hour': math.pi / 12.,
    'arcmin': math.pi / 180. / 60.,
    'arcsec': math.pi / 180. / 3600.,
}


def parse_unit(value):
    """Return the value in radians of an angle-unit name ('deg', 'degrees', 'arcmin', ...)."""
    if value is None:
        return 1.0
    if isinstance(value, (int, float)):
        return float(value)
    v = value.strip().lower()
    if v.startswith('radian'):
        return 1.0
    for name, val in ANGLE_UNITS.items():
        if v.startswith(name):
            return val
    raise ValueError("Unable to parse %s as an angle unit" % value)
