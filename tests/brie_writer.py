"""Test access to the .brie writer of the package."""
from parahpop_amd.brie_io import cbor, circular_geocentric, fit_unit, ssb_units, write_brie  # noqa: F401
