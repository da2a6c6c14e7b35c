"""Test access to the .brie writer of the package."""
from parahpop_amd.brie_io import cbor, circular_geocentric, fit_unit, iau2006_channels, ssb_units, write_brie, write_brot2  # noqa: F401,E501
