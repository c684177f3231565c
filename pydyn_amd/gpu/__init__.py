"""The device side of pydyn_amd under the reference's pydyn.gpu names."""
