"""MI355X-native PatchMatch depth/normal estimation with the HPM-MVS
per-problem interface. The directory name is not a Python identifier, so the
entry points load it by path under the module name ``hpm_mvs_amd``; the
bindings live in ``hpm_mvs_amd.hpm``."""
