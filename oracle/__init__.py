"""CPU restatement of the HPM-MVS PatchMatch path: the checker used by the
tests, the smoke check and the CPU-baseline leg of bench.py. Not part of the
product."""
