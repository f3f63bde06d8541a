"""CPU restatement of the PDAC hot path: the checker of the HIP library.

TEST INFRASTRUCTURE ONLY -- imported by tests/, the smoke check and the
CPU-baseline leg of bench.py; never by the product package."""
