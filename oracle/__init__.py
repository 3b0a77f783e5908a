"""CPU oracle for the NECG energy path (test infrastructure only).

Loaded by tests, by the smoke check and by the CPU-baseline leg of the
benchmark; never by the product package."""
