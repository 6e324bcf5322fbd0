"""Host side (Python) of the MI355X solver for the visionary-sim substep:
ctypes binding of the C ABI (native) and scene assembly (scenes)."""
