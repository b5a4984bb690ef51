"""MI355X-native engine for propaq's per-gate Pauli / Majorana term propagation."""
