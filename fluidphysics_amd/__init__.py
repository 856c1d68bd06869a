"""MI355X-native unified-particle solver (placeholder; host class added below)."""
