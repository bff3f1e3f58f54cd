"""CPU oracle for the time-evolution hot path. TEST INFRASTRUCTURE ONLY."""
