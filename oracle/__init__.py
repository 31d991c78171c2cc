"""CPU oracle for the Rydberg H·v path: test infrastructure, never product code."""
