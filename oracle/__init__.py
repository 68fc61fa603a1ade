"""CPU oracle (test infrastructure only): C++ restatement of MiniLake's CPU executor."""
