"""CPU oracle of kettle's simulate() hot path -- test infrastructure only (see oracle.cpp)."""
