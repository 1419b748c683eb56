"""CPU oracle of the YallaSQL operator hot path — test infrastructure only."""
