"""CPU oracle of the DEMON hot path -- test infrastructure only (see bind.py)."""
