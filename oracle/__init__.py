"""The oracle: a CPU restatement of particle-stabilizer's CPU step (psx_oracle.cpp), used only as
the checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. Never shipped,
never on the product path."""
