"""CPU oracle of the optical-photon transport (test infrastructure only).

Importable by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg;
never by the product package.
"""
