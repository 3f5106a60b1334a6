"""Loader for the C-ABI library (filled in below)."""
def load():
    raise RuntimeError("libmsbg_hip.so not built")
